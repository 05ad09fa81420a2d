"""Writes the golden fixtures of tests/golden from the reference's own code.

Runs oracle/_ref/ref_harness (the reference's bounds/, multicore/ivm*,
common/ and heuristic/ sources compiled unmodified by oracle/Makefile, with a
restated sequentialbb driver loop) in this container, where the reference
tree exists.  The fixtures are plain data (inputs and expected outputs):

  bounds_<seed>_<n>x<m>.npz   LB1 children (both directions), LB1', LB2
                              (all pairs, with the idle sum), makespan on
                              seeded random partial schedules
  solve_cases.json            optimum / decomposed / leaves / LB1 bounds /
                              refinement bounds of single-thread solves
  taillard_ptm.json           sha256 of the ta001..ta120 processing times
  findall_cases.json          single-thread solves with findAll (prune on
                              >, the bound frozen at the optimum)
  neh_cases.json              the reference fastNEH (initialSort + runNEH
                              from the identity order, its uninitialised
                              tail cell zeroed by the harness) on the 20x5
                              seed-0 instance, the 512 batched instances,
                              the 256-request serving stream, one random
                              instance per n = 3..63 and instances built
                              to tie: equal job totals and equal insertion
                              makespans (explicit times)

Usage: python tests/golden/make_golden.py [bounds|solve|taillard|neh|findall ...]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

BOUND_SHAPES = [(0, 20, 5, 3000), (1, 20, 10, 3000), (2, 50, 20, 600), (3, 9, 4, 3000), (6, 30, 17, 600)]

# (seed, n, m, bound_mode, branching, refine, johnson_pairs, ub_kind)
SOLVE_CASES = []
for s in range(24):
    n, m = 8 + s % 5, 3 + s % 4
    for (bm, br, rf, jp) in [(0, 1, 0, 0), (0, 2, 0, 0), (0, 3, 0, 0), (0, -1, 0, 0), (2, 1, 0, 0),
                             (2, 3, 1, 0), (2, 2, 1, 1), (1, 1, 0, 0), (1, 3, 1, 2)]:
        SOLVE_CASES.append((9000 + s, n, m, bm, br, rf, jp, "opt"))
        SOLVE_CASES.append((9000 + s, n, m, bm, br, rf, jp, "neh"))
for (bm, br, rf, jp) in [(0, 1, 0, 0), (0, 2, 0, 0), (0, 3, 0, 0)]:
    SOLVE_CASES.append((0, 20, 5, bm, br, rf, jp, "neh"))
    SOLVE_CASES.append((0, 20, 5, bm, br, rf, jp, "opt"))
for (bm, br, rf, jp) in [(2, 1, 1, 0), (2, 2, 1, 0), (2, 3, 1, 0), (2, 1, 0, 0), (0, 3, 0, 0)]:
    SOLVE_CASES.append((1, 20, 10, bm, br, rf, jp, "opt"))
SOLVE_CASES.append((42, 16, 8, 0, 1, 0, 0, "opt"))
SOLVE_CASES.append((3, 16, 6, 2, 3, 1, 0, "opt"))


def ref_bounds_fixture(seed, n, m, count):
    ptm = O.synth_ptm(seed, n, m)
    rows = O.ref_bounds(ptm, count, 77 + seed, pairs=0)
    arr = lambda i: np.array([r[i] for r in rows], dtype=np.int32)
    path = os.path.join(HERE, f"bounds_{seed}_{n}x{m}.npz")
    np.savez_compressed(path, ptm=ptm, l1=arr(0), l2=arr(1), perm=arr(2), cb=arr(3), ce=arr(4),
                        lb1n=arr(5), lb2=arr(6), idle=arr(7), mk=arr(8))
    return path


def tie_instances():
    """Instances whose NEH runs through ties: job totals that repeat (the
    stable initial sort) and insertion positions of equal makespan (the
    first-minimum rule), from a seeded stream.  Returned with their times."""
    rng = np.random.default_rng(4711)
    out = []
    for i in range(48):
        n = int(rng.integers(3, 31))
        m = int(rng.integers(2, 11))
        kind = i % 4
        if kind == 0:    # every time equal: all totals and all positions tie
            ptm = np.full((m, n), int(rng.integers(1, 100)), dtype=np.int32)
        elif kind == 1:  # times in {1, 2}
            ptm = rng.integers(1, 3, size=(m, n)).astype(np.int32)
        elif kind == 2:  # columns that are machine permutations of one job: equal totals
            base = rng.integers(1, 100, size=m)
            ptm = np.stack([rng.permutation(base) for _ in range(n)], axis=1).astype(np.int32)
        else:            # a few distinct jobs, each repeated
            kinds = rng.integers(1, 100, size=(m, max(1, n // 4)))
            ptm = kinds[:, rng.integers(0, kinds.shape[1], size=n)].astype(np.int32)
        out.append(np.ascontiguousarray(ptm))
    return out


def neh_fixture():
    cases = []  # (kind, seed, ptm, store_ptm)
    cases.append(("c1_20x5", 0, O.synth_ptm(0, 20, 5), False))
    for seed in range(100, 612):
        n, m = O.synth_shape(seed, 10, 22, (5, 10))
        cases.append(("batch_c4", seed, O.synth_ptm(seed, n, m), False))
    for seed in range(20000, 20256):
        n, m = O.synth_shape(seed, 12, 16, (5, 10, 20))
        cases.append(("serve_small", seed, O.synth_ptm(seed, n, m), False))
    for n in range(3, 64):
        m = 2 + (7 * n) % 19
        cases.append(("random", 30000 + n, O.synth_ptm(30000 + n, n, m), False))
    for i, ptm in enumerate(tie_instances()):
        cases.append(("ties", i, ptm, True))
    got = O.ref_neh_many([c[2] for c in cases])
    rows = []
    for (kind, seed, ptm, store), (cost, perm) in zip(cases, got):
        d = dict(kind=kind, seed=seed, n=int(ptm.shape[1]), m=int(ptm.shape[0]), cost=cost, perm=perm)
        if store:
            d["ptm"] = ptm.tolist()
        rows.append(d)
    with open(os.path.join(HERE, "neh_cases.json"), "w") as f:
        json.dump(rows, f, separators=(",", ":"))
    print("neh cases", len(rows), flush=True)


# findAll (multicore/src/ivm_bound.cpp:678-682): pruning with > instead of
# >=, with the bound frozen at the optimum, so every optimal schedule's branch
# is explored.  (seed, n, m, bound_mode, branching, refine, johnson_pairs)
FINDALL_CASES = [(9100 + s, 7 + s % 4, 3 + s % 3, bm, br, rf, jp)
                 for s in range(8)
                 for (bm, br, rf, jp) in [(0, 1, 0, 0), (0, 2, 0, 0), (0, 3, 0, 0), (0, -1, 0, 0),
                                          (2, 1, 0, 0), (2, 3, 1, 0), (1, 1, 0, 0), (1, 3, 1, 2)]]


def findall_fixture():
    rows = []
    for (seed, n, m, bm, br, rf, jp) in FINDALL_CASES:
        ptm = O.synth_ptm(seed, n, m)
        neh, _ = O.Instance(ptm).neh()
        opt = O.ref_solve(ptm, neh, 0, 1).best
        r = O.ref_solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp, find_all=True)
        assert r.finished
        rows.append(dict(seed=seed, n=n, m=m, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp,
                         ub=opt, best=r.best, decomposed=r.decomposed, leaves=r.leaves,
                         simple_bounds=r.simple_bounds, johnson_bounds=r.johnson_bounds))
    with open(os.path.join(HERE, "findall_cases.json"), "w") as f:
        json.dump(rows, f, indent=0)
    print("findall cases", len(rows), flush=True)


def main():
    if not O.have_ref():
        O.build(ref=True)
    which = set(sys.argv[1:]) or {"bounds", "solve", "taillard", "neh", "findall"}
    if "neh" in which:
        neh_fixture()
    if "findall" in which:
        findall_fixture()
    if not which & {"bounds", "solve", "taillard"}:
        return
    for seed, n, m, count in BOUND_SHAPES:
        print("bounds", seed, n, m, ref_bounds_fixture(seed, n, m, count), flush=True)

    cases = []
    opt_cache = {}
    for (seed, n, m, bm, br, rf, jp, ubk) in SOLVE_CASES:
        ptm = O.synth_ptm(seed, n, m)
        neh_ref_free, _ = O.Instance(ptm).neh()
        if (seed, n, m) not in opt_cache:
            opt_cache[(seed, n, m)] = O.ref_solve(ptm, neh_ref_free, 0, 1).best
        ub = opt_cache[(seed, n, m)] if ubk == "opt" else neh_ref_free
        r = O.ref_solve(ptm, ub, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
        cases.append(dict(seed=seed, n=n, m=m, bound_mode=bm, branching=br, refine=rf,
                          johnson_pairs=jp, ub=ub, ub_kind=ubk, best=r.best, decomposed=r.decomposed,
                          leaves=r.leaves, simple_bounds=r.simple_bounds,
                          johnson_bounds=r.johnson_bounds))
    with open(os.path.join(HERE, "solve_cases.json"), "w") as f:
        json.dump(cases, f, indent=0)
    print("solve cases", len(cases), flush=True)

    tai = {}
    for i in range(1, 121):
        ptm = O.ref_taillard(i)
        tai[str(i)] = dict(n=int(ptm.shape[1]), m=int(ptm.shape[0]),
                           sha256=hashlib.sha256(ptm.astype("<i4").tobytes()).hexdigest())
    with open(os.path.join(HERE, "taillard_ptm.json"), "w") as f:
        json.dump(tai, f, indent=0)
    print("taillard", len(tai), flush=True)


if __name__ == "__main__":
    main()
