"""Pins the CPU oracle to the reference's own outputs (CPU only).

Fixtures in tests/golden were written by tests/golden/make_golden.py from the
reference's bound, IVM and bound-driver sources compiled unmodified
(oracle/_ref), plus the optima the reference ships in
parameters/instances.data (Taillard ta001-ta020).
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# parameters/instances.data:4-24 (optimal makespans of ta001..ta020)
TAILLARD_OPT = {1: 1278, 2: 1359, 3: 1081, 4: 1293, 5: 1235, 6: 1195, 7: 1234, 8: 1206, 9: 1230,
                10: 1108, 11: 1582, 12: 1659, 13: 1496, 14: 1377, 15: 1419, 16: 1397, 18: 1538,
                19: 1593, 20: 1591}


def _bound_files():
    return sorted(f for f in os.listdir(GOLDEN) if f.startswith("bounds_") and f.endswith(".npz"))


@pytest.mark.parametrize("fname", _bound_files())
def test_oracle_bounds_match_reference(oracle_mod, fname):
    d = np.load(os.path.join(GOLDEN, fname))
    inst = oracle_mod.Instance(d["ptm"])
    out = inst.bounds_batch(d["perm"], d["l1"], d["l2"], pairs=0)
    np.testing.assert_array_equal(out["cb"], d["cb"])
    np.testing.assert_array_equal(out["ce"], d["ce"])
    np.testing.assert_array_equal(out["lb1n"], d["lb1n"])
    np.testing.assert_array_equal(out["lb2"], d["lb2"])
    np.testing.assert_array_equal(out["mk"], d["mk"])
    # idle sum couts[1] (written only when more than one job is unscheduled)
    for i in range(0, len(d["l1"]), 7):
        l1, l2 = int(d["l1"][i]), int(d["l2"][i])
        if l2 - l1 > 2:
            assert inst.lb2_node(d["perm"][i], l1, l2, 0)[1] == d["idle"][i]


def test_oracle_solves_match_reference(oracle_mod):
    with open(os.path.join(GOLDEN, "solve_cases.json")) as f:
        cases = json.load(f)
    assert len(cases) > 400
    for c in cases:
        ptm = oracle_mod.synth_ptm(c["seed"], c["n"], c["m"])
        r = oracle_mod.solve(ptm, c["ub"], bound_mode=c["bound_mode"], branching=c["branching"],
                             refine=c["refine"], johnson_pairs=c["johnson_pairs"])
        got = (r.best, r.decomposed, r.leaves, r.simple_bounds, r.johnson_bounds)
        want = (c["best"], c["decomposed"], c["leaves"], c["simple_bounds"], c["johnson_bounds"])
        assert got == want, (c, got)


def test_taillard_generator_matches_reference(oracle_mod):
    with open(os.path.join(GOLDEN, "taillard_ptm.json")) as f:
        ref = json.load(f)
    for i in range(1, 121):
        ptm = oracle_mod.taillard(i)
        h = hashlib.sha256(ptm.astype("<i4").tobytes()).hexdigest()
        assert (ptm.shape[1], ptm.shape[0], h) == (ref[str(i)]["n"], ref[str(i)]["m"], ref[str(i)]["sha256"])


@pytest.mark.parametrize("ta", sorted(TAILLARD_OPT))
def test_taillard_known_optima(oracle_mod, ta):
    """UB = optimum + 1: the search must find a schedule of exactly the optimum."""
    ptm = oracle_mod.taillard(ta)
    r = oracle_mod.solve(ptm, TAILLARD_OPT[ta] + 1, bound_mode=0, branching=1)
    assert r.finished and r.best == TAILLARD_OPT[ta]
    assert oracle_mod.Instance(ptm).makespan(r.perm) == r.best


def _neh_cases(oracle_mod):
    with open(os.path.join(GOLDEN, "neh_cases.json")) as f:
        cases = json.load(f)
    for c in cases:
        if "ptm" in c:
            ptm = np.asarray(c["ptm"], dtype=np.int32)
        else:
            ptm = oracle_mod.synth_ptm(c["seed"], c["n"], c["m"])
        yield c, np.ascontiguousarray(ptm)


def test_oracle_neh_matches_reference(oracle_mod):
    """orc_neh gives the reference fastNEH's order and makespan on every
    fixture case: the 20x5 instance, the 512 batched instances, the serving
    stream, n = 3..63, and instances built to tie (equal totals, equal
    insertion makespans)."""
    kinds = set()
    for c, ptm in _neh_cases(oracle_mod):
        cost, perm = oracle_mod.Instance(ptm).neh()
        assert (cost, list(perm)) == (c["cost"], c["perm"]), (c["kind"], c["seed"])
        kinds.add(c["kind"])
    assert kinds == {"c1_20x5", "batch_c4", "serve_small", "random", "ties"}


def _neh_variant(ptm, first_min=True, stable=True, keep_two_if_not_worse=False):
    """A small NEH written out in Python with one rule switched, to show the
    fixture separates the rules it pins (the reference's rules are
    first_min=True, stable=True, keep_two_if_not_worse=False)."""
    m, n = ptm.shape

    def mk(seq):
        c = [0] * m
        for j in seq:
            c[0] += int(ptm[0, j])
            for k in range(1, m):
                c[k] = max(c[k - 1], c[k]) + int(ptm[k, j])
        return c[-1]

    tot = ptm.sum(axis=0)
    order = sorted(range(n), key=lambda j: (-int(tot[j]), j if stable else -j))
    a, b = order[0], order[1]
    keep = mk([a, b]) <= mk([b, a]) if keep_two_if_not_worse else mk([a, b]) < mk([b, a])
    seq = [a, b] if keep else [b, a]
    for job in order[2:]:
        costs = [mk(seq[:p] + [job] + seq[p:]) for p in range(len(seq) + 1)]
        best = min(costs)
        pos = costs.index(best) if first_min else len(costs) - 1 - costs[::-1].index(best)
        seq.insert(pos, job)
    return mk(seq), seq


def test_neh_fixture_separates_tie_rules(oracle_mod):
    """The tie cases pin each tie rule: switching any one of them changes the
    order on some fixture instance, while the rules as stated reproduce all."""
    cases = [(c, p) for c, p in _neh_cases(oracle_mod) if c["kind"] in ("ties", "c1_20x5")]
    for c, ptm in cases:
        assert _neh_variant(ptm) == (c["cost"], c["perm"])
    for kw in (dict(first_min=False), dict(stable=False), dict(keep_two_if_not_worse=True)):
        assert any(_neh_variant(ptm, **kw)[1] != c["perm"] for c, ptm in cases), kw


def test_host_neh_matches_reference(oracle_mod):
    """The product's host NEH (pbb_neh) on the same fixture."""
    import pbbperm_amd as P
    for c, ptm in _neh_cases(oracle_mod):
        cost, perm = P.neh(ptm)
        assert (cost, list(perm)) == (c["cost"], c["perm"]), (c["kind"], c["seed"])


def test_oracle_findall_solves_match_reference(oracle_mod):
    """findAll (prune on >, multicore/src/ivm_bound.cpp:678-682) with the
    bound frozen at the optimum: the oracle's counts equal the reference's,
    and the fixture tells the two pruning rules apart (>= gives fewer nodes
    on most cases)."""
    with open(os.path.join(GOLDEN, "findall_cases.json")) as f:
        cases = json.load(f)
    differ = 0
    for c in cases:
        ptm = oracle_mod.synth_ptm(c["seed"], c["n"], c["m"])
        kw = dict(bound_mode=c["bound_mode"], branching=c["branching"], refine=c["refine"],
                  johnson_pairs=c["johnson_pairs"])
        r = oracle_mod.solve(ptm, c["ub"], find_all=True, **kw)
        got = (r.best, r.decomposed, r.leaves, r.simple_bounds, r.johnson_bounds)
        want = (c["best"], c["decomposed"], c["leaves"], c["simple_bounds"], c["johnson_bounds"])
        assert got == want, (c, got)
        differ += oracle_mod.solve(ptm, c["ub"], **kw).decomposed != c["decomposed"]
    assert differ > len(cases) // 2
