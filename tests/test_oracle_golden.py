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


def _ig_moves():
    with open(os.path.join(GOLDEN, "ig_moves.json")) as f:
        return json.load(f)["shapes"]


def test_oracle_ig_moves_match_reference(oracle_mod):
    """The oracle's iterated-greedy moves against the reference's own
    fastInsertRemove (tests/golden/ig_moves.json, written by
    make_ig_golden.py from oracle/_ref): the IG visit order, and the
    insertion and removal makespans of every sampled partial sequence, for
    n = 3..63 and m = 2..20."""
    O = oracle_mod
    for sh in _ig_moves():
        ptm = O.synth_ptm(sh["seed"], sh["n"], sh["m"])
        inst = O.Instance(ptm)
        assert list(O.ig_visit_order(sh["n"])) == sh["visit_order"]
        for s in sh["samples"]:
            assert list(inst.insert_makespans(s["seq"], s["job"])) == s["ins"], (sh["n"], sh["m"], s)
            assert list(inst.remove_makespans(s["seq"])) == s["rem"], (sh["n"], sh["m"], s)


def test_oracle_ig_moves_end_positions_by_recomputation(oracle_mod):
    """The fixture's end cases, recomputed from scratch by the textbook
    recurrence: inserting the job after the whole sequence (the cell the
    reference leaves unwritten, here zero), before it, and removing the
    first and the last job; every insertion is no shorter and every removal
    no longer than the sequence itself."""
    O = oracle_mod
    for sh in _ig_moves():
        inst = O.Instance(O.synth_ptm(sh["seed"], sh["n"], sh["m"]))
        for s in sh["samples"]:
            seq, job = s["seq"], s["job"]
            whole = _partial(inst, seq)
            assert s["ins"][-1] == _partial(inst, seq + [job])
            assert s["ins"][0] == _partial(inst, [job] + seq)
            assert s["rem"][0] == _partial(inst, seq[1:]) and s["rem"][-1] == _partial(inst, seq[:-1])
            assert min(s["ins"]) >= whole and max(s["rem"]) <= whole


def _partial(inst, seq):
    c = np.zeros(inst.m, np.int64)
    for j in seq:
        c[0] += inst.ptm[0, j]
        for k in range(1, inst.m):
            c[k] = max(c[k - 1], c[k]) + inst.ptm[k, j]
    return int(c[-1])


def test_oracle_ig_properties(oracle_mod):
    """orc_ig (IG.cpp runIG with a seeded stream): never worse than NEH, its
    order realises its cost, the same seed gives the same result, and on 40
    tiny instances (n = 6..8) it reaches the brute-force optimum on at
    least 38 (a broken move, acceptance or destruction loses many)."""
    import itertools
    O = oracle_mod
    hits = 0
    for seed in range(40):
        n, m = 6 + seed % 3, 3 + seed % 4
        ptm = O.synth_ptm(1000 + seed, n, m)
        inst = O.Instance(ptm)
        opt = min(inst.makespan(np.array(p, np.intc)) for p in itertools.permutations(range(n)))
        st = O.ig_chain_seed(3, seed, 0)
        c, p = inst.ig(st, 30)
        assert (c, list(p)) == tuple((lambda r: (r[0], list(r[1])))(inst.ig(st, 30)))
        neh, _ = inst.neh()
        assert opt <= c <= neh and inst.makespan(p) == c and sorted(p) == list(range(n))
        hits += c == opt
    assert hits >= 38
    for seed, n, m in [(0, 20, 5), (2, 50, 20), (11, 63, 20)]:
        inst = O.Instance(O.synth_ptm(seed, n, m))
        neh, _ = inst.neh()
        c, p = inst.ig(O.ig_chain_seed(5, 0, 0), 3)
        assert c <= neh and inst.makespan(p) == c


def test_oracle_ig_thresholds(oracle_mod):
    """Acceptance thresholds of IG::acceptance: exp(-d / T) scaled to 2^32,
    T = 0.2 * mean time / 10, equal costs always accepted, decreasing."""
    O = oracle_mod
    ptm = O.synth_ptm(4, 20, 10)
    thr = O.Instance(ptm).ig_thresholds()
    T = 0.2 * ptm.mean() / 10.0
    assert int(thr[0]) == 1 << 32
    for d in (1, 2, 5):
        assert abs(int(thr[d]) - np.floor(np.exp(-d / T) * 2.0 ** 32)) <= 1
    assert all(int(a) >= int(b) for a, b in zip(thr[:-1], thr[1:]))
