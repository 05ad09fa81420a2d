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
