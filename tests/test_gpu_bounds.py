"""Bit-exact parity of the batched device bounds (the explorer's own bound
routines) against the CPU oracle on seeded random partial schedules.

Covers bound_fsp_weak::boundChildren (both directions and front only),
bound_fsp_weak::bornes_calculer (LB1'), bound_fsp_strong::bornes_calculer
(Johnson LB2, machine-pair modes 0/1/2) and evalSolution, including the edge
cases the reference special-cases: empty front (l1=-1), empty back (l2=n),
l2-l1 in {1, 2}.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(0, 20, 5), (1, 20, 10), (2, 50, 20), (3, 12, 4), (4, 63, 3), (5, 7, 2), (6, 30, 17)]


def _to_dev(*arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda() for a in arrs]


@pytest.mark.parametrize("seed,n,m", SHAPES)
def test_bounds_match_oracle(oracle_mod, seed, n, m):
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(seed, n, m)
    inst = O.Instance(ptm)
    perms, l1, l2 = O.random_partials(n, 4000, 1000 + seed, edge_frac=0.05)
    ref = inst.bounds_batch(perms, l1, l2, pairs=0)
    b = P.Bounds(ptm)
    dp, d1, d2 = _to_dev(perms, l1, l2)
    cb, ce = b.lb1_children(dp, d1, d2, both=True)
    np.testing.assert_array_equal(cb.cpu().numpy(), ref["cb"])
    np.testing.assert_array_equal(ce.cpu().numpy(), ref["ce"])
    cbf, _ = b.lb1_children(dp, d1, d2, both=False)
    np.testing.assert_array_equal(cbf.cpu().numpy(), ref["cbf"])
    np.testing.assert_array_equal(b.lb1_node(dp, d1, d2).cpu().numpy(), ref["lb1n"])
    np.testing.assert_array_equal(b.lb2_node(dp, d1, d2, 0).cpu().numpy(), ref["lb2"])
    np.testing.assert_array_equal(b.makespan(dp).cpu().numpy(), ref["mk"])
    for pairs in (1, 2):
        refp = inst.bounds_batch(perms, l1, l2, pairs=pairs, which=("lb2",))
        np.testing.assert_array_equal(b.lb2_node(dp, d1, d2, pairs).cpu().numpy(), refp["lb2"])


def test_bounds_1m_partial_schedules_50x20(oracle_mod):
    """The 50x20 throughput instance: LB1 children, LB1' and makespan on 1M
    seeded partial schedules, LB2 on 100k."""
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(2, 50, 20)
    inst = O.Instance(ptm)
    perms, l1, l2 = O.random_partials(50, 1_000_000, 2024)
    ref = inst.bounds_batch(perms, l1, l2, which=("cb", "ce", "lb1n", "mk"))
    b = P.Bounds(ptm)
    dp, d1, d2 = _to_dev(perms, l1, l2)
    cb, ce = b.lb1_children(dp, d1, d2, both=True)
    assert np.array_equal(cb.cpu().numpy(), ref["cb"])
    assert np.array_equal(ce.cpu().numpy(), ref["ce"])
    assert np.array_equal(b.lb1_node(dp, d1, d2).cpu().numpy(), ref["lb1n"])
    assert np.array_equal(b.makespan(dp).cpu().numpy(), ref["mk"])
    k = 100_000
    ref2 = inst.bounds_batch(perms[:k], l1[:k], l2[:k], which=("lb2",))
    assert np.array_equal(b.lb2_node(dp[:k].contiguous(), d1[:k].contiguous(), d2[:k].contiguous()).cpu().numpy(),
                          ref2["lb2"])


def test_bounds_reject_bad_inputs():
    import torch
    import pbbperm_amd as P
    ptm = P.synth(0, 10, 3)
    b = P.Bounds(ptm)
    x = torch.zeros((4, 10), dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError):
        b.makespan(x)
    with pytest.raises(P.PbbError):
        P.Bounds(np.ones((3, 70), np.int32))
