"""Device iterated greedy (pbb_ig_device) against the CPU restatement of
the same seeded steps (the oracle's orc_ig, whose moves are pinned to the
reference's own fastInsertRemove), chain by chain: every chain's makespan
and order must be identical, and the kept result is the best chain's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_ig_matches_oracle_chain_by_chain(oracle_mod):
    import pbbperm_amd.pool as PP
    O = oracle_mod
    # one batch of uneven shapes (n = 3..63, m = 2..20: zero-time padding
    # up to the batch's largest m), a few chains each
    shapes = [(7, 3, 2), (3, 12, 3), (0, 20, 5), (1, 20, 10), (9, 30, 20), (11, 63, 20)]
    ptms = [O.synth_ptm(s, n, m) for s, n, m in shapes]
    chains, iters, seed = 4, 6, 12345
    best, costs, cc, cp = PP.ig_device(ptms, chains=chains, iterations=iters, seed=seed, per_chain=True)
    for i, (ptm, (s, n, m)) in enumerate(zip(ptms, shapes)):
        inst = O.Instance(ptm)
        neh, _ = inst.neh()
        for c in range(chains):
            want, wp = inst.ig(O.ig_chain_seed(seed, i, c), iters)
            print(f"instance {i} ({n}x{m}) chain {c}: device {cc[i, c]} oracle {want} (NEH {neh})", flush=True)
            assert cc[i, c] == want and list(cp[i][c]) == list(wp)
        assert costs[i] == cc[i].min() <= neh
        assert sorted(best[i].tolist()) == list(range(n)) and inst.makespan(best[i]) == costs[i]


def test_device_ig_many_chains_properties(oracle_mod):
    """64 chains on the 50x20 headline instance: every chain is no worse
    than NEH and its order realises its cost; a second call with the same
    seed repeats every chain exactly."""
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptm = O.synth_ptm(2, 50, 20)
    inst = O.Instance(ptm)
    neh, _ = inst.neh()
    best, costs, cc, cp = PP.ig_device([ptm], chains=64, iterations=4, seed=7, per_chain=True)
    again = PP.ig_device([ptm], chains=64, iterations=4, seed=7, per_chain=True)
    assert (again[2] == cc).all()
    for c in range(64):
        assert cc[0, c] <= neh and inst.makespan(cp[0][c]) == cc[0, c]
    print(f"50x20: NEH {neh}, IG best of 64 chains x 4 rounds {costs[0]}", flush=True)
    assert costs[0] < neh


def test_device_ig_rejects_bad_shapes():
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    with pytest.raises(P.PbbError):
        PP.ig_device([np.ones((3, 2), np.int32)], chains=2, iterations=1)
    with pytest.raises(P.PbbError):
        PP.ig_device([np.ones((3, 64), np.int32)], chains=2, iterations=1)
    with pytest.raises(P.PbbError):
        PP.ig_device([np.ones((3, 10), np.int32)], chains=0, iterations=1)
