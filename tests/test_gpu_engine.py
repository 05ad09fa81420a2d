"""GPU engine parity with the CPU oracle (sequentialbb semantics).

With the incumbent frozen at the optimum (every leaf of cost >= UB is pruned,
so the UB never moves), the set of decomposed nodes is independent of how
the interval is split among explorers, and the engine must reproduce the
oracle's optimum, decomposed-node, leaf, LB1-bound and refinement-bound
counts exactly.  With an improving incumbent only the optimum is fixed.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INT_MAX = 2**31 - 1

# (bound_mode, branching, refine, johnson_pairs)
MODES = [
    (0, 1, 0, 0), (0, 2, 0, 0), (0, 3, 0, 0), (0, -1, 0, 0),
    (2, 1, 0, 0), (2, 3, 1, 0), (2, 2, 1, 1), (2, 1, 1, 2),
    (1, 1, 0, 0), (1, 3, 1, 0), (0, -2, 0, 0), (0, -3, 0, 0),
]


def _cfg(P, bm, br, rf, jp, nb=4096, **kw):
    return P.default_config(bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp,
                            early_stop_johnson=0, nb_ivms=nb, **kw)


def _gpu(P, ptm, ub, bm, br, rf, jp, nb=4096, **kw):
    eng = P.GpuBB(ptm, _cfg(P, bm, br, rf, jp, nb, **kw), ub=ub)
    done = eng.solve(time_budget=60.0)
    assert done
    st = eng.stats()
    best, perm = eng.best()
    eng.close()
    return best, perm, st


def _opt(O, ptm):
    inst = O.Instance(ptm)
    c, _ = inst.neh()
    return O.solve(ptm, c, bound_mode=0, branching=1).best


def _same_counts(st, ref):
    return (st.decomposed, st.leaves, st.simple_bounds, st.johnson_bounds) == \
        (ref.decomposed, ref.leaves, ref.simple_bounds, ref.johnson_bounds)


@pytest.mark.parametrize("seed", range(10))
def test_small_instances_all_modes_frozen_ub(oracle_mod, seed):
    import pbbperm_amd as P
    O = oracle_mod
    n, m = 8 + seed % 6, 3 + seed % 4
    ptm = O.synth_ptm(5000 + seed, n, m)
    opt = _opt(O, ptm)
    for (bm, br, rf, jp) in MODES:
        if br in (-2, -3) and n > 9:
            continue  # back rows are never pruned in these modes: near-full enumeration
        ref = O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
        best, _, st = _gpu(P, ptm, opt, bm, br, rf, jp)
        print(f"seed={seed} n={n} m={m} mode={(bm, br, rf, jp)} dec={st.decomposed}/{ref.decomposed}",
              flush=True)
        assert best == ref.best == opt
        assert _same_counts(st, ref), (st, ref)


def test_find_all_frozen_ub_matches_reference():
    """findAll (prune on >, multicore/src/ivm_bound.cpp:678-682) with the
    bound frozen at the optimum, every bounding/branching mode of the
    reference fixture: counts equal the reference's single-thread run, for a
    single explorer and for many."""
    import json
    import os
    import pbbperm_amd as P
    with open(os.path.join(os.path.dirname(__file__), "golden", "findall_cases.json")) as f:
        cases = json.load(f)
    for c in cases:
        ptm = P.synth(c["seed"], c["n"], c["m"])
        want = (c["best"], c["decomposed"], c["leaves"], c["simple_bounds"], c["johnson_bounds"])
        for nb in (1, 4096):
            best, _, st = _gpu(P, ptm, c["ub"], c["bound_mode"], c["branching"], c["refine"], c["johnson_pairs"],
                               nb=nb, find_all=1)
            got = (best, st.decomposed, st.leaves, st.simple_bounds, st.johnson_bounds)
            assert got == want, (c, nb, got)


@pytest.mark.parametrize("seed,n,m", [(42, 16, 8), (3, 16, 6)])
def test_larger_trees_frozen_ub(oracle_mod, seed, n, m):
    """Trees of 10^4-10^6 nodes: work sharing is exercised for real."""
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(seed, n, m)
    opt = _opt(O, ptm)
    for (bm, br, rf, jp) in [(0, 1, 0, 0), (0, 3, 0, 0), (2, 3, 1, 0), (2, 1, 0, 0), (0, 2, 0, 0)]:
        ref = O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
        best, _, st = _gpu(P, ptm, opt, bm, br, rf, jp)
        print(f"seed={seed} mode={(bm, br, rf, jp)} dec={st.decomposed}/{ref.decomposed} "
              f"explorers={st.explorers}", flush=True)
        assert best == opt and _same_counts(st, ref), (st, ref)


@pytest.mark.parametrize("seed", range(6))
def test_small_instances_improving_ub(oracle_mod, seed):
    """UB from NEH (and from infinity): the optimum and a permutation realising it."""
    import pbbperm_amd as P
    O = oracle_mod
    n, m = 9 + seed % 4, 4 + seed % 3
    ptm = O.synth_ptm(7000 + seed, n, m)
    inst = O.Instance(ptm)
    neh, _ = inst.neh()
    for ub in (neh, INT_MAX):
        for (bm, br, rf, jp) in [(0, 1, 0, 0), (2, 3, 1, 0), (0, 2, 0, 0)]:
            ref = O.solve(ptm, ub, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
            best, perm, _ = _gpu(P, ptm, ub, bm, br, rf, jp)
            assert best == ref.best
            if best < ub:
                assert sorted(perm.tolist()) == list(range(n))
                assert inst.makespan(perm) == best


def test_explorer_count_invariance(oracle_mod):
    """Work stealing between explorers must not change the counts (frozen UB)."""
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(1, 20, 10)
    opt = _opt(O, ptm)
    ref = O.solve(ptm, opt, bound_mode=0, branching=3)
    for nb in (1, 2, 64, 512, 100000):
        best, _, st = _gpu(P, ptm, opt, 0, 3, 0, 0, nb=nb)
        print(f"nb_ivms={nb} explorers={st.explorers} dec={st.decomposed}", flush=True)
        assert best == opt and _same_counts(st, ref)


@pytest.mark.parametrize("mode", [(0, 3, 0, 0), (2, 1, 1, 0)])
def test_team_phase_hands_over_to_the_device(oracle_mod, mode):
    """A search from pbb_init_full starts on the one-workgroup team and, past
    its node budget, continues on the whole device from the team's saved
    intervals: with the incumbent frozen, counts equal the oracle's for every
    team budget (none, one that stops the team after its first check, one
    the tree outgrows, one it never reaches), with and without the donation
    ramp, and for a node-budgeted pbb_run that stops inside the team."""
    import pbbperm_amd as P
    O = oracle_mod
    bm, br, rf, jp = mode
    ptm = O.synth_ptm(1, 20, 10)
    opt = _opt(O, ptm)
    ref = O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
    print(f"mode={mode} reference decomposed={ref.decomposed}", flush=True)
    for team, ramp, nb in [(0, 0, 4096), (64, 0, 4096), (2048, 0, 4096), (1 << 30, 0, 4096), (2048, 48, 4096),
                           (2048, 0, 3)]:
        best, _, st = _gpu(P, ptm, opt, bm, br, rf, jp, nb=nb, team_nodes=team, ramp_nodes=ramp)
        print(f"  team={team} ramp={ramp} nb={nb} dec={st.decomposed}", flush=True)
        assert best == opt and _same_counts(st, ref)
    # a node budget smaller than the team's: the first pbb_run returns from
    # the team, the next ones continue on the device
    eng = P.GpuBB(ptm, _cfg(P, bm, br, rf, jp, 4096, team_nodes=8192), ub=opt)
    eng.init_full_interval()
    runs = 1
    while not eng.next(node_budget=500, time_budget=60.0):
        runs += 1
        assert runs < 100000
    st = eng.stats()
    eng.close()
    assert runs > 1 and _same_counts(st, ref)


def test_full_grid_at_the_lds_edge():
    """12x20 with LB2 and the Johnson table in LDS fills a CU's LDS to within
    2 KB at three workgroups.  The persistent grid must still be co-resident:
    with one workgroup per CU too many, the resident explorers idle while the
    rest hold work, and the solve runs into its deadline.  Counts are the
    committed frozen-bound reference counts of that request."""
    import json
    import os
    import pbbperm_amd as P
    with open(os.path.join(os.path.dirname(__file__), "golden", "serve_small.json")) as f:
        fx = json.load(f)
    sp = fx["spec"]
    d = next(x for x in fx["instances"] if x["seed"] == 20194)
    ptm = P.synth(d["seed"], d["n"], d["m"])
    for nb in (4096, 100000):
        eng = P.GpuBB(ptm, _cfg(P, sp["bound_mode"], sp["branching"], sp["refine"], sp["pairs"], nb), ub=d["opt"])
        done = eng.solve(time_budget=20.0)
        st = eng.stats()
        eng.close()
        print(f"nb_ivms={nb} explorers={st.explorers} done={done} dec={st.decomposed}", flush=True)
        assert done
        assert [st.decomposed, st.leaves, st.simple_bounds, st.johnson_bounds] == d["counts_frozen"]


def test_config1_20x5_lb1_neh(oracle_mod):
    """20 jobs x 5 machines, seed 0, LB1, UB from NEH, 4096 explorers."""
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(0, 20, 5)
    neh, _ = O.Instance(ptm).neh()
    assert P.neh(ptm)[0] == neh
    ref = O.solve(ptm, neh, bound_mode=0, branching=1)
    best, perm, st = _gpu(P, ptm, neh, 0, 1, 0, 0, nb=4096)
    assert best == ref.best
    assert P.makespan(ptm, perm) == best
    ref_opt = O.solve(ptm, ref.best, bound_mode=0, branching=1)
    _, _, st2 = _gpu(P, ptm, ref.best, 0, 1, 0, 0, nb=4096)
    assert _same_counts(st2, ref_opt)


@pytest.mark.parametrize("branching", [1, 2, 3])
def test_config2_20x10_lb1_johnson(oracle_mod, branching):
    """20 jobs x 10 machines, seed 1, LB1 + Johnson LB2 over all pairs, UB = optimum."""
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(1, 20, 10)
    opt = _opt(O, ptm)
    ref = O.solve(ptm, opt, bound_mode=2, branching=branching, refine=1, johnson_pairs=0)
    best, _, st = _gpu(P, ptm, opt, 2, branching, 1, 0, nb=16384)
    print(f"branching={branching} dec={st.decomposed} ref={ref.decomposed} jb={st.johnson_bounds}")
    assert best == opt and _same_counts(st, ref)


def test_depth_histogram(oracle_mod):
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(11, 14, 5)
    opt = _opt(O, ptm)
    inst = O.Instance(ptm)
    bb = O.BB(inst, opt, bound_mode=0, branching=1)
    bb.init_full()
    bb.run()
    ref = bb.depth_histo().tolist()
    _, _, st = _gpu(P, ptm, opt, 0, 1, 0, 0)
    assert st.depth_histo == ref


def test_budgeted_runs_resume(oracle_mod):
    """next() with small node budgets, repeated until done, equals one run."""
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(3, 16, 6)
    opt = _opt(O, ptm)
    ref = O.solve(ptm, opt, bound_mode=2, branching=3, refine=1)
    eng = P.GpuBB(ptm, _cfg(P, 2, 3, 1, 0), ub=opt)
    eng.init_full_interval()
    calls = 0
    while not eng.next(node_budget=2000, time_budget=60.0):
        calls += 1
        assert calls < 10000
    st = eng.stats()
    eng.close()
    assert calls > 0
    assert _same_counts(st, ref)


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7])
def test_smallest_instances_all_modes_frozen_ub(oracle_mod, n):
    """n from 3 (the smallest the engine accepts) up: the root row, the
    last-line leaf and the l2 - l1 <= 2 LB2 case meet at once."""
    import pbbperm_amd as P
    O = oracle_mod
    for m in (2, 3, 5):
        ptm = O.synth_ptm(900 + 10 * n + m, n, m)
        opt = _opt(O, ptm)
        for (bm, br, rf, jp) in MODES:
            ref = O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
            best, _, st = _gpu(P, ptm, opt, bm, br, rf, jp, nb=64)
            assert best == opt and _same_counts(st, ref), (n, m, bm, br, rf, jp)


@pytest.mark.parametrize("m,tail,seed,mode", [(5, 12, 77, (0, 1, 0, 0)), (20, 14, 80, (0, 1, 0, 0)),
                                              (20, 14, 80, (2, 3, 1, 0))])
def test_largest_row_deep_interval(oracle_mod, m, tail, seed, mode):
    """n = 63, the longest IVM row a wavefront holds: one interval that fixes a
    random prefix and spans the last `tail` levels, given to one explorer of
    64 (the rest take work from it).  With the incumbent frozen at the
    interval's optimum the counts equal the oracle's (here the unfold down the
    prefix, bounded at full length, decides them); with an improving
    incumbent the engine finds that optimum and a schedule realising it
    (node counts then depend on the exploration order, so only the optimum is
    compared)."""
    import pbbperm_amd as P
    O = oracle_mod
    n = 63
    bm, br, rf, jp = mode
    ptm = O.synth_ptm(seed, n, m)
    inst = O.Instance(ptm)
    rng = np.random.default_rng(5)
    pre = [int(rng.integers(0, n - i)) for i in range(n - tail)]
    pos = np.array(pre + [0] * tail, np.intc)
    end = np.array(pre + list(range(tail - 1, -1, -1)), np.intc)
    bb = O.BB(inst, INT_MAX, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
    bb.init_interval(pos, end)
    bb.run()
    opt = bb.best()[0]
    assert bb.stats().decomposed > 50
    bb = O.BB(inst, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
    bb.init_interval(pos, end)
    bb.run()
    ref = bb.stats()
    cfg = _cfg(P, bm, br, rf, jp, nb=64)
    with P.GpuBB(ptm, cfg, ub=opt) as eng:
        eng.init_from_fac(np.array([0]), pos[None, :], end[None, :])
        assert eng.next(time_budget=60.0)
        assert _same_counts(eng.stats(), ref)
    with P.GpuBB(ptm, cfg) as eng:
        eng.init_from_fac(np.array([0]), pos[None, :], end[None, :])
        assert eng.next(time_budget=60.0)
        best, perm = eng.best()
        print(f"n=63 m={m} tail={tail} optimum={opt} found={best} dec={eng.stats().decomposed}", flush=True)
    assert best == opt
    assert sorted(perm.tolist()) == list(range(n)) and inst.makespan(perm) == opt


def test_factoradic_intervals_import(oracle_mod):
    """initFromFac with the full interval cut in pieces: each piece is unfolded
    (bounded, uncounted); counts equal the oracle summed over the pieces."""
    import pbbperm_amd as P
    from pbbperm_amd import factoradic as F
    O = oracle_mod
    n, m = 16, 8
    ptm = O.synth_ptm(42, n, m)
    opt = _opt(O, ptm)
    pos, end = F.split(n, 37)
    inst = O.Instance(ptm)
    tot = [0, 0, 0]
    for p, e in zip(pos, end):
        bb = O.BB(inst, opt, bound_mode=0, branching=1)
        bb.init_interval(p, e)
        bb.run()
        st = bb.stats()
        tot[0] += st.decomposed
        tot[1] += st.leaves
        tot[2] += st.simple_bounds
    eng = P.GpuBB(ptm, _cfg(P, 0, 1, 0, 0, nb=256), ub=opt)
    eng.init_from_fac(np.arange(len(pos)), pos, end)
    assert eng.next(time_budget=60.0)
    st = eng.stats()
    eng.close()
    assert [st.decomposed, st.leaves, st.simple_bounds] == tot


def test_export_import_intervals(oracle_mod):
    """Interrupted run -> getIntervals -> fresh engine initFromFac: the node
    total equals the first part plus the oracle over the exported pieces."""
    import pbbperm_amd as P
    O = oracle_mod
    n, m = 16, 6
    ptm = O.synth_ptm(3, n, m)
    opt = _opt(O, ptm)
    inst = O.Instance(ptm)
    eng = P.GpuBB(ptm, _cfg(P, 0, 1, 0, 0, nb=128), ub=opt)
    eng.init_full_interval()
    done = eng.next(node_budget=20000, time_budget=60.0)
    assert not done
    ids, pos, end, dec1, _ = eng.get_intervals()
    eng.close()
    assert len(ids) > 0
    expect = dec1
    for p, e in zip(pos, end):
        bb = O.BB(inst, opt, bound_mode=0, branching=1)
        bb.init_interval(p, e)
        bb.run()
        expect += bb.stats().decomposed
    eng2 = P.GpuBB(ptm, _cfg(P, 0, 1, 0, 0, nb=256), ub=opt)
    eng2.init_from_fac(np.arange(len(ids)), pos, end)
    assert eng2.next(time_budget=60.0)
    assert dec1 + eng2.stats().decomposed == expect
    eng2.close()


def test_checkpoint_save_load_resume(oracle_mod, tmp_path):
    """works::save after a budgeted run, works::init_infile into a fresh
    engine (incumbent and intervals from the file), run to the end.  A
    restored interval is unfolded along its pos (bounded, uncounted, as
    sequentialbb::initAtInterval does), so the expected total is the file's
    explored count plus the oracle's count over the same restored intervals;
    the optimum is unchanged."""
    import pbbperm_amd as P
    O = oracle_mod
    for (seed, n, m, mode) in [(42, 16, 8, (0, 1, 0, 0)), (3, 16, 6, (2, 3, 1, 0))]:
        ptm = O.synth_ptm(seed, n, m)
        opt = _opt(O, ptm)
        bm, br, rf, jp = mode
        inst = O.Instance(ptm)
        ref = O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
        eng = P.GpuBB(ptm, _cfg(P, bm, br, rf, jp, nb=256), ub=opt)
        eng.init_full_interval()
        assert not eng.next(node_budget=max(1000, ref.decomposed // 5), time_budget=60.0)
        path = str(tmp_path / f"bab{seed}.save")
        eng.save_checkpoint(path)
        first = eng.stats()
        eng.close()
        with open(path) as f:
            hdr, ids, pos, end = P.checkpoint_parse(f.read(), n)
        assert hdr["nb"] > 0
        expect = first.decomposed
        for p, e in zip(pos, end):
            bb = O.BB(inst, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
            bb.init_interval(p, e)
            bb.run()
            expect += bb.stats().decomposed
        eng2 = P.GpuBB(ptm, _cfg(P, bm, br, rf, jp, nb=512))
        info = eng2.load_checkpoint(path)
        assert info["explored"] == first.decomposed and info["best_cost"] == opt
        assert info["simple_bounds"] == first.simple_bounds
        # the file's counters carry on into the engine (works::init_infile)
        assert eng2.stats().decomposed == first.decomposed
        assert eng2.stats().simple_bounds == first.simple_bounds
        assert eng2.next(time_budget=60.0)
        assert eng2.stats().decomposed == expect
        # a second save holds the whole run's counts, not the last segment's
        eng2.save_checkpoint(path + ".2")
        with open(path + ".2") as f:
            hdr2, _, _, _ = P.checkpoint_parse(f.read(), n)
        assert hdr2["explored"] == expect and hdr2["nb"] == 0
        assert eng2.best()[0] == opt
        eng2.close()


def test_deep_subproblems(oracle_mod):
    """getDeepSubproblem after a budgeted run: complete permutations of the
    active explorers' current schedules, sorted by makespan; each schedule's
    prefix/suffix is the explorer's exported interval position."""
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(2, 50, 20)
    neh, _ = O.Instance(ptm).neh()
    eng = P.GpuBB(ptm, _cfg(P, 0, 1, 0, 0, nb=1024), ub=neh)
    eng.init_full_interval()
    assert not eng.next(node_budget=2_000_000, time_budget=60.0)
    ids, pos, end, _, _ = eng.get_intervals()
    sched = eng.get_deep_subproblems(64)
    eng.close()
    assert 0 < len(sched) <= min(64, len(ids))
    costs = [P.makespan(ptm, s) for s in sched]
    assert costs == sorted(costs)
    for s in sched:
        assert sorted(s.tolist()) == list(range(50))
    assert len(P.GpuBB(ptm, _cfg(P, 0, 1, 0, 0, nb=64), ub=neh).get_deep_subproblems(8)) == 0


def test_interrupt_and_continue(oracle_mod):
    import time
    import pbbperm_amd as P
    O = oracle_mod
    ptm = O.synth_ptm(2, 50, 20)
    neh, _ = O.Instance(ptm).neh()
    eng = P.GpuBB(ptm, _cfg(P, 0, 1, 0, 0), ub=neh)
    eng.init_full_interval()
    t = threading.Timer(0.5, eng.interrupt_exploration)
    t.start()
    t0 = time.time()
    done = eng.next(time_budget=30.0)
    el = time.time() - t0
    t.join()
    assert not done and el < 10.0
    st = eng.stats()
    assert st.decomposed > 0 and st.active > 0
    assert not eng.next(time_budget=0.3)
    assert eng.stats().decomposed > st.decomposed
    eng.close()


def test_engine_rejects_bad_config():
    import pbbperm_amd as P
    ptm = P.synth(0, 10, 3)
    with pytest.raises(P.PbbError):
        P.GpuBB(ptm, P.default_config(sorted_dfs=1))
    with pytest.raises(P.PbbError):
        P.GpuBB(ptm, P.default_config(johnson_pairs=3))
    eng = P.GpuBB(ptm, P.default_config(nb_ivms=8))
    with pytest.raises(P.PbbError):
        eng.init_from_fac([0], np.full((1, 10), 9), np.zeros((1, 10)))
    with pytest.raises(P.PbbError):
        eng.next()
    eng.close()


def test_host_coworker_shares_one_search_with_the_engine(oracle_mod):
    """The engine and the host co-worker on one search with the incumbent
    frozen at the optimum: after a budgeted engine run, its exported
    intervals are split between the host threads and the engine, both run
    to the end at the same time, and the counts of both sides add up to the
    engine's first segment plus the oracle's counts of every exported
    interval (sequentialbb::initAtInterval semantics on each).  Then
    pbb_hybrid_run from the root with an improving incumbent finds the
    optimum and a schedule realising it."""
    import threading
    import pbbperm_amd as P
    O = oracle_mod
    for (seed, n, m, mode) in [(42, 16, 8, (0, 1, 0, 0)), (3, 16, 6, (2, 3, 1, 0))]:
        bm, br, rf, jp = mode
        ptm = O.synth_ptm(seed, n, m)
        opt = _opt(O, ptm)
        print(f"hybrid seed={seed} mode={mode}", flush=True)
        cfg = _cfg(P, bm, br, rf, jp, nb=512)
        eng = P.GpuBB(ptm, cfg, ub=opt)
        eng.init_full_interval()
        assert not eng.next(node_budget=20000, time_budget=60.0)
        ids, pos, end, dec1, _ = eng.get_intervals()
        assert len(ids) >= 4
        want = dec1 + sum(O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp,
                                  pos=p, end=e).decomposed for p, e in zip(pos, end))
        h = min(8, len(ids) // 2)
        cpu = P.CpuBB(ptm, cfg, threads=8, ub=opt)
        cpu.init_from_fac(np.arange(h), pos[:h], end[:h])
        eng.init_from_fac(np.arange(len(ids) - h), pos[h:], end[h:])
        res = {}
        t = threading.Thread(target=lambda: res.update(cpu=cpu.next(time_budget=120.0)))
        t.start()
        assert eng.next(time_budget=60.0)
        t.join()
        assert res["cpu"]
        # the engine's counters restart at pbb_init_intervals: the first
        # segment (dec1) is added back, as in test_export_import_intervals
        gdec, cdec = eng.stats().decomposed, cpu.stats().decomposed
        print(f"  first segment {dec1}, engine {gdec}, host {cdec}, oracle total {want}", flush=True)
        assert dec1 + gdec + cdec == want
        eng.close()
        cpu.close()
    ptm = O.synth_ptm(42, 16, 8)
    opt = _opt(O, ptm)
    nehc, nehp = O.Instance(ptm).neh()
    cfg = _cfg(P, 0, 1, 0, 0, nb=1024)
    with P.GpuBB(ptm, cfg, ub=nehc, ub_perm=nehp) as eng, P.CpuBB(ptm, cfg, threads=8, ub=nehc, ub_perm=nehp) as cpu:
        eng.init_full_interval()
        done, gdec, cdec = P.hybrid_run(eng, cpu, period=0.002, time_budget=120.0)
        assert done and gdec > 0
        gb, gp = eng.best()
        cb, cp = cpu.best()
        assert min(gb, cb) == opt
        best_perm = gp if gb <= cb else cp
        assert O.Instance(ptm).makespan(best_perm) == opt
