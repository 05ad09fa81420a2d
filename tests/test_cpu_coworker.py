"""The host co-worker (pbb_cpu_*): IVM explorers on host threads that take
the engine's factoradic work units.  Runs without a GPU.

Its node processing must equal the reference's, so the tests compare it with
the reference's own fixtures (solve_cases.json, findall_cases.json: the
reference's single-thread runs) and with the oracle over intervals; with the
incumbent frozen, counts must not depend on the thread count (threads split
intervals among themselves as bbthread::shareWork does) nor on how the
search space is cut into intervals and handed over."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cfg(P, bm, br, rf, jp, **kw):
    return P.default_config(bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp, early_stop_johnson=0, **kw)


def _counts(st):
    return (st.decomposed, st.leaves, st.simple_bounds, st.johnson_bounds)


def test_cpu_coworker_matches_reference_solves():
    """Every single-thread solve of the reference fixture (all bounding and
    branching modes, frozen and NEH incumbents): optimum and counts equal
    on one thread; frozen-incumbent counts equal on four threads."""
    import pbbperm_amd as P
    with open(os.path.join(GOLDEN, "solve_cases.json")) as f:
        cases = json.load(f)
    with open(os.path.join(GOLDEN, "findall_cases.json")) as f:
        cases += [dict(c, ub_kind="opt", find_all=1) for c in json.load(f)]
    checked = 0
    for c in cases:
        if c["n"] > 16:
            continue  # the 20-job cases take seconds on one thread; covered on the GPU
        ptm = P.synth(c["seed"], c["n"], c["m"])
        cfg = _cfg(P, c["bound_mode"], c["branching"], c["refine"], c["johnson_pairs"],
                   find_all=c.get("find_all", 0))
        want = (c["decomposed"], c["leaves"], c["simple_bounds"], c["johnson_bounds"])
        threads = (1, 4) if c["ub_kind"] == "opt" else (1,)
        for t in threads:
            with P.CpuBB(ptm, cfg, threads=t, ub=c["ub"]) as cpu:
                assert cpu.solve()
                best, perm = cpu.best()
                st = cpu.stats()
            assert best == c["best"], (c, t)
            assert _counts(st) == want, (c, t, _counts(st))
            if best < c["ub"]:
                assert P.makespan(ptm, perm) == best
            checked += 1
    assert checked > 400


def test_cpu_coworker_interval_handover_matches_oracle(oracle_mod):
    """Intervals cut from the search space and handed to the co-worker: the
    sum of its counts over them equals the oracle's count of each interval
    explored on its own (sequentialbb::initAtInterval semantics), for any
    thread count; a budgeted run exports its unfinished intervals, whose
    oracle counts complete the total."""
    import pbbperm_amd as P
    from pbbperm_amd import factoradic as F
    O = oracle_mod
    for seed, n, m, mode in [(42, 12, 5, (0, 1, 0, 0)), (3, 11, 6, (2, 3, 1, 0)), (7, 11, 4, (2, 2, 0, 0))]:
        bm, br, rf, jp = mode
        ptm = O.synth_ptm(seed, n, m)
        opt = O.solve(ptm, O.Instance(ptm).neh()[0]).best
        pos, end = F.split(n, 6)
        want = sum(O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp,
                           pos=p, end=e).decomposed for p, e in zip(pos, end))
        for t in (6, 8):
            with P.CpuBB(ptm, _cfg(P, bm, br, rf, jp), threads=t, ub=opt) as cpu:
                cpu.init_from_fac(np.arange(6), pos, end)
                assert cpu.next()
                assert cpu.stats().decomposed == want, (seed, t)
        # budgeted run, export, finish each exported interval on the oracle
        with P.CpuBB(ptm, _cfg(P, bm, br, rf, jp), threads=4, ub=opt) as cpu:
            cpu.init_full_interval()
            done = cpu.next(node_budget=300)
            ids, p2, e2, dec, _ = cpu.get_intervals()
        if not done:
            assert len(ids) > 0
        total = dec + sum(O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp,
                                  pos=p, end=e).decomposed for p, e in zip(p2, e2))
        ref_total = O.solve(ptm, opt, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp).decomposed
        assert total >= ref_total  # re-unfolded paths are bounded uncounted: never fewer nodes


def test_cpu_coworker_improving_incumbent_and_interrupt():
    import threading
    import pbbperm_amd as P
    for seed in range(5, 9):
        ptm = P.synth(seed, 13, 5)
        nehc, nehp = P.neh(ptm)
        with P.CpuBB(ptm, _cfg(P, 0, 1, 0, 0), threads=4, ub=nehc, ub_perm=nehp) as cpu:
            assert cpu.solve()
            best, perm = cpu.best()
        assert best <= nehc and P.makespan(ptm, perm) == best
        with P.CpuBB(ptm, _cfg(P, 0, 1, 0, 0), threads=1, ub=nehc) as one:
            assert one.solve()
            assert one.best()[0] == best
    # a long search stops on interrupt and keeps its intervals
    big = P.synth(2, 50, 20)
    with P.CpuBB(big, _cfg(P, 0, 1, 0, 0), threads=4, ub=P.neh(big)[0]) as cpu:
        cpu.init_full_interval()
        t = threading.Timer(0.3, cpu.interrupt_exploration)
        t.start()
        assert not cpu.next(time_budget=30.0)
        t.join()
        st = cpu.stats()
        assert st.decomposed > 0 and st.active > 0
        assert not cpu.next(time_budget=0.2)
        assert cpu.stats().decomposed > st.decomposed


def test_cpu_coworker_rejects_bad_arguments():
    import pbbperm_amd as P
    ptm = P.synth(1, 8, 3)
    with pytest.raises(P.PbbError):
        P.CpuBB(ptm, ub=-1)
    with pytest.raises(P.PbbError):
        P.CpuBB(ptm, threads=-2)
    with P.CpuBB(ptm, threads=2) as cpu:
        with pytest.raises(P.PbbError):
            cpu.next()
        with pytest.raises(P.PbbError):
            cpu.init_from_fac([0, 1, 2], np.zeros((3, 8)), np.zeros((3, 8)))
        with pytest.raises(P.PbbError):
            cpu.init_from_fac([0], np.full((1, 8), 9), np.zeros((1, 8)))
