"""GPU parity of the multi-instance pool (batched solves and the serving
loop) with the CPU oracle and the reference's own fixtures.

Every instance in a pool is solved with the one-instance semantics
(initFullInterval + next until idle, identity root order), so with the
incumbent frozen at the optimum its decomposed / leaf / LB1 / refinement
counts must equal the oracle's (and the reference's, in the committed
fixtures) whatever else shares the launch; with an improving incumbent the
optimum must be exact and the returned schedule must realise it.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BUDGET = 30.0  # every launch ends by itself within this many seconds


def _cfg(P, bm, br, rf, jp):
    return P.default_config(bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp, early_stop_johnson=0)


def _counts(r):
    return [r.decomposed, r.leaves, r.simple_bounds, r.johnson_bounds]


def _check_schedule(P, ptm, r):
    if r.improved:
        perm = np.array(r.perm)
        assert sorted(perm.tolist()) == list(range(ptm.shape[1]))
        assert P.makespan(ptm, perm) == r.best


def _small_batch(O, count, seed0, nlo=8, nhi=12, mch=(3, 4, 5)):
    ptms = []
    for i in range(count):
        n, m = O.synth_shape(seed0 + i, nlo, nhi, mch)
        ptms.append(O.synth_ptm(seed0 + i, n, m))
    return ptms


def _opt(O, ptm):
    c, _ = O.Instance(ptm).neh()
    return O.solve(ptm, c, bound_mode=0, branching=1).best


MODES = [(0, 1, 0, 0), (0, 3, 0, 0), (0, 2, 0, 0), (2, 1, 1, 0), (2, 3, 1, 0), (2, 2, 1, 1), (2, 1, 0, 0),
         (1, 3, 1, 2)]


@pytest.mark.parametrize("mode", MODES)
def test_pool_counts_match_oracle_frozen_ub(oracle_mod, mode):
    import pbbperm_amd.pool as PP
    O = oracle_mod
    bm, br, rf, jp = mode
    ptms = _small_batch(O, 40, 7000)
    opts = [_opt(O, p) for p in ptms]
    want = []
    for p, u in zip(ptms, opts):
        r = O.solve(p, u, bound_mode=bm, branching=br, refine=rf, johnson_pairs=jp)
        want.append([r.decomposed, r.leaves, r.simple_bounds, r.johnson_bounds])
    import pbbperm_amd as P
    with PP.Pool(_cfg(P, bm, br, rf, jp), max_jobs=12, machines=5, slots=64, waves_per_group=2,
                 explorers_per_instance=32) as pool:
        res, _ = pool.solve(ptms, opts, time_budget=BUDGET)
    for i, r in enumerate(res):
        assert r.status == "solved", (i, r)
        assert r.best == opts[i]
        assert _counts(r) == want[i], (i, mode, _counts(r), want[i])


@pytest.mark.parametrize("machines", [5, 10, 20])
def test_pool_zero_time_padding_is_exact(oracle_mod, machines):
    """m = 3..5 instances in pools of 5, 10 and 20 machines give the same counts."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptms = _small_batch(O, 24, 7100)
    opts = [_opt(O, p) for p in ptms]
    want = [O.solve(p, u, bound_mode=2, branching=3, refine=1) for p, u in zip(ptms, opts)]
    with PP.Pool(_cfg(P, 2, 3, 1, 0), max_jobs=12, machines=machines, slots=32, waves_per_group=4,
                 explorers_per_instance=64) as pool:
        res, _ = pool.solve(ptms, opts, time_budget=BUDGET)
    for r, w, u in zip(res, want, opts):
        assert r.status == "solved" and r.best == u
        assert _counts(r) == [w.decomposed, w.leaves, w.simple_bounds, w.johnson_bounds]


def test_pool_improving_ub_finds_the_optimum(oracle_mod):
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptms = _small_batch(O, 48, 7200, 9, 13)
    opts = [_opt(O, p) for p in ptms]
    nehs = [O.Instance(p).neh()[0] for p in ptms]
    for W, epi in [(1, 1), (4, 16), (4, 256)]:
        with PP.Pool(_cfg(P, 2, 1, 1, 0), max_jobs=13, machines=5, slots=48, waves_per_group=W,
                     explorers_per_instance=epi) as pool:
            res, _ = pool.solve(ptms, nehs, time_budget=BUDGET)
        for p, r, u, h in zip(ptms, res, opts, nehs):
            assert r.status == "solved" and r.best == u
            assert r.improved == (u < h)
            _check_schedule(P, p, r)


def _batch_fixture():
    with open(os.path.join(HERE, "golden", "batch_c4.json")) as f:
        return json.load(f)


def test_batch_c4_512_instances_match_reference():
    """The 512-instance batch (n in [10,22], m in {5,10}, LB1 + Johnson LB2)
    in one graph launch: every optimum equals the reference's; with the
    bound frozen at the optimum every node count equals the reference's."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    fx = _batch_fixture()
    sp = fx["spec"]
    inst = fx["instances"]
    assert len(inst) == 512
    ptms = [P.synth(d["seed"], d["n"], d["m"]) for d in inst]
    cfg = _cfg(P, sp["bound_mode"], sp["branching"], sp["refine"], sp["pairs"])
    with PP.Pool(cfg, max_jobs=22, machines=10, slots=512, waves_per_group=4,
                 explorers_per_instance=256) as pool:
        res, ms = pool.solve(ptms, [d["neh"] for d in inst], time_budget=BUDGET)
        for p, r, d in zip(ptms, res, inst):
            assert r.status == "solved", d
            assert r.best == d["opt"], d
            _check_schedule(P, p, r)
        res, ms = pool.solve(ptms, [d["opt"] for d in inst], time_budget=BUDGET)
        for r, d in zip(res, inst):
            assert r.status == "solved" and r.best == d["opt"]
            assert _counts(r) == d["counts_frozen"], d
    print(f"batch of 512: {ms:.1f} ms")


def test_batch_histograms_match_the_oracle(oracle_mod):
    """pbb_pool_batch_histograms after a batch: with frozen incumbents, each
    instance's decomposed-node and refinement histograms by depth equal the
    oracle's (its depth histogram; refinements counted by the depth of the
    bounded node), and they sum to the instance's counters -- the figures
    the batched roofline is computed from."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptms = _small_batch(O, 24, 700, nlo=9, nhi=13)
    opts = [_opt(O, p) for p in ptms]
    cfg = _cfg(P, 2, 3, 1, 0)
    with PP.Pool(cfg, max_jobs=13, machines=5, slots=24, waves_per_group=2, explorers_per_instance=16) as pool:
        res, _ = pool.solve(ptms, opts, time_budget=BUDGET)
        hd, hr = pool.batch_histograms(len(ptms))
    for p, u, r, h, g in zip(ptms, opts, res, hd, hr):
        assert r.status == "solved"
        assert int(h.sum()) == r.decomposed and int(g.sum()) == r.johnson_bounds
        bb = O.BB(O.Instance(p), u, bound_mode=2, branching=3, refine=1)
        bb.init_full()
        bb.run()
        ref = bb.depth_histo()
        assert h[:len(ref)].tolist() == list(ref) and not h[len(ref):].any()


def test_repeated_graph_launches_are_identical(oracle_mod):
    """The captured graph replayed several times, with other GPU work and
    host synchronisation between the replays, gives the same results."""
    import torch
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptms = _small_batch(O, 32, 7300)
    opts = [_opt(O, p) for p in ptms]
    with PP.Pool(_cfg(P, 2, 3, 1, 0), max_jobs=12, machines=5, slots=32, waves_per_group=4,
                 explorers_per_instance=64) as pool:
        first, _ = pool.solve(ptms, opts, time_budget=BUDGET)
        for k in range(4):
            x = torch.arange(1 << 20, device="cuda", dtype=torch.int64).sum()
            torch.cuda.synchronize()
            assert int(x.item()) == (1 << 20) * ((1 << 20) - 1) // 2
            again, _ = pool.solve(ptms, opts, time_budget=BUDGET)
            assert [(_counts(a), a.best) for a in again] == [(_counts(f), f.best) for f in first]


def _serve_fixture():
    with open(os.path.join(HERE, "golden", "serve_small.json")) as f:
        return json.load(f)


def test_serving_results_alone_equal_results_in_the_mix():
    """A 256-request Poisson stream (n in [12,16], m in {5,10,20}) served by
    the resident loop with requests overlapping on the device: each
    request's optimum and frozen-bound node counts equal the reference's,
    equal the same request solved alone in a pool, and equal the plain
    one-instance-per-launch path."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    fx = _serve_fixture()
    sp = fx["spec"]
    inst = fx["instances"]
    ptms = [P.synth(d["seed"], d["n"], d["m"]) for d in inst]
    ubs = [d["opt"] for d in inst]
    cfg = _cfg(P, sp["bound_mode"], sp["branching"], sp["refine"], sp["pairs"])
    arrivals = PP.poisson_arrivals(len(ptms), rate=4000.0, seed=1)
    with PP.Pool(cfg, max_jobs=16, machines=20, slots=64, waves_per_group=4, explorers_per_instance=128) as pool:
        mixed, lat, wall = pool.replay(ptms, ubs, arrivals, time_budget=BUDGET)
        overlap = sum(1 for i in range(1, len(mixed)) if mixed[i].admit_ms < mixed[i - 1].done_ms)
        alone = []
        for i in range(0, len(ptms), 37):
            r, _ = pool.solve([ptms[i]], [ubs[i]], time_budget=BUDGET)
            alone.append((i, r[0]))
    for r, d in zip(mixed, inst):
        assert r.status == "solved", d
        assert r.best == d["opt"]
        assert _counts(r) == d["counts_frozen"], d
    for i, r in alone:
        assert _counts(r) == _counts(mixed[i])
    plain, plat, pwall = PP.plain_replay(ptms, ubs, arrivals, config=P.default_config(
        bound_mode=sp["bound_mode"], branching=sp["branching"], refine=sp["refine"],
        johnson_pairs=sp["pairs"], early_stop_johnson=0, nb_ivms=4096), time_budget=BUDGET)
    for r, d in zip(plain, inst):
        assert r.status == "solved" and r.best == d["opt"]
        assert _counts(r) == d["counts_frozen"], d
    assert np.all(lat > 0) and np.all(plat > 0)
    print(f"served 256 in {wall:.3f} s (p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms, "
          f"overlapping admissions {overlap}); plain {pwall:.3f} s")


def test_serving_wide_stream_matches_reference():
    """The serving stream's whole shape range (n in [12,30], m in
    {5,10,20}), stratified by machine count and job band (tests/golden/
    serve_wide.json: the first requests of the stream, seeds 10000+, of each
    stratum whose search the C oracle finishes within 3 M nodes; 40
    requests, none at m = 20 with n >= 21, where none of the 40 probed per
    band finished).  Served through one 20-machine pool of up to 30 jobs as
    a Poisson stream: from the NEH incumbent every optimum is the
    reference's and the schedule realises it; with the incumbent frozen at
    the optimum every node count is the reference's; the plain
    one-instance-per-launch path gives the same counts."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    with open(os.path.join(HERE, "golden", "serve_wide.json")) as f:
        fx = json.load(f)
    sp = fx["spec"]
    inst = fx["instances"]
    assert max(d["n"] for d in inst) >= 29 and {d["m"] for d in inst} == {5, 10, 20}
    ptms = [P.synth(d["seed"], d["n"], d["m"]) for d in inst]
    cfg = _cfg(P, sp["bound_mode"], sp["branching"], sp["refine"], sp["pairs"])
    arrivals = PP.poisson_arrivals(len(ptms), rate=200.0, seed=4)
    with PP.Pool(cfg, max_jobs=30, machines=20, slots=16, waves_per_group=4, explorers_per_instance=512) as pool:
        imp, _, _ = pool.replay(ptms, [d["neh"] for d in inst], arrivals, time_budget=120.0)
        frz, lat, wall = pool.replay(ptms, [d["opt"] for d in inst], arrivals, time_budget=120.0)
    for p, r, d in zip(ptms, imp, inst):
        assert r.status == "solved" and r.best == d["opt"], d
        _check_schedule(P, p, r)
    for r, d in zip(frz, inst):
        assert r.status == "solved" and _counts(r) == d["counts_frozen"], d
    plain, _, _ = PP.plain_replay(ptms, [d["opt"] for d in inst], arrivals, config=P.default_config(
        bound_mode=sp["bound_mode"], branching=sp["branching"], refine=sp["refine"],
        johnson_pairs=sp["pairs"], early_stop_johnson=0, nb_ivms=4096), time_budget=120.0)
    for r, d in zip(plain, inst):
        assert r.status == "solved" and _counts(r) == d["counts_frozen"], d
    print(f"wide stream: {len(inst)} requests, {sum(d['counts_frozen'][0] for d in inst)} frozen-bound nodes, "
          f"served in {wall:.2f} s (p99 {np.percentile(lat, 99):.1f} ms)")


def test_ig_helper_wave_keeps_every_optimum(oracle_mod):
    """Pools whose admitting groups run an iterated-greedy helper wave beside
    the search (ig_iterations > 0): the wide serving stream from the NEH
    incumbent gives the reference's optimum for every request, with a
    schedule that realises it; frozen at the optimum (nothing for the chain
    to improve) every node count is still the reference's; a batch of small
    instances, where the helper often outlives the search and records the
    result itself, gives the oracle's optima."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    with open(os.path.join(HERE, "golden", "serve_wide.json")) as f:
        fx = json.load(f)
    sp = fx["spec"]
    inst = fx["instances"]
    ptms = [P.synth(d["seed"], d["n"], d["m"]) for d in inst]
    cfg = _cfg(P, sp["bound_mode"], sp["branching"], sp["refine"], sp["pairs"])
    arrivals = PP.poisson_arrivals(len(ptms), rate=200.0, seed=4)
    with PP.Pool(cfg, max_jobs=30, machines=20, slots=16, waves_per_group=4, explorers_per_instance=512,
                 ig_iterations=30) as pool:
        imp, _, _ = pool.replay(ptms, [d["neh"] for d in inst], arrivals, time_budget=120.0)
        frz, _, _ = pool.replay(ptms, [d["opt"] for d in inst], arrivals, time_budget=120.0)
    for p, r, d in zip(ptms, imp, inst):
        assert r.status == "solved" and r.best == d["opt"], d
        _check_schedule(P, p, r)
    for r, d in zip(frz, inst):
        assert r.status == "solved" and _counts(r) == d["counts_frozen"], d
    small = _small_batch(O, 48, 8800, 8, 12)
    opts = [_opt(O, q) for q in small]
    with PP.Pool(_cfg(P, 2, 3, 1, 0), max_jobs=12, machines=5, slots=48, waves_per_group=2,
                 explorers_per_instance=32, ig_iterations=5) as pool:
        for ubs in ([PP.UB_NEH] * len(small), [P.neh(q)[0] for q in small]):
            res, _ = pool.solve(small, ubs, time_budget=BUDGET)
            for q, r, u in zip(small, res, opts):
                assert r.status == "solved" and r.best == u
                _check_schedule(P, q, r)


def test_serving_improving_ub_optimum_exact():
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    fx = _serve_fixture()
    sp = fx["spec"]
    inst = fx["instances"][:128]
    ptms = [P.synth(d["seed"], d["n"], d["m"]) for d in inst]
    cfg = _cfg(P, sp["bound_mode"], sp["branching"], sp["refine"], sp["pairs"])
    arrivals = PP.poisson_arrivals(len(ptms), rate=2000.0, seed=2)
    with PP.Pool(cfg, max_jobs=16, machines=20, slots=32, waves_per_group=2, explorers_per_instance=64) as pool:
        res, lat, wall = pool.replay(ptms, [d["neh"] for d in inst], arrivals, time_budget=BUDGET)
    for p, r, d in zip(ptms, res, inst):
        assert r.status == "solved" and r.best == d["opt"]
        _check_schedule(P, p, r)


def test_routed_serving_matches_fixture():
    """The 256-request stream through a 5-, a 10- and a 20-machine pool
    resident together, each request routed to the smallest pool that holds
    it: every request is solved with the reference's frozen-bound counts,
    however the stream is split.  A request that fits no pool, a repeated
    pool or an empty pool list is refused before anything runs."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    fx = _serve_fixture()
    sp = fx["spec"]
    inst = fx["instances"]
    ptms = [P.synth(d["seed"], d["n"], d["m"]) for d in inst]
    ubs = [d["opt"] for d in inst]
    assert {p.shape[0] for p in ptms} == {5, 10, 20}
    cfg = _cfg(P, sp["bound_mode"], sp["branching"], sp["refine"], sp["pairs"])
    arrivals = PP.poisson_arrivals(len(ptms), rate=4000.0, seed=3)
    kw = dict(max_jobs=16, slots=32, waves_per_group=4, explorers_per_instance=128)
    with PP.Pool(cfg, machines=5, groups=64, **kw) as p5, PP.Pool(cfg, machines=20, groups=64, **kw) as p20, \
            PP.Pool(cfg, machines=10, groups=64, **kw) as p10:
        res, lat, wall = PP.replay_routed([p20, p5, p10], ptms, ubs, arrivals, time_budget=BUDGET)
        with pytest.raises(P.PbbError):
            PP.replay_routed([p5, p10], ptms, ubs, arrivals, time_budget=BUDGET)   # m = 20 fits neither
        with pytest.raises(P.PbbError):
            PP.replay_routed([p20, p20], ptms, ubs, arrivals, time_budget=BUDGET)
        with pytest.raises(P.PbbError):
            PP.replay_routed([], ptms, ubs, arrivals, time_budget=BUDGET)
        # the pools serve again after the refusals
        again, _, _ = PP.replay_routed([p10, p20], ptms[:32], ubs[:32], arrivals[:32], time_budget=BUDGET)
    for r, d in zip(res, inst):
        assert r.status == "solved", d
        assert r.best == d["opt"]
        assert _counts(r) == d["counts_frozen"], d
    for r, d in zip(again, inst[:32]):
        assert r.status == "solved" and _counts(r) == d["counts_frozen"], d
    assert np.all(lat > 0)
    print(f"routed 256 in {wall:.3f} s (p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms)")


def test_routed_serving_with_a_partly_resident_pool_ends_promptly():
    """A 10-machine pool sized for the whole device (groups=0), listed
    before a 20-machine pool: the 20-machine pool's kernel starts first, so
    some 10-machine groups wait for a CU for the whole replay.  Every
    request is still solved, and stopping closes both pools before waiting
    for either, so the call returns without waiting for the kernels' time
    budget (BUDGET + 5 s)."""
    import time
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    fx = _serve_fixture()
    sp = fx["spec"]
    inst = fx["instances"][:48]
    ptms = [P.synth(d["seed"], d["n"], d["m"]) for d in inst]
    cfg = _cfg(P, sp["bound_mode"], sp["branching"], sp["refine"], sp["pairs"])
    arrivals = PP.poisson_arrivals(len(ptms), rate=2000.0, seed=4)
    kw = dict(max_jobs=16, slots=32, waves_per_group=4, explorers_per_instance=128)
    with PP.Pool(cfg, machines=10, groups=0, **kw) as p10, PP.Pool(cfg, machines=20, groups=64, **kw) as p20:
        assert p10.info()["groups"] > 64
        t0 = time.time()
        res, lat, wall = PP.replay_routed([p10, p20], ptms, [d["opt"] for d in inst], arrivals,
                                          time_budget=BUDGET)
        took = time.time() - t0
    for r, d in zip(res, inst):
        assert r.status == "solved" and _counts(r) == d["counts_frozen"], d
    assert took < BUDGET / 2, took


def test_serve_admission_and_poll_api(oracle_mod):
    """Submit/poll by hand: a full pool refuses requests (PBB_E_BUSY) until
    slots free up; the footprint follows explorers x (n^2 + 2n) int32."""
    import time
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptms = _small_batch(O, 6, 7400)
    opts = [_opt(O, p) for p in ptms]
    with PP.Pool(_cfg(P, 0, 1, 0, 0), max_jobs=12, machines=5, slots=2, waves_per_group=1,
                 explorers_per_instance=8) as pool:
        n = ptms[0].shape[1]
        fp = pool.footprint(n)
        assert fp >= 8 * (n * n + 2 * n) * 4
        pool.serve_start(BUDGET)
        got = {}
        pending = list(range(len(ptms)))
        refused = 0
        t0 = time.time()
        while len(got) < len(ptms) and time.time() - t0 < BUDGET:
            while pending:
                if not pool.submit(ptms[pending[0]], opts[pending[0]], tag=pending[0]):
                    refused += 1
                    break
                pending.pop(0)
            for r in pool.poll():
                got[r.tag] = r
        pool.serve_stop()
    assert len(got) == len(ptms)
    for i, r in got.items():
        assert r.status == 1 and r.best == opts[i]
    assert refused >= 0


def test_pool_rejects_bad_requests():
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    with PP.Pool(_cfg(P, 0, 1, 0, 0), max_jobs=10, machines=5, slots=4, waves_per_group=1,
                 explorers_per_instance=4) as pool:
        with pytest.raises(P.PbbError):
            pool.solve([P.synth(1, 12, 5)], [10**6], time_budget=BUDGET)   # n above max_jobs
        with pytest.raises(P.PbbError):
            pool.solve([P.synth(1, 8, 7)], [10**6], time_budget=BUDGET)    # m above machines
        bad = P.synth(1, 8, 5).copy()
        bad[0, 0] = -3
        with pytest.raises(P.PbbError):
            pool.solve([bad], [10**6], time_budget=BUDGET)
    with pytest.raises(P.PbbError):
        PP.Pool(_cfg(P, 0, 1, 0, 0), max_jobs=10, machines=5, slots=4, waves_per_group=3)


def test_autotuner_picks_a_valid_configuration(oracle_mod):
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptms = _small_batch(O, 24, 7500)
    opts = [_opt(O, p) for p in ptms]
    base = PP.default_params(max_jobs=12, machines=5, slots=24)
    best, trials = PP.autotune(ptms, opts, config=_cfg(P, 2, 3, 1, 0), base=base, time_budget=BUDGET)
    nbest, ntrials = PP.autotune_native(ptms, opts, config=_cfg(P, 2, 3, 1, 0), base=base)
    assert len(ntrials) == len(trials)
    assert [tuple(r[:3]) for r in ntrials] == [tuple(r[:3]) for r in trials]
    assert best.waves_per_group in (1, 2, 4)
    assert best.explorers_per_instance >= best.waves_per_group
    assert len(trials) >= 9
    assert np.all(np.isfinite(trials[:, 3])) and np.all(trials[:, 3] < 1e100)
    with PP.Pool(_cfg(P, 2, 3, 1, 0), params=best) as pool:
        res, _ = pool.solve(ptms, opts, time_budget=BUDGET)
    assert all(r.status == "solved" and r.best == u for r, u in zip(res, opts))
    # the latency objective replays the instances as a request stream and
    # searches the same grid
    arrivals = PP.poisson_arrivals(len(ptms), 400.0, seed=3)
    rbest, rtrials = PP.autotune(ptms, opts, config=_cfg(P, 2, 3, 1, 0), base=base, time_budget=BUDGET,
                                 objective="replay_mean", arrivals=arrivals)
    assert [tuple(r[:3]) for r in rtrials] == [tuple(r[:3]) for r in trials]
    assert np.all(np.isfinite(rtrials[:, 3])) and np.all(rtrials[:, 3] > 0)
    assert (rbest.waves_per_group, rbest.explorers_per_instance, rbest.steal_margin) in \
        [tuple(int(x) for x in r[:3]) for r in rtrials]
    with pytest.raises(ValueError):
        PP.autotune(ptms, opts, objective="p99")
    with pytest.raises(ValueError):
        PP.autotune(ptms, opts, objective="replay_mean")


def test_device_neh_matches_host_neh():
    """NEH on the device (one wavefront per instance) gives the host NEH's
    order and makespan exactly, from 3 to 63 jobs and 2 to 20 machines."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    shapes = [(3, 2), (5, 3), (12, 5), (20, 5), (20, 10), (30, 20), (50, 20), (63, 20), (63, 7), (17, 13)]
    ptms = [P.synth(900 + i, n, m) for i, (n, m) in enumerate(shapes)]
    ptms += PP.serving_stream(120, 40000)
    orders, costs = PP.neh_device(ptms)
    for p, o, c in zip(ptms, orders, costs):
        hc, hp = P.neh(p)
        assert int(c) == hc
        assert list(o) == list(hp)
        assert P.makespan(p, o) == hc


def test_pool_find_all_matches_reference():
    """findAll in the pool (all instances of the fixture in one launch, each
    with its own mode's pool): counts equal the reference's."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    with open(os.path.join(HERE, "golden", "findall_cases.json")) as f:
        cases = json.load(f)
    modes = sorted({(c["bound_mode"], c["branching"], c["refine"], c["johnson_pairs"]) for c in cases})
    for mode in modes:
        sel = [c for c in cases if (c["bound_mode"], c["branching"], c["refine"], c["johnson_pairs"]) == mode]
        cfg = _cfg(P, *mode)
        cfg.find_all = 1
        ptms = [P.synth(c["seed"], c["n"], c["m"]) for c in sel]
        with PP.Pool(cfg, max_jobs=12, machines=5, slots=len(sel), waves_per_group=2,
                     explorers_per_instance=32) as pool:
            res, _ = pool.solve(ptms, [c["ub"] for c in sel], time_budget=BUDGET)
        for c, r in zip(sel, res):
            assert r.status == "solved" and r.best == c["best"]
            assert _counts(r) == [c["decomposed"], c["leaves"], c["simple_bounds"], c["johnson_bounds"]], (c, r)


def test_pool_tags_are_returned_and_never_decide_retirement(oracle_mod):
    """Caller tags come back as given, -1 included; whether an instance
    retired is decided by the request's index, so with a deadline that ends
    the launch first the unfinished instances read ABORTED whatever their
    tag (a tag of -1 once equalled the reset value of an unwritten result)."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptms = _small_batch(O, 6, 7800)
    opts = [_opt(O, p) for p in ptms]
    tags = [-1, 5, -1, 7, 123456, -1]
    with PP.Pool(_cfg(P, 0, 1, 0, 0), max_jobs=12, machines=5, slots=8, waves_per_group=2,
                 explorers_per_instance=32) as pool:
        res, _ = pool.solve(ptms, opts, tags=tags, time_budget=BUDGET)
    assert [r.tag for r in res] == tags
    assert all(r.status == "solved" and r.best == u for r, u in zip(res, opts))
    # a 30x20 instance from infinity cannot finish in 0.2 s: it must read aborted
    big = [P.synth(31, 30, 20)] + ptms[:2]
    with PP.Pool(_cfg(P, 0, 1, 0, 0), max_jobs=30, machines=20, slots=4, waves_per_group=2,
                 explorers_per_instance=64) as pool:
        res, _ = pool.solve(big, [2**31 - 1, opts[0], opts[1]], tags=[-1, -1, -1], time_budget=0.2)
    assert res[0].status == "aborted" and res[0].tag == -1
    assert res[1].status == "solved" and res[2].status == "solved"


def test_request_budget_retires_long_requests_and_keeps_the_rest_exact(oracle_mod):
    """A per-request time budget: requests whose trees cannot finish within
    it (30x20 from an infinite incumbent) retire ABORTED shortly after it,
    with a schedule of the reported cost when one was found, while the small
    requests around them keep exact optima and frozen-bound counts; the slot
    of an aborted request serves later requests.  Serving and batch mode."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    small = _small_batch(O, 24, 7900)
    opts = [_opt(O, p) for p in small]
    want = [O.solve(p, u, bound_mode=0, branching=1) for p, u in zip(small, opts)]
    big = [P.synth(41 + i, 30, 20) for i in range(3)]
    ptms, ubs, kinds = [], [], []
    for i, (p, u) in enumerate(zip(small, opts)):
        ptms.append(p); ubs.append(u); kinds.append(i)
        if i % 8 == 3:
            ptms.append(big[i // 8]); ubs.append(2**31 - 1); kinds.append(-1)
    budget = 0.25
    with PP.Pool(_cfg(P, 0, 1, 0, 0), max_jobs=30, machines=20, slots=8, waves_per_group=2,
                 explorers_per_instance=64, request_budget_s=budget) as pool:
        arrivals = np.linspace(0.0, 0.5, len(ptms))
        res, lat, wall = pool.replay(ptms, ubs, arrivals, time_budget=BUDGET)
        for r, l, k, p in zip(res, lat, kinds, ptms):
            if k < 0:
                assert r.status == "aborted", r
                assert budget * 1e3 <= l < budget * 1e3 + 2000, l
                _check_schedule(P, p, r)
            else:
                w = want[k]
                assert r.status == "solved" and r.best == opts[k]
                assert _counts(r) == [w.decomposed, w.leaves, w.simple_bounds, w.johnson_bounds]
        res, _ = pool.solve(ptms[:5], ubs[:5], time_budget=BUDGET)
        assert [r.status for r in res] == ["solved"] * 4 + ["aborted"]
        assert [r.best for r in res[:4]] == opts[:4]
    plain, plat, _ = PP.plain_replay(ptms[:5], ubs[:5], np.zeros(5), config=_cfg(P, 0, 1, 0, 0),
                                     time_budget=BUDGET, request_budget=budget)
    assert [r.status for r in plain] == ["solved"] * 4 + ["aborted"]
    assert [r.best for r in plain[:4]] == opts[:4]


def test_device_neh_matches_reference_fixture(oracle_mod):
    """pbb_neh_device gives the reference fastNEH's order and makespan on the
    committed fixture (20x5, the 512 batched instances, the serving stream,
    n = 3..63 and instances built to tie)."""
    import pbbperm_amd.pool as PP
    with open(os.path.join(HERE, "golden", "neh_cases.json")) as f:
        cases = json.load(f)
    ptms = []
    for c in cases:
        p = np.asarray(c["ptm"], dtype=np.int32) if "ptm" in c else oracle_mod.synth_ptm(c["seed"], c["n"], c["m"])
        ptms.append(np.ascontiguousarray(p))
    orders, costs = PP.neh_device(ptms)
    for c, o, cost in zip(cases, orders, costs):
        assert (int(cost), list(o)) == (c["cost"], c["perm"]), (c["kind"], c["seed"])


def test_pool_computes_neh_at_admission(oracle_mod):
    """Requests with ub = UB_NEH: the pool's NEH incumbent then the search;
    the optimum is exact and the returned schedule realises it."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    O = oracle_mod
    ptms = _small_batch(O, 40, 7600, 9, 13)
    opts = [_opt(O, p) for p in ptms]
    with PP.Pool(_cfg(P, 2, 3, 1, 0), max_jobs=13, machines=5, slots=40, waves_per_group=4,
                 explorers_per_instance=64) as pool:
        res, _ = pool.solve(ptms, [PP.UB_NEH] * len(ptms), time_budget=BUDGET)
    for p, r, u in zip(ptms, res, opts):
        assert r.status == "solved" and r.best == u
        assert r.improved
        _check_schedule(P, p, r)
    # the plain path takes the same stream: a UB_NEH request gets the host
    # NEH's incumbent (never a cost of -1 that would prune every child)
    plain, _, _ = PP.plain_replay(ptms[:8], [PP.UB_NEH] * 8, np.zeros(8), config=_cfg(P, 2, 3, 1, 0),
                                  time_budget=BUDGET)
    for p, r, u in zip(ptms[:8], plain, opts[:8]):
        assert r.status == "solved" and r.best == u and r.improved
        _check_schedule(P, p, r)


def test_single_request_pool_recruits_every_group():
    """A batch of one request: while its owner group admits it the other
    groups find the queue closed and drained, and must wait for the slot to
    run rather than leave the launch (before they did, and the 19x10 tree
    below took 5.5 s on the owner's 4 explorers).  The counts with the
    incumbent frozen at the optimum equal the plain engine's, and the solve
    takes a small fraction of that time."""
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    p = PP.batch_stream(512, 100)[311]
    assert p.shape == (10, 19)
    opt = 1434
    cfg = PP.workload_config("batched")
    with P.GpuBB(p, PP.workload_config("batched", nb_ivms=1024), ub=opt) as eng:
        eng.solve(time_budget=BUDGET)
        want = [eng.stats().decomposed, eng.stats().leaves]
    with PP.Pool(cfg, max_jobs=22, machines=10, slots=1, waves_per_group=4, explorers_per_instance=1024,
                 steal_margin=2) as pool:
        pool.solve([p], [opt], time_budget=BUDGET)
        r, ms = pool.solve([p], [opt], time_budget=BUDGET)
    print(f"19x10 frozen at {opt}: {r[0].decomposed} nodes in {ms:.1f} ms (plain {want})", flush=True)
    assert r[0].status == "solved" and r[0].best == opt
    assert [r[0].decomposed, r[0].leaves] == want
    assert ms < 1500.0
