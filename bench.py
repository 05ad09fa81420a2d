"""Benchmark of the IVM branch-and-bound engine on one MI355X.  Prints one
JSON line.

Headline (the line's top-level fields): decomposed nodes/s on the 50 jobs x
20 machines instance (times uniform in [1, 99], seed 2), LB1 bounding,
MaxSum branching, incumbent from NEH.  A step is one resident-engine launch
(pbb_run) exploring a fixed budget of decomposed nodes; the tree is far too
large to finish, so the engine keeps its IVMs, intervals and incumbent in
HBM across steps and each step continues the exploration.

Three more legs, reported as objects of the same line:
  "small_configs"  the 20x5 (LB1, NEH incumbent, 4096 explorers) and 20x10
             (LB1 + Johnson LB2, incumbent = optimum, 16384 explorers)
             parity configurations timed end to end, checked against the
             reference's fixture, with the reference's single-thread time;
  "batched"  the 512-instance batch (seeds 100-611, n ~ U[10,22],
             m ~ U{5,10}, LB1 + Johnson LB2 over all machine pairs, MinMin
             branching, NEH incumbents) solved to optimality in one
             graph-captured launch of the pool with autotuned parameters,
             against the plain one-instance-per-launch engine on the same
             instances, with the reference's own code on the host cores
             beside it;
  "serving"  a Poisson stream of requests (seeds 10000+, n ~ U[12,30],
             m ~ U{5,10,20}; --serve-requests at --serve-rate per second,
             each stopped --serve-budget seconds after its admission and then
             counted late) through the resident serving loop, through a
             10-machine and a 20-machine pool resident together with each
             request routed to the smaller one that holds it ("routed"), and
             through the plain path, same stream, same arrival times, same
             per-request budget: solves/s, p50/p99 over solved requests, and
             p99 with the late ones counted (with PBB_REPLAY_PROGRESS=1 in the
             environment the replays report every 500 served requests on
             stderr).
Two legs run only when named: "treesizes" (the serving stream's tree-size
histogram) and "long" (the headline workload for --long-seconds, step by
step).
Pool parameters come from the autotuner's stored choice
(pbbperm_amd/tuned.json, python -m pbbperm_amd.tune), or --autotune.
Every optimum of both legs is compared between the pool and the plain path
(for the serving stream: every request both solved).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--node-budget B]
                  [--legs nodes,small,batched,serving,treesizes,long]
                  [--serve-requests N] [--serve-rate R] [--serve-budget S]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 int32 lanes/clk x 2.4 GHz
PEAK_HBM_GBS = 8000.0
# HBM bytes per decomposed node of k_explore on this workload, from two
# rocprofv3 --pmc passes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a
# plain 256 MiB device copy): profiles/k_explore_50x20_lb1_pmc.md
HBM_BYTES_PER_NODE = 0.80
# HBM bytes of one k_pool<10,2> launch of the 512-instance batch, from the
# same two counters and factors: profiles/k_pool_batch512_lb2_pmc.md
BATCH_HBM_BYTES = 0.93e9
# The serving stream of the default run: requests, Poisson rate and the
# per-request time budget after admission.  Every path's replay ends at the
# arrival span + 2 budgets + 15 s; what is not served by then counts late.
SERVE_REQUESTS = 600
SERVE_RATE = 20.0
SERVE_BUDGET_S = 2.0


def lb1_ops(depth: int, n: int, m: int) -> int:
    """Algorithmic int32 ops of one LB1 node at depth D (D fixed jobs,
    k = n - D children): front+back recompute (2m-1)D, remain m*k, the two
    child chains 2k(5m-3) (bound_fsp_weak.cpp:61-221)."""
    k = n - depth
    return (2 * m - 1) * depth + m * k + 2 * k * (5 * m - 3)


def lb2_ops(depth: int, n: int, m: int, pairs: int) -> int:
    """Algorithmic int32 ops of one Johnson LB2 refinement of a child of a
    node at depth D (the child has D+1 fixed jobs, k' = n - D - 1 unscheduled):
    per machine pair a two-machine pass over the k' jobs plus the front/back
    offsets, P(4k' + 4), and the child's front and back, (2m-1)(D+1)
    (bound_fsp_strong.cpp:50-162, :441-560)."""
    kp = n - depth - 1
    return pairs * (4 * kp + 4) + (2 * m - 1) * (depth + 1)


def pool_ops(ptms, dec_histo, ref_histo, pairs_of_m) -> float:
    """Algorithmic ops of a pool batch from its per-instance depth
    histograms: interior nodes (depths 1..n-2) x W1 plus refinements x W2."""
    ops = 0.0
    for p, hd, hr in zip(ptms, dec_histo, ref_histo):
        m, n = p.shape
        P = pairs_of_m(m)
        for d in range(1, n - 1):
            ops += float(hd[d]) * lb1_ops(d, n, m) + float(hr[d]) * lb2_ops(d, n, m, P)
    return ops


def host_cpus():
    """What the CPU baseline runs on: the lscpu model, sockets, physical
    cores, hardware threads and SMT state of the host, and the CPUs this
    process may use (its affinity set, at most 16: the host's share per GPU)
    with how many distinct physical cores they are."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            kv[k.strip()] = v.strip()
        tpc = int(kv.get("Thread(s) per core", "1") or 1)
        cps = int(kv.get("Core(s) per socket", "0") or 0)
        sk = int(kv.get("Socket(s)", "1") or 1)
        info.update(model=kv.get("Model name"), sockets=sk, physical_cores=cps * sk,
                    hardware_threads=int(kv.get("CPU(s)", "0") or 0), threads_per_core=tpc)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            info["smt_active"] = f.read().strip() == "1"
    except OSError:
        info["smt_active"] = info.get("threads_per_core", 1) > 1
    cpus = sorted(os.sched_getaffinity(0))[:16]
    cores = set()
    for c in cpus:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/thread_siblings_list") as f:
                cores.add(f.read().strip())
        except OSError:
            cores.add(str(c))
    info["cpus_used"] = len(cpus)
    info["physical_cores_used"] = len(cores)
    info["smt_used"] = len(cores) < len(cpus)
    return info, len(cpus)


def cpu_baseline(ptm, ub, seconds: float):
    """The reference's own bound / IVM code (oracle/_ref/ref_harness, one
    single-threaded process per core) on the same instance and incumbent for
    a bounded time: the search space [0, n!-1] is cut into 64 equal closed
    factoradic intervals per core (as the reference's master cuts its work,
    distributed/src/works.cpp:96-128, and hands intervals to
    matrix_controller::initFromFac, multicore/src/matrix_controller.cpp:112-140);
    a pool of one worker per core takes them in order, each explored from
    its start until done or the deadline.  nodes/s = decomposed nodes of all
    workers / wall.  Also a single-thread run from the root (the parity
    configuration).  Falls back to the C oracle port where the reference
    build is absent."""
    import concurrent.futures as cf
    import oracle as O
    from pbbperm_amd import factoradic as F
    O.lib()
    n = ptm.shape[1]
    ref = O.have_ref()
    kind = "reference" if ref else "port"
    host, cores = host_cpus()
    # one thread, from the root
    if ref:
        r1 = O.ref_solve(ptm, ub, bound_mode=0, branching=1, time_limit=seconds / 2)
    else:
        t = time.time()
        r1 = O.solve(ptm, ub, bound_mode=0, branching=1, max_nodes=int(seconds / 2 * 2e5))
        r1.wall_s = time.time() - t
    single = {"value": r1.decomposed / r1.wall_s, "unit": "nodes/s", "cores": 1,
              "sample": f"DFS from the root for {r1.wall_s:.1f} s ({r1.decomposed} decomposed nodes)"}
    pos, end = F.split(n, 64 * cores)
    t0 = time.time()
    deadline = t0 + seconds

    def run(k):
        left = deadline - time.time()
        if left <= 0.05:
            return 0
        if ref:
            r = O.ref_solve(ptm, ub, bound_mode=0, branching=1, time_limit=left, pos=pos[k], end=end[k])
        else:
            r = O.solve(ptm, ub, bound_mode=0, branching=1, max_nodes=int(left * 2e5), pos=pos[k], end=end[k])
        return r.decomposed

    with cf.ThreadPoolExecutor(cores) as ex:
        counts = list(ex.map(run, range(len(pos))))
    wall = time.time() - t0
    started = sum(1 for c in counts if c > 0)
    return {"value": sum(counts) / wall, "unit": "nodes/s", "cores": cores, "kind": kind,
            "sample": f"{cores} single-thread processes over {started} of {len(pos)} equal factoradic intervals "
                      f"of the same instance and incumbent for {wall:.1f} s ({sum(counts)} decomposed nodes)",
            "host": host, "single_thread": single}


def note(msg: str):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _latency(lat, wall, res=None):
    """Throughput and latency figures of a replay.  A request is solved when
    its result says so; the others are "late": still running when the
    replay's time budget ended (latency -1), or stopped by the per-request
    budget (status aborted, latency = when that result came back).  p50 /
    p99 / mean are over the solved requests; p99_all also counts every late
    request, at its latency or, when it never came back, at the replay's
    end, so it is a lower bound of the stream's true p99."""
    import numpy as np
    a = np.asarray(lat, dtype=float)
    ok = a >= 0
    if res is not None:
        ok = ok & np.array([r.status == "solved" for r in res], dtype=bool)
    done = a[ok]
    has = len(done) > 0
    allv = np.where(a >= 0, a, max(float(a.max()) if len(a) else 0.0, wall * 1e3))
    return {"solves_per_s": len(done) / wall if wall > 0 else 0.0,
            "p50_ms": float(np.percentile(done, 50)) if has else None,
            "p99_ms": float(np.percentile(done, 99)) if has else None,
            "p99_all_ms": float(np.percentile(allv, 99)) if len(allv) else None,
            "mean_ms": float(np.mean(done)) if has else None,
            "max_ms": float(np.max(done)) if has else None,
            "late": int(len(a) - len(done))}


def cpu_solves(ptms, ubs, kw, seconds: float):
    """The reference's own code (oracle/_ref, one process per instance, as
    its process-global bound tables require) over the host cores on a
    bounded sample: instances in order until the time is up; solves/s =
    instances finished / wall."""
    import concurrent.futures as cf
    import oracle as O
    O.lib()
    kind = "reference" if O.have_ref() else "port"
    cores = max(1, min(16, os.cpu_count() or 1))
    t0 = time.time()
    done = 0

    def run(i):
        left = seconds - (time.time() - t0)
        if left <= 0.05:
            return False
        if kind == "reference":
            r = O.ref_solve(ptms[i], ubs[i], time_limit=left, **kw)
        else:
            r = O.solve(ptms[i], ubs[i], **kw)
        return bool(r.finished) and time.time() - t0 <= seconds * 1.05

    with cf.ThreadPoolExecutor(cores) as ex:
        for ok in ex.map(run, range(len(ptms))):
            done += int(ok)
    wall = time.time() - t0
    return {"value": done / wall, "unit": "solves/s", "cores": cores, "kind": kind,
            "sample": f"{done} of the same instances in order, one process each, {wall:.1f} s"}


def _params(PP, name, args):
    """Pool parameters of a workload: the autotuner's stored choice
    (pbbperm_amd/tuned.json, written by python -m pbbperm_amd.tune), or a
    fresh autotuning run with --autotune or when none is stored."""
    best = None if args.autotune else PP.tuned_params(name)
    src = "stored autotune (pbbperm_amd/tuned.json)"
    if best is None:
        note(f"{name}: autotuning")
        best, _ = PP.tune_workload(name, progress=lambda r: note(f"  trial {r}"))
        src = "autotuned in this run"
    return best, src


def leg_small(P, args):
    """The two parity configurations timed end to end (one solve each):
    20x5 seed 0, LB1, MaxSum, NEH incumbent, 4096 explorers; and 20x10
    seed 1, LB1 + Johnson LB2, MaxSum, incumbent = the optimum, 16384
    explorers.  Optimum and node counts against the reference's fixture
    (counts are order-free only with the incumbent frozen at the optimum)."""
    fx = os.path.join(ROOT, "tests", "golden", "solve_cases.json")
    cases = json.load(open(fx)) if os.path.exists(fx) else []

    def ref(seed, m, bm, br, rf, kind):
        for c in cases:
            if (c["seed"], c["m"], c["bound_mode"], c["branching"], c["refine"], c["ub_kind"]) == \
                    (seed, m, bm, br, rf, kind):
                return c
        return None

    out = {}
    for name, seed, n, m, bm, rf, nb, kind in [("20x5_lb1_neh", 0, 20, 5, 0, 0, 4096, "neh"),
                                               ("20x10_lb1_lb2_opt", 1, 20, 10, 2, 1, 16384, "opt")]:
        ptm = P.synth(seed, n, m)
        r = ref(seed, m, bm, 1, rf, kind)
        ub = r["ub"] if r else P.neh(ptm)[0]
        cfg = P.default_config(bound_mode=bm, branching=1, refine=rf, johnson_pairs=0, early_stop_johnson=0,
                               nb_ivms=nb)
        # solve_ms: pbb_init_full + pbb_run on an engine created beforehand
        # (the reference's time is its search alone); e2e_ms adds pbb_create
        # (device buffers, tables, stream) and pbb_destroy
        times, e2e, decs = [], [], []
        for _ in range(5):
            t0 = time.perf_counter()
            with P.GpuBB(ptm, cfg, ub=ub) as eng:
                t1 = time.perf_counter()
                eng.solve(time_budget=60.0)
                t2 = time.perf_counter()
                st = eng.stats()
                best, _ = eng.best()
            e2e.append(time.perf_counter() - t0)
            times.append(t2 - t1)
            decs.append(st.decomposed)
        d = {"solve_ms": min(times) * 1e3, "solve_ms_median": sorted(times)[len(times) // 2] * 1e3,
             "e2e_ms": min(e2e) * 1e3, "kernel_ms": st.kernel_ms, "explorers": st.explorers,
             "ramp_nodes": cfg.ramp_nodes, "best": best, "decomposed": st.decomposed,
             "decomposed_runs": decs}
        if r:
            d["optimum_equals_reference"] = best == r["best"]
            if kind == "opt":
                d["counts_equal_reference"] = [st.decomposed, st.leaves, st.simple_bounds, st.johnson_bounds] == \
                    [r["decomposed"], r["leaves"], r["simple_bounds"], r["johnson_bounds"]]
        if not args.no_cpu:
            import oracle as O
            O.lib()
            if O.have_ref():
                rr = O.ref_solve(ptm, ub, bound_mode=bm, branching=1, refine=rf, johnson_pairs=0)
                d["cpu_reference_ms"] = rr.wall_s * 1e3
                d["cpu_reference_decomposed"] = rr.decomposed
        out[name] = d
    return out


def leg_batched(P, PP, args):
    import numpy as np
    cfg = PP.workload_config("batched")
    ptms = PP.batch_stream(512, 100)
    ubs = [P.neh(p)[0] for p in ptms]
    best, src = _params(PP, "batched", args)
    note(f"batched: 512 instances, W={best.waves_per_group} epi={best.explorers_per_instance} "
         f"margin={best.steal_margin}")
    with PP.Pool(cfg, params=best) as pool:
        res, _ = pool.solve(ptms, ubs, time_budget=120.0)
        walls, kms, opss = [], [], []
        for _ in range(3):
            t = time.perf_counter()
            res, km = pool.solve(ptms, ubs, time_budget=120.0)
            walls.append(time.perf_counter() - t)
            kms.append(km)
            hd, hr = pool.batch_histograms(len(ptms))
            opss.append(pool_ops(ptms, hd, hr, lambda m: m * (m - 1) // 2))
        info = pool.info()
    # the same pool without the IG helper wave: the plain path below has only
    # the NEH incumbents, so this is the like-for-like comparison
    noig = None
    if best.ig_iterations > 0:
        import ctypes
        prm0 = PP.PoolParams()
        ctypes.memmove(ctypes.byref(prm0), ctypes.byref(best), ctypes.sizeof(PP.PoolParams))
        prm0.ig_iterations = 0
        with PP.Pool(cfg, params=prm0) as pool0:
            pool0.solve(ptms, ubs, time_budget=120.0)
            w0, r0 = [], None
            for _ in range(3):
                t = time.perf_counter()
                r0, _ = pool0.solve(ptms, ubs, time_budget=120.0)
                w0.append(time.perf_counter() - t)
        noig = {"wall_ms": min(w0) * 1e3, "solves_per_s": len(ptms) / min(w0),
                "decomposed": int(sum(r.decomposed for r in r0)),
                "optima_equal": all(a.best == b.best for a, b in zip(res, r0))}
    note(f"batched: pool {min(walls) * 1e3:.1f} ms; plain path")
    # the plain path gets the same per-instance explorer count
    pcfg = PP.workload_config("batched", nb_ivms=best.explorers_per_instance)
    plain, _, pwall = PP.plain_replay(ptms, ubs, np.zeros(len(ptms)), config=pcfg, time_budget=300.0)
    note(f"batched: plain {pwall * 1e3:.1f} ms")
    ok = all(r.status == "solved" for r in res) and all(r.status == "solved" for r in plain)
    same = all(a.best == b.best for a, b in zip(res, plain))
    ref_match = None
    fx = os.path.join(ROOT, "tests", "golden", "batch_c4.json")
    if os.path.exists(fx):
        with open(fx) as f:
            opts = [d["opt"] for d in json.load(f)["instances"]]
        ref_match = [r.best for r in res] == opts
    wall = min(walls)
    nodes = sum(r.decomposed for r in res)
    # how the batch ends: when instances retire (device clock, ms after the
    # earliest admission) -- the last few trees set the batch time
    done = np.sort(np.array([r.done_ms for r in res]))
    tail = {"retired_p50_ms": float(np.percentile(done, 50)), "retired_p90_ms": float(np.percentile(done, 90)),
            "retired_p99_ms": float(np.percentile(done, 99)), "retired_last_ms": float(done[-1]),
            "largest_trees": sorted((int(r.decomposed) for r in res), reverse=True)[:3]}
    # the roofline of k_pool<10,2>: the batch's algorithmic int32 ops over the
    # graph's device time (HIP events on the pool's stream around the graph:
    # the k_pool launch plus k_pool_init and the 2.6 MB request upload and
    # 0.2 MB result download, which take well under 1 % of it)
    km_avg = sum(kms) / len(kms)
    ops_avg = sum(opss) / len(opss)
    achieved = ops_avg / (km_avg / 1e3) / 1e12
    out = {
        "roofline": {"bound": "valu", "kernel": "k_pool<10,2>", "achieved": achieved, "peak": PEAK_INT32_TOPS,
                     "unit": "TOP/s (int32)", "frac": achieved / PEAK_INT32_TOPS,
                     "ops_per_launch": ops_avg, "graph_ms_avg": km_avg,
                     "traffic": BATCH_HBM_BYTES, "traffic_unit": "HBM bytes per launch (rocprofv3 "
                     "FETCH_SIZE/WRITE_SIZE, calibrated)" if BATCH_HBM_BYTES else None},
        "workload": "512 instances, seeds 100-611, n~U[10,22], m~U{5,10}, LB1 + Johnson LB2 (all pairs), "
                    "MinMin branching, NEH incumbents (the pool adds an IG helper wave per instance, "
                    f"{best.ig_iterations} rounds at most), solved to optimality in one graph launch",
        "solves_per_s": len(ptms) / wall, "wall_ms": wall * 1e3, "graph_ms": min(kms),
        "decomposed": nodes, "nodes_per_s": nodes / wall, "tail": tail,
        "plain_solves_per_s": len(ptms) / pwall, "plain_wall_ms": pwall * 1e3,
        "plain_explorers_per_instance": best.explorers_per_instance,
        "speedup_vs_plain": pwall / wall,
        "ig_iterations": best.ig_iterations, "without_ig_helper": noig,
        "speedup_vs_plain_without_ig_helper": (pwall * 1e3 / noig["wall_ms"]) if noig else None,
        "all_solved": ok, "optima_equal_plain": same, "optima_equal_reference": ref_match,
        "tuned": {"waves_per_group": best.waves_per_group, "explorers_per_instance": best.explorers_per_instance,
                  "steal_margin": best.steal_margin, "ig_iterations": best.ig_iterations, "source": src},
        "explorers": info["explorers"],
    }
    if not args.no_cpu:
        note("batched: reference on the host cores")
        out["cpu_baseline"] = cpu_solves(ptms, ubs, dict(bound_mode=2, branching=3, refine=1, johnson_pairs=0),
                                         args.cpu_seconds)
    return out


def leg_serving(P, PP, args):
    """The serving stream (n ~ U[12,30], m ~ U{5,10,20}, not folded) at a
    fixed Poisson rate through the resident pool, through the two routed
    pools, and through the plain one-instance-per-launch path, same arrivals
    and the same per-request time budget everywhere: a request still running
    that long after its admission is stopped and counted late."""
    import numpy as np
    N = args.serve_requests
    cfg = PP.workload_config("serving")
    ptms = PP.serving_stream(N, 10000)
    ubs = [P.neh(p)[0] for p in ptms]
    ig = None
    if args.serve_ub == "ig":
        # incumbents from the device iterated greedy, computed for the whole
        # stream in one launch before the replay (its time is reported, not
        # added to any request's latency)
        PP.ig_device(ptms[:1], chains=args.ig_chains, iterations=1, seed=1)
        t = time.perf_counter()
        _, costs = PP.ig_device(ptms, chains=args.ig_chains, iterations=args.ig_iters, seed=1)
        ig = {"chains": args.ig_chains, "iterations": args.ig_iters,
              "seconds_all_requests": round(time.perf_counter() - t, 4),
              "below_neh": int(sum(int(c) < u for c, u in zip(costs, ubs)))}
        ubs = [int(c) for c in costs]
    best, src = _params(PP, "serving", args)
    best.request_budget_s = args.serve_budget
    if args.serve_ig is not None:
        best.ig_iterations = args.serve_ig
    pcfg = PP.workload_config("serving", nb_ivms=best.explorers_per_instance)
    rate = args.serve_rate
    arrivals = PP.poisson_arrivals(N, rate, seed=7)
    span = float(arrivals[-1]) if N else 0.0
    tb = span + 2 * args.serve_budget + 15.0
    paths = set(args.serve_paths.split(","))
    note(f"serving: {N} requests at {rate:.1f}/s ({span:.0f} s of arrivals), budget {args.serve_budget} s "
         f"per request, paths {sorted(paths)}")
    res = rres = plain = None
    if "pool" in paths:
        with PP.Pool(cfg, params=best) as pool:
            res, lat, wall = pool.replay(ptms, ubs, arrivals, time_budget=tb)
        note(f"serving: pool done in {wall:.1f} s")
    rprm = PP.serving_routed_params(best)
    if "routed" in paths:
        with PP.Pool(cfg, params=rprm[1]) as p20, PP.Pool(cfg, params=rprm[0]) as p10:
            rres, rlat, rwall = PP.replay_routed([p20, p10], ptms, ubs, arrivals, time_budget=tb)
        note(f"serving: routed done in {rwall:.1f} s")
    if "plain" in paths:
        plain, plat, pwall = PP.plain_replay(ptms, ubs, arrivals, config=pcfg, time_budget=tb,
                                             request_budget=args.serve_budget)
        note(f"serving: plain done in {pwall:.1f} s")

    def agree(a, b):
        if a is None or b is None:
            return None
        return all(x.best == y.best for x, y in zip(a, b) if x.status == "solved" and y.status == "solved")

    # requests of the stream in the reference fixture (tests/golden/
    # serve_wide.json: the first ones of each machine count and job band whose
    # search a CPU finishes): every solved one against the reference's optimum
    ref_check = None
    fx = os.path.join(ROOT, "tests", "golden", "serve_wide.json")
    if os.path.exists(fx):
        with open(fx) as f:
            want = {d["seed"]: d["opt"] for d in json.load(f)["instances"]}
        pairs = [(r.best, want[10000 + i]) for path in (res, rres, plain) if path is not None
                 for i, r in enumerate(path) if 10000 + i in want and r.status == "solved"]
        ref_check = {"checked": len(pairs), "equal": sum(a == b for a, b in pairs),
                     "fixture_requests_in_stream": sum(1 for i in range(N) if 10000 + i in want)}

    out = {
        "workload": f"{N} Poisson requests at {rate:g}/s, seeds 10000-{10000 + N - 1}, n~U[12,30], "
                    f"m~U{{5,10,20}}, LB1 + Johnson LB2 (all pairs), MinMin, {'IG' if ig else 'NEH'} incumbents, per-request budget "
                    f"{args.serve_budget:g} s after admission (over-budget requests are stopped and counted late); "
                    f"the pools add an IG helper wave per request ({best.ig_iterations} rounds at most), "
                    "the plain path has the incumbents alone",
        "requests": N, "rate_per_s": rate, "request_budget_s": args.serve_budget, "paths": sorted(paths),
        "incumbent": args.serve_ub, "ig": ig,
    }
    if res is not None:
        by_m = {}
        for p, r, l in zip(ptms, res, lat):
            d = by_m.setdefault(int(p.shape[0]), {"requests": 0, "late": 0})
            d["requests"] += 1
            d["late"] += int(r.status != "solved" or l < 0)
        out.update(_latency(lat, wall, res))
        out["late_by_machines"] = by_m
        out["decomposed"] = int(sum(r.decomposed for r in res))
    if rres is not None:
        out["routed"] = {"pools": [{"machines": q.machines, "max_jobs": q.max_jobs, "groups": q.groups} for q in rprm],
                         **_latency(rlat, rwall, rres), "optima_equal_plain": agree(rres, plain),
                         "optima_equal_pool": agree(rres, res)}
    if plain is not None:
        out["plain"] = {**_latency(plat, pwall, plain), "explorers_per_instance": best.explorers_per_instance}
    out.update({
        "optima_equal_plain": agree(res, plain),
        "optima_vs_reference": ref_check,
        "tuned": {"waves_per_group": best.waves_per_group, "explorers_per_instance": best.explorers_per_instance,
                  "steal_margin": best.steal_margin, "ig_iterations": best.ig_iterations, "source": src},
    })
    return out


def leg_treesizes(P, PP, args):
    """Tree sizes of the serving stream's distribution (its first
    --tree-requests requests): every request solved alone in turn through a
    pool with the serving tuning and a per-request budget of
    --tree-budget s, NEH incumbent, or with --tree-ub ig the device iterated
    greedy's (--ig-chains chains of --ig-iters iterations per request, all
    requests in one launch, timed).  Reports decomposed nodes per request
    (a lower bound for the requests the budget stopped) as a histogram by
    machine count and decade."""
    import numpy as np
    K = args.tree_requests
    ptms = PP.serving_stream(K, 10000)
    nehs = [P.neh(p)[0] for p in ptms]
    ubs, ig = nehs, None
    if args.tree_ub == "ig":
        PP.ig_device(ptms[:1], chains=args.ig_chains, iterations=1, seed=1)  # load the code object
        t = time.perf_counter()
        _, costs = PP.ig_device(ptms, chains=args.ig_chains, iterations=args.ig_iters, seed=1)
        ig_s = time.perf_counter() - t
        ubs = [int(c) for c in costs]
        ig = {"chains": args.ig_chains, "iterations": args.ig_iters, "seconds_all_requests": round(ig_s, 4),
              "below_neh": sum(u < h for u, h in zip(ubs, nehs))}
        note(f"tree sizes: IG incumbents in {ig_s:.3f} s, {ig['below_neh']} of {K} below NEH")
    best, _ = _params(PP, "serving", args)
    best.request_budget_s = args.tree_budget
    best.slots = 1
    rows = []
    with PP.Pool(PP.workload_config("serving"), params=best) as pool:
        for i, (p, u) in enumerate(zip(ptms, ubs)):
            t = time.perf_counter()
            r, _ = pool.solve([p], [u], time_budget=args.tree_budget + 30.0)
            rows.append((int(p.shape[0]), int(p.shape[1]), r[0].status, int(r[0].decomposed),
                         time.perf_counter() - t, int(u), int(r[0].best)))
            if (i + 1) % 25 == 0:
                note(f"tree sizes: {i + 1} / {K}")
    hist = {}
    for m, n, st, dec, sec, _, _ in rows:
        dcd = int(np.floor(np.log10(max(dec, 1))))
        key = f"m{m}"
        h = hist.setdefault(key, {})
        k2 = f"1e{dcd}" + ("+" if st != "solved" else "")
        h[k2] = h.get(k2, 0) + 1
    return {"requests": K, "budget_s": args.tree_budget, "incumbent": args.tree_ub, "ig": ig,
            "histogram_by_machines": hist,
            "solved": sum(r[2] == "solved" for r in rows),
            "rows": [{"m": m, "n": n, "status": st, "decomposed": dec, "seconds": round(sec, 4), "ub": u,
                      "best": b} for m, n, st, dec, sec, u, b in rows]}


def leg_long(P, args):
    """The headline workload run for --long-seconds of node-budgeted steps:
    the rate of every step, to show whether the short timed window is
    representative of a long (120 s-capped) run."""
    n, m, seed = 50, 20, 2
    ptm = P.synth(seed, n, m)
    ub, _ = P.neh(ptm)
    cfg = P.default_config(bound_mode=0, branching=1, refine=0, johnson_pairs=0, early_stop_johnson=0,
                           nb_ivms=args.explorers)
    curve = []
    with P.GpuBB(ptm, cfg, ub=ub) as eng:
        eng.init_full_interval()
        t0 = time.perf_counter()
        prev = eng.stats().decomposed
        while time.perf_counter() - t0 < args.long_seconds:
            t = time.perf_counter()
            done = eng.next(node_budget=args.node_budget)
            st = eng.stats()
            curve.append((round(time.perf_counter() - t0, 2), (st.decomposed - prev) / (time.perf_counter() - t)))
            prev = st.decomposed
            if len(curve) % 100 == 0:
                note(f"long: {curve[-1][0]:.0f} s, {st.decomposed / 1e9:.1f} G nodes")
            if done:
                break
        total = time.perf_counter() - t0
        st = eng.stats()
        best, _ = eng.best()
    rates = [r for _, r in curve]
    return {"seconds": total, "steps": len(curve), "decomposed": st.decomposed, "nodes_per_s": st.decomposed / total,
            "step_rate_min": min(rates), "step_rate_max": max(rates),
            "step_rate_first": rates[0], "step_rate_last": rates[-1], "best": int(best),
            "curve": [[t, round(r)] for t, r in curve]}


def leg_nodes(P, args):
    """The headline: decomposed nodes/s on the 50x20 instance, K timed
    node-budgeted launches after W warm-up launches, with the int32 VALU
    roofline of k_explore<20,0> and the reference CPU baseline."""
    import torch
    n, m, seed = 50, 20, 2
    ptm = P.synth(seed, n, m)
    ub, _ = P.neh(ptm)
    cfg = P.default_config(bound_mode=0, branching=1, refine=0, johnson_pairs=0, early_stop_johnson=0,
                           nb_ivms=args.explorers)
    eng = P.GpuBB(ptm, cfg, ub=ub)
    eng.init_full_interval()
    for _ in range(args.warmup):
        eng.next(node_budget=args.node_budget)

    torch.cuda.synchronize()
    st0 = eng.stats()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.next(node_budget=args.node_budget)
        kernel_ms += eng.stats().kernel_ms
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st1 = eng.stats()
    best, _ = eng.best()
    eng.close()

    nodes = st1.decomposed - st0.decomposed
    value = nodes / wall
    ops = sum((st1.depth_histo[d] - st0.depth_histo[d]) * lb1_ops(d, n, m) for d in range(1, n - 1))
    achieved = ops / (kernel_ms / 1e3) / 1e12
    nodes_per_launch = nodes / args.steps
    traffic = HBM_BYTES_PER_NODE * nodes_per_launch
    line = {
        "metric": "decomposed nodes/sec (50x20 PFSP, LB1)",
        "value": value,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (SplitMix64 seed 2, times uniform [1,99])",
        "config": {"workload": "one 50x20 PFSP instance, LB1 bounding, MaxSum branching, NEH incumbent, "
                               "node-budgeted resident launches",
                   "n": n, "m": m, "seed": seed, "explorers": st1.explorers,
                   "explorers_stated": "16384-65536",
                   "explorers_note": "one explorer per 64-lane wavefront; 4096 is the resident maximum of "
                                     "k_explore<20,0> on gfx950 (128 VGPRs: 4 waves per SIMD x 4 SIMDs x 256 CUs), "
                                     "and a persistent launch needs every explorer resident",
                   "ub": int(ub), "best": int(best), "node_budget_per_step": args.node_budget},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT32_TOPS, "unit": "TOP/s (int32)",
                     "frac": achieved / PEAK_INT32_TOPS, "traffic": traffic,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
                     "hbm_GBs": traffic / (kernel_ms / args.steps / 1e3) / 1e9,
                     "hbm_frac": traffic / (kernel_ms / args.steps / 1e3) / 1e9 / PEAK_HBM_GBS,
                     "kernel": "k_explore<20,0>",
                     "kernel_ms_per_step": kernel_ms / args.steps,
                     "ops_per_step": ops / args.steps},
    }
    note(f"nodes leg: {value / 1e6:.1f} M nodes/s, roofline frac {achieved / PEAK_INT32_TOPS:.3f}")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ptm, ub, args.cpu_seconds)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--node-budget", type=int, default=400_000_000)
    ap.add_argument("--explorers", type=int, default=1 << 20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="nodes,small,batched,serving",
                    help="any of nodes,small,batched,serving (default) and treesizes,long")
    ap.add_argument("--serve-requests", type=int, default=SERVE_REQUESTS)
    ap.add_argument("--serve-rate", type=float, default=SERVE_RATE, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-budget", type=float, default=SERVE_BUDGET_S,
                    help="per-request time budget after admission (s)")
    ap.add_argument("--tree-ub", choices=("neh", "ig"), default="neh",
                    help="incumbent of the tree-size leg: NEH, or the device iterated greedy")
    ap.add_argument("--serve-ig", type=int, default=None,
                    help="IG helper rounds per request in the serving pools (default: the stored tuning's)")
    ap.add_argument("--serve-ub", choices=("neh", "ig"), default="neh",
                    help="incumbents of the serving stream: NEH, or the device iterated greedy (precomputed)")
    ap.add_argument("--ig-chains", type=int, default=64)
    ap.add_argument("--ig-iters", type=int, default=100)
    ap.add_argument("--serve-paths", default="pool,routed,plain",
                    help="which of pool,routed,plain the serving leg replays")
    ap.add_argument("--autotune", action="store_true", help="re-run the pool autotuner instead of tuned.json")
    ap.add_argument("--tree-requests", type=int, default=300)
    ap.add_argument("--tree-budget", type=float, default=20.0)
    ap.add_argument("--long-seconds", type=float, default=120.0)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import pbbperm_amd as P

    if args.copy_probe:
        src = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    legs = set(args.legs.split(","))
    line = leg_nodes(P, args) if "nodes" in legs else {"legs": sorted(legs)}
    if "small" in legs:
        line["small_configs"] = leg_small(P, args)
        note(f"small configs: {line['small_configs']}")
    if "long" in legs:
        line["long"] = leg_long(P, args)
        note(f"long run: {line['long']['nodes_per_s'] / 1e6:.1f} M nodes/s over {line['long']['seconds']:.0f} s")
    if legs & {"batched", "serving", "treesizes"}:
        import pbbperm_amd.pool as PP
        if "batched" in legs:
            line["batched"] = leg_batched(P, PP, args)
            print(json.dumps({"leg": "batched", **line["batched"]}), file=sys.stderr, flush=True)
        if "treesizes" in legs:
            line["treesizes"] = leg_treesizes(P, PP, args)
        if "serving" in legs:
            line["serving"] = leg_serving(P, PP, args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
