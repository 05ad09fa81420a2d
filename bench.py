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
  "serving"  a Poisson stream of requests (n ~ U[12,30], m ~ U{5,10,20},
             20-machine requests folded to n <= 16; 600 by default, the
             10,000-request stream with --serve-requests 10000) through the
             resident serving loop, through a 10-machine and a 20-machine
             pool resident together with each request routed to the smaller
             one that holds it ("routed"), and through the plain path, same
             stream, same arrival times: solves/s and p50/p99 latency (with
             PBB_REPLAY_PROGRESS=1 in the environment both replays report
             every 500 served requests on stderr).
Pool parameters come from the autotuner's stored choice
(pbbperm_amd/tuned.json, python -m pbbperm_amd.tune), or --autotune.
Every optimum of both legs is compared between the pool and the plain path.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--node-budget B]
                  [--legs nodes,batched,serving] [--serve-requests N]
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


def lb1_ops(depth: int, n: int, m: int) -> int:
    """Algorithmic int32 ops of one LB1 node at depth D (D fixed jobs,
    k = n - D children): front+back recompute (2m-1)D, remain m*k, the two
    child chains 2k(5m-3) (bound_fsp_weak.cpp:61-221)."""
    k = n - depth
    return (2 * m - 1) * depth + m * k + 2 * k * (5 * m - 3)


def cpu_baseline(ptm, ub, seconds: float):
    """The reference's own bound / IVM code (oracle/_ref) single-threaded on a
    bounded sample of the same workload; falls back to the C oracle port."""
    import oracle as O
    O.lib()
    if O.have_ref():
        r = O.ref_solve(ptm, ub, bound_mode=0, branching=1, time_limit=seconds)
        kind = "reference"
    else:
        t0 = time.time()
        r = O.solve(ptm, ub, bound_mode=0, branching=1, max_nodes=int(seconds * 2e5))
        r.wall_s = time.time() - t0
        kind = "port"
    return {"value": r.decomposed / r.wall_s, "unit": "nodes/s", "cores": 1, "kind": kind,
            "sample": f"single-thread DFS from the root of the same instance/UB for {r.wall_s:.1f} s "
                      f"({r.decomposed} decomposed nodes)"}


def note(msg: str):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _latency(lat, wall):
    """Throughput and latency figures of a replay over the requests completed
    within its time budget (the replay marks the others with latency -1;
    they are counted in "late")."""
    import numpy as np
    a = np.asarray(lat, dtype=float)
    done = a[a >= 0]
    has = len(done) > 0
    return {"solves_per_s": len(done) / wall if wall > 0 else 0.0,
            "p50_ms": float(np.percentile(done, 50)) if has else None,
            "p99_ms": float(np.percentile(done, 99)) if has else None,
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
        times = []
        for _ in range(3):
            t = time.perf_counter()
            with P.GpuBB(ptm, cfg, ub=ub) as eng:
                eng.solve(time_budget=60.0)
                st = eng.stats()
                best, _ = eng.best()
            times.append(time.perf_counter() - t)
        d = {"solve_ms": min(times) * 1e3, "explorers": st.explorers, "best": best, "decomposed": st.decomposed}
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
        walls, kms = [], []
        for _ in range(3):
            t = time.perf_counter()
            res, km = pool.solve(ptms, ubs, time_budget=120.0)
            walls.append(time.perf_counter() - t)
            kms.append(km)
        info = pool.info()
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
    out = {
        "workload": "512 instances, seeds 100-611, n~U[10,22], m~U{5,10}, LB1 + Johnson LB2 (all pairs), "
                    "MinMin branching, NEH incumbents, solved to optimality in one graph launch",
        "solves_per_s": len(ptms) / wall, "wall_ms": wall * 1e3, "graph_ms": min(kms),
        "decomposed": nodes, "nodes_per_s": nodes / wall,
        "plain_solves_per_s": len(ptms) / pwall, "plain_wall_ms": pwall * 1e3,
        "plain_explorers_per_instance": best.explorers_per_instance,
        "speedup_vs_plain": pwall / wall,
        "all_solved": ok, "optima_equal_plain": same, "optima_equal_reference": ref_match,
        "tuned": {"waves_per_group": best.waves_per_group, "explorers_per_instance": best.explorers_per_instance,
                  "steal_margin": best.steal_margin, "source": src},
        "explorers": info["explorers"],
    }
    if not args.no_cpu:
        note("batched: reference on the host cores")
        out["cpu_baseline"] = cpu_solves(ptms, ubs, dict(bound_mode=2, branching=3, refine=1, johnson_pairs=0),
                                         args.cpu_seconds)
    return out


def leg_serving(P, PP, args):
    import numpy as np
    N = args.serve_requests
    cfg = PP.workload_config("serving")
    ptms = PP.serving_stream(N, 10000)
    ubs = [P.neh(p)[0] for p in ptms]
    best, src = _params(PP, "serving", args)
    pcfg = PP.workload_config("serving", nb_ivms=best.explorers_per_instance)
    # arrival rate: a fraction of the plain path's closed-loop rate
    k = min(N, 200)
    _, _, w0 = PP.plain_replay(ptms[:k], ubs[:k], np.zeros(k), config=pcfg, time_budget=300.0)
    rate = args.serve_load * k / w0
    arrivals = PP.poisson_arrivals(N, rate, seed=7)
    note(f"serving: {N} requests at {rate:.1f}/s through the pool")
    with PP.Pool(cfg, params=best) as pool:
        res, lat, wall = pool.replay(ptms, ubs, arrivals, time_budget=900.0)
    note(f"serving: pool done in {wall:.1f} s; routed pools")
    rprm = PP.serving_routed_params(best)
    with PP.Pool(cfg, params=rprm[1]) as p20, PP.Pool(cfg, params=rprm[0]) as p10:
        rres, rlat, rwall = PP.replay_routed([p20, p10], ptms, ubs, arrivals, time_budget=900.0)
    note(f"serving: routed done in {rwall:.1f} s; plain path")
    plain, plat, pwall = PP.plain_replay(ptms, ubs, arrivals, config=pcfg, time_budget=900.0)
    note(f"serving: plain done in {pwall:.1f} s")
    ok = all(r.status == "solved" for r in res + rres + plain)
    same = all(a.best == b.best for a, b in zip(res, plain))
    rsame = all(a.best == b.best for a, b in zip(rres, plain))
    return {
        "workload": f"{N} Poisson requests, seeds 10000+, n~U[12,30], m~U{{5,10,20}} (m=20 folded to n<=16), "
                    "LB1 + Johnson LB2, MinMin, NEH incumbents, every request solved to optimality",
        "rate_per_s": rate, "load_vs_plain": args.serve_load,
        **_latency(lat, wall),
        "decomposed": int(sum(r.decomposed for r in res)),
        "routed": {"pools": [{"machines": q.machines, "max_jobs": q.max_jobs, "groups": q.groups} for q in rprm],
                   **_latency(rlat, rwall), "optima_equal_plain": rsame},
        "plain": {**_latency(plat, pwall), "explorers_per_instance": best.explorers_per_instance},
        "all_solved": ok, "optima_equal_plain": same,
        "tuned": {"waves_per_group": best.waves_per_group, "explorers_per_instance": best.explorers_per_instance,
                  "steal_margin": best.steal_margin, "source": src},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--node-budget", type=int, default=400_000_000)
    ap.add_argument("--explorers", type=int, default=1 << 20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="nodes,small,batched,serving")
    ap.add_argument("--serve-requests", type=int, default=600)
    ap.add_argument("--autotune", action="store_true", help="re-run the pool autotuner instead of tuned.json")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="arrival rate as a fraction of the plain path's closed-loop solves/s")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    import pbbperm_amd as P

    if args.copy_probe:
        src = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

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
                   "n": n, "m": m, "seed": seed, "explorers": st1.explorers, "ub": int(ub), "best": int(best),
                   "node_budget_per_step": args.node_budget},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT32_TOPS, "unit": "TOP/s (int32)",
                     "frac": achieved / PEAK_INT32_TOPS, "traffic": traffic,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
                     "hbm_GBs": traffic / (kernel_ms / args.steps / 1e3) / 1e9,
                     "hbm_frac": traffic / (kernel_ms / args.steps / 1e3) / 1e9 / PEAK_HBM_GBS,
                     "kernel": "k_explore<20,0>",
                     "kernel_ms_per_step": kernel_ms / args.steps,
                     "ops_per_step": ops / args.steps},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ptm, ub, args.cpu_seconds)
    legs = set(args.legs.split(","))
    if "small" in legs:
        line["small_configs"] = leg_small(P, args)
        note(f"small configs: {line['small_configs']}")
    note(f"nodes leg: {value / 1e6:.1f} M nodes/s")
    if legs & {"batched", "serving"}:
        import pbbperm_amd.pool as PP
        if "batched" in legs:
            line["batched"] = leg_batched(P, PP, args)
            print(json.dumps({"leg": "batched", **line["batched"]}), file=sys.stderr, flush=True)
        if "serving" in legs:
            line["serving"] = leg_serving(P, PP, args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
