"""Throughput benchmark: decomposed nodes/s of the IVM branch-and-bound engine
on one MI355X, on the 50 jobs x 20 machines instance (times uniform in
[1, 99], seed 2), LB1 bounding, MaxSum branching, incumbent from NEH.

A step is one resident-engine launch (pbb_run) exploring a fixed budget of
decomposed nodes; the tree is far too large to finish, so the engine keeps
its IVMs, intervals and incumbent in HBM across steps and each step
continues the exploration.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--node-budget B]
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
HBM_BYTES_PER_NODE = 5.32


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--node-budget", type=int, default=400_000_000)
    ap.add_argument("--explorers", type=int, default=1 << 20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
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
                     "kernel": "k_explore<20,false>",
                     "kernel_ms_per_step": kernel_ms / args.steps,
                     "ops_per_step": ops / args.steps},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ptm, ub, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
