"""Autotune the multi-instance workloads and store the choices.

  python -m pbbperm_amd.tune [--out PATH] [batched] [serving]

For each workload the autotuner (pbbperm_amd.pool.autotune) scores every
(waves per workgroup, explorers per instance, steal margin) candidate on a
tuning sample of the workload's distribution with seeds disjoint from the
measured stream -- the batch's makespan for "batched"; for "serving", the
mean latency of a Poisson replay of the sample at the serving leg's rate (20
requests/s) and per-request budget (2 s, stopped requests counted at their
latency) -- and the best is written to PATH
(default pbbperm_amd/tuned.json) with the trial table.  bench.py reads that
file.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("workloads", nargs="*", default=["batched", "serving"])
    args = ap.parse_args(argv)
    from . import pool as PP
    out = args.out or PP.TUNED_PATH
    data = {}
    if os.path.exists(out):
        with open(out) as f:
            data = json.load(f)
    for name in args.workloads:
        t0 = time.time()
        best, trials = PP.tune_workload(
            name, progress=lambda r: print(f"[tune {name}] W={r[0]} epi={r[1]} margin={r[2]} -> {r[3]:.1f} ms",
                                           file=sys.stderr, flush=True))
        data[name] = {
            "waves_per_group": best.waves_per_group,
            "explorers_per_instance": best.explorers_per_instance,
            "steal_margin": best.steal_margin,
            "trials": [[int(r[0]), int(r[1]), int(r[2]), (float(r[3]) if r[3] < 1e300 else None)] for r in trials],
            "objective": PP.WORKLOADS[name]["objective"],
            "tune_s": round(time.time() - t0, 1),
        }
        print(json.dumps({name: data[name]}), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    with open(out, "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
