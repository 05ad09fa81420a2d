"""Writes the multi-instance fixtures of tests/golden from the reference's
own code (oracle/_ref/ref_harness) and the C oracle, in this container.

  batch_c4.json   the 512-instance batch: seeds 100-611, shape n ~ U[10,22],
                  m ~ U{5,10} from each seed's shape stream, times from its
                  PTM stream.  boundingMode 2 with the Johnson refinement
                  (LB1 + LB2, all machine pairs), MinMin branching.  Per
                  instance: the NEH value, the optimum found by the
                  reference from the NEH bound, and the decomposed / leaf /
                  LB1 / LB2 counts of the reference and of the oracle with
                  the bound frozen at the optimum (they must agree).
  serve_small.json  a 256-request stream for serving tests: seeds
                  20000-20255, n ~ U[12,16], m ~ U{5,10,20}, same bounding,
                  MaxSum branching; optimum and frozen-bound counts.

  serve_wide.json   the serving stream's full shape range, stratified: for
                  m in {5, 10, 20} and n in 12-16, 17-20, 21-25, 26-30, the
                  first requests of the stream (seeds 10000+) of that stratum
                  whose search from the NEH bound ends within a node cap on
                  the C oracle (at most 40 probed per stratum: the long tail
                  of the stream cannot be solved on a CPU in reasonable time;
                  the fixture records how many were probed and kept), then
                  solved by the reference (optimum from NEH; counts with the
                  bound frozen at it).

Usage: python tests/golden/make_pool_golden.py [batch|serve|wide]
"""
from __future__ import annotations

import json
import os
import sys
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

BATCH = dict(seeds=list(range(100, 612)), nlo=10, nhi=22, mch=(5, 10), bound_mode=2, branching=3, refine=1,
             pairs=0)
SERVE = dict(seeds=list(range(20000, 20256)), nlo=12, nhi=16, mch=(5, 10, 20), bound_mode=2, branching=1,
             refine=1, pairs=0)


def one(args):
    seed, spec = args
    n, m = O.synth_shape(seed, spec["nlo"], spec["nhi"], spec["mch"])
    ptm = O.synth_ptm(seed, n, m)
    neh, _ = O.Instance(ptm).neh()
    kw = dict(bound_mode=spec["bound_mode"], branching=spec["branching"], refine=spec["refine"],
              johnson_pairs=spec["pairs"])
    r_ref = O.ref_solve(ptm, neh, **kw)
    opt = r_ref.best
    o_imp = O.solve(ptm, neh, **kw)
    assert o_imp.best == opt, (seed, o_imp.best, opt)
    f_ref = O.ref_solve(ptm, opt, **kw)
    f_orc = O.solve(ptm, opt, **kw)
    c_ref = [f_ref.decomposed, f_ref.leaves, f_ref.simple_bounds, f_ref.johnson_bounds]
    c_orc = [f_orc.decomposed, f_orc.leaves, f_orc.simple_bounds, f_orc.johnson_bounds]
    assert c_ref == c_orc, (seed, c_ref, c_orc)
    return {"seed": seed, "n": n, "m": m, "neh": neh, "opt": opt, "counts_frozen": c_ref,
            "counts_neh_ref": [r_ref.decomposed, r_ref.leaves, r_ref.simple_bounds, r_ref.johnson_bounds]}


def make(name, spec, path):
    O.build(ref=True)
    assert O.have_ref(), "oracle/_ref/ref_harness is required"
    with Pool(8) as p:
        rows = p.map(one, [(s, spec) for s in spec["seeds"]], chunksize=1)
    meta = {k: v for k, v in spec.items() if k != "seeds"}
    meta["mch"] = list(meta["mch"])
    with open(path, "w") as f:
        json.dump({"spec": meta, "instances": rows}, f, separators=(",", ":"))
    print(name, len(rows), "instances ->", path)


WIDE = dict(seeds=range(10000, 14000), nlo=12, nhi=30, mch=(5, 10, 20), bound_mode=2, branching=3, refine=1,
            pairs=0, per_stratum=4, node_cap=3_000_000, max_probes=40)
BANDS = [(12, 16), (17, 20), (21, 25), (26, 30)]


def _wide_probe(seed):
    n, m = O.synth_shape(seed, WIDE["nlo"], WIDE["nhi"], WIDE["mch"])
    ptm = O.synth_ptm(seed, n, m)
    neh, _ = O.Instance(ptm).neh()
    r = O.solve(ptm, neh, bound_mode=WIDE["bound_mode"], branching=WIDE["branching"], refine=WIDE["refine"],
                johnson_pairs=WIDE["pairs"], max_nodes=WIDE["node_cap"])
    return seed, n, m, bool(r.finished)


def make_wide(path):
    """Stratified fixture over the whole serving shape range."""
    O.build(ref=True)
    assert O.have_ref(), "oracle/_ref/ref_harness is required"
    want = {(m, b): [] for m in WIDE["mch"] for b in range(len(BANDS))}
    tried = {k: 0 for k in want}
    seeds = list(WIDE["seeds"])
    with Pool(8) as pool:
        for i in range(0, len(seeds), 64):
            chunk = []
            for s in seeds[i:i + 64]:
                n, m = O.synth_shape(s, WIDE["nlo"], WIDE["nhi"], WIDE["mch"])
                b = next(k for k, (lo, hi) in enumerate(BANDS) if lo <= n <= hi)
                if len(want[(m, b)]) < WIDE["per_stratum"] and tried[(m, b)] + \
                        sum(1 for c in chunk if c[1] == (m, b)) < WIDE["max_probes"]:
                    chunk.append((s, (m, b)))
            if not chunk:
                continue
            for seed, n, m, ok in pool.map(_wide_probe, [c[0] for c in chunk], chunksize=1):
                b = next(k for k, (lo, hi) in enumerate(BANDS) if lo <= n <= hi)
                tried[(m, b)] += 1
                if ok and len(want[(m, b)]) < WIDE["per_stratum"]:
                    want[(m, b)].append(seed)
            if all(len(v) >= WIDE["per_stratum"] or tried[k] >= WIDE["max_probes"] for k, v in want.items()):
                break
        picked = sorted(s for v in want.values() for s in v)
        spec = dict(nlo=WIDE["nlo"], nhi=WIDE["nhi"], mch=WIDE["mch"], bound_mode=WIDE["bound_mode"],
                    branching=WIDE["branching"], refine=WIDE["refine"], pairs=WIDE["pairs"])
        rows = pool.map(one, [(s, spec) for s in picked], chunksize=1)
    meta = {k: v for k, v in spec.items()}
    meta["mch"] = list(meta["mch"])
    meta["bands"] = BANDS
    meta["node_cap"] = WIDE["node_cap"]
    meta["probed"] = {f"m{m}_n{BANDS[b][0]}-{BANDS[b][1]}": [tried[(m, b)], len(want[(m, b)])] for (m, b) in want}
    with open(path, "w") as f:
        json.dump({"spec": meta, "instances": rows}, f, separators=(",", ":"))
    print("wide", len(rows), "instances ->", path, meta["probed"])


if __name__ == "__main__":
    which = sys.argv[1:] or ["batch", "serve"]
    if "wide" in which:
        make_wide(os.path.join(HERE, "serve_wide.json"))
    if "batch" in which:
        make("batch", BATCH, os.path.join(HERE, "batch_c4.json"))
    if "serve" in which:
        make("serve", SERVE, os.path.join(HERE, "serve_small.json"))
