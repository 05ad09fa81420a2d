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

Usage: python tests/golden/make_pool_golden.py [batch|serve]
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


if __name__ == "__main__":
    which = sys.argv[1:] or ["batch", "serve"]
    if "batch" in which:
        make("batch", BATCH, os.path.join(HERE, "batch_c4.json"))
    if "serve" in which:
        make("serve", SERVE, os.path.join(HERE, "serve_small.json"))
