"""Writes tests/golden/ig_moves.json from the reference's own iterated-greedy
code (oracle/_ref/ref_harness moves: heuristic/src/IG.cpp,
fspnhoods.cpp and fastinsertremove.cpp compiled unmodified), in this
container.

For each shape: the IG object's visit order (IG.cpp:38-52) and, for seeded
partial sequences (the harness's splitmix64 stream), the makespans of
fastInsertRemove::insertMakespans (the job inserted before each position,
with the unwritten tail cell at the end of the sequence zeroed) and
removeMakespans (each position removed).

Usage: python tests/golden/make_ig_golden.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

# (instance seed, n, m, samples); n = 3 and 63 are the engine's size limits
SHAPES = [(7, 3, 2, 40), (5, 9, 4, 80), (0, 20, 5, 80), (1, 20, 10, 80), (2, 50, 20, 40), (11, 63, 20, 20)]


def main():
    O.build(ref=True)
    assert O.have_ref(), "oracle/_ref/ref_harness is required"
    out = []
    for k, (seed, n, m, ns) in enumerate(SHAPES):
        ptm = O.synth_ptm(seed, n, m)
        vo, rows = O.ref_moves(ptm, ns, 11 + k)
        out.append({"seed": seed, "n": n, "m": m, "visit_order": vo, "samples": rows})
    path = os.path.join(HERE, "ig_moves.json")
    with open(path, "w") as f:
        json.dump({"shapes": out}, f, separators=(",", ":"))
    print("ig moves:", sum(len(s["samples"]) for s in out), "samples ->", path)


if __name__ == "__main__":
    main()
