"""CPU oracle for the IVM branch-and-bound path (TEST INFRASTRUCTURE ONLY).

ctypes bindings over ``liboracle.so`` (``pbb_oracle.c``, a C restatement of the
reference's bounds/, multicore/ivm*, sequentialbb and fastNEH code) and a
runner for ``_ref/ref_harness`` (the reference's own sources, compiled
unmodified).  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this package, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import itertools
import json
import os
import subprocess
import threading
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
REF_TREE = "/root/reference"

_lib = None

I32P = C.POINTER(C.c_int32)
IP = C.POINTER(C.c_int)


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref/ref_harness when the reference tree exists)."""
    targets = ["all"]
    if ref and os.path.isdir(REF_TREE):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_taillard.argtypes = [C.c_int, IP, I32P]
        L.orc_synth_ptm.argtypes = [C.c_uint64, C.c_int, C.c_int, I32P]
        L.orc_synth_shape.argtypes = [C.c_uint64, C.c_int, C.c_int, IP, C.c_int, IP, IP]
        L.orc_inst_new.argtypes = [C.c_int, C.c_int, I32P]
        L.orc_inst_new.restype = C.c_void_p
        L.orc_inst_free.argtypes = [C.c_void_p]
        L.orc_makespan.argtypes = [C.c_void_p, IP]
        L.orc_lb1_children.argtypes = [C.c_void_p, IP, C.c_int, C.c_int, C.c_int, IP, IP, IP, IP]
        L.orc_lb1_node.argtypes = [C.c_void_p, IP, C.c_int, C.c_int]
        L.orc_lb2_node.argtypes = [C.c_void_p, IP, C.c_int, C.c_int, C.c_int, IP]
        L.orc_neh.argtypes = [C.c_void_p, IP]
        L.orc_bb_new.argtypes = [C.c_void_p, C.c_void_p, C.c_int, IP]
        L.orc_bb_new.restype = C.c_void_p
        L.orc_bb_free.argtypes = [C.c_void_p]
        L.orc_bb_set_root.argtypes = [C.c_void_p, IP]
        L.orc_bb_init_interval.argtypes = [C.c_void_p, IP, IP]
        L.orc_bb_init_full.argtypes = [C.c_void_p]
        L.orc_bb_next.argtypes = [C.c_void_p]
        L.orc_bb_run.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_bb_stats.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_bb_best.argtypes = [C.c_void_p, IP]
        L.orc_bb_line.argtypes = [C.c_void_p]
        L.orc_bb_interval.argtypes = [C.c_void_p, IP, IP]
        L.orc_bb_share.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.orc_bb_depth_histo.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.orc_bounds_batch.argtypes = [C.c_void_p, C.c_int, IP, IP, IP, C.c_int, IP, IP, IP, IP, IP, IP]
        L.orc_insert_makespans.argtypes = [C.c_void_p, IP, C.c_int, C.c_int, IP]
        L.orc_remove_makespans.argtypes = [C.c_void_p, IP, C.c_int, IP]
        L.orc_ig_visit_order.argtypes = [C.c_int, IP]
        L.orc_ig_below.argtypes = [C.POINTER(C.c_uint64), C.c_uint32]
        L.orc_ig_below.restype = C.c_uint32
        L.orc_ig_chain_seed.argtypes = [C.c_uint64, C.c_int, C.c_int]
        L.orc_ig_chain_seed.restype = C.c_uint64
        L.orc_ig_thresholds.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.orc_ig.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_uint64), IP]
        _lib = L
    return _lib


def _ip(a: np.ndarray):
    return a.ctypes.data_as(IP)


def _i32(a: np.ndarray):
    return a.ctypes.data_as(I32P)


# ---------------------------------------------------------------- generators

def taillard(ta_id: int) -> np.ndarray:
    """PTM (m x n, machine-major) of Taillard instance ta<id>."""
    buf = np.zeros(64 * 512, dtype=np.int32)
    m = C.c_int()
    n = lib().orc_taillard(ta_id, C.byref(m), _i32(buf))
    return buf[: m.value * n].reshape(m.value, n).copy()


def synth_ptm(seed: int, n: int, m: int) -> np.ndarray:
    buf = np.zeros(n * m, dtype=np.int32)
    lib().orc_synth_ptm(seed, n, m, _i32(buf))
    return buf.reshape(m, n)


def synth_shape(seed: int, nlo: int, nhi: int, mchoices: Sequence[int]) -> Tuple[int, int]:
    mc = np.asarray(mchoices, dtype=np.intc)
    n = C.c_int()
    m = C.c_int()
    lib().orc_synth_shape(seed, nlo, nhi, _ip(mc), len(mc), C.byref(n), C.byref(m))
    return n.value, m.value


# ---------------------------------------------------------------- instance

class Instance:
    def __init__(self, ptm: np.ndarray):
        ptm = np.ascontiguousarray(ptm, dtype=np.int32)
        self.m, self.n = ptm.shape
        self.ptm = ptm
        self.h = lib().orc_inst_new(self.n, self.m, _i32(self.ptm))
        if not self.h:
            raise ValueError("unsupported instance shape")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_inst_free(self.h)
            self.h = None

    def makespan(self, perm) -> int:
        p = np.ascontiguousarray(perm, dtype=np.intc)
        return lib().orc_makespan(self.h, _ip(p))

    def lb1_children(self, sched, l1: int, l2: int, both: bool = True):
        s = np.ascontiguousarray(sched, dtype=np.intc)
        cb = np.zeros(self.n, np.intc)
        ce = np.zeros(self.n, np.intc)
        pb = np.zeros(self.n, np.intc)
        pe = np.zeros(self.n, np.intc)
        lib().orc_lb1_children(self.h, _ip(s), l1, l2, int(both), _ip(cb), _ip(ce), _ip(pb), _ip(pe))
        return cb, ce, pb, pe

    def lb1_node(self, perm, l1: int, l2: int) -> int:
        p = np.ascontiguousarray(perm, dtype=np.intc)
        return lib().orc_lb1_node(self.h, _ip(p), l1, l2)

    def lb2_node(self, perm, l1: int, l2: int, pairs: int = 0) -> Tuple[int, int]:
        p = np.ascontiguousarray(perm, dtype=np.intc)
        idle = C.c_int(-7)
        v = lib().orc_lb2_node(self.h, _ip(p), l1, l2, pairs, C.byref(idle))
        return v, idle.value

    def bounds_batch(self, perms, l1, l2, pairs: int = 0,
                     which=("cb", "ce", "cbf", "lb1n", "lb2", "mk")):
        """Bound values for arrays of partial schedules: dict of numpy arrays."""
        perms = np.ascontiguousarray(perms, np.intc)
        l1 = np.ascontiguousarray(l1, np.intc)
        l2 = np.ascontiguousarray(l2, np.intc)
        B = perms.shape[0]
        out = {k: np.zeros((B, self.n), np.intc) for k in ("cb", "ce", "cbf") if k in which}
        out.update({k: np.zeros(B, np.intc) for k in ("lb1n", "lb2", "mk") if k in which})
        if "cb" in out and "ce" not in out:
            out["ce"] = np.zeros((B, self.n), np.intc)
        ptr = lambda k: _ip(out[k]) if k in out else None
        lib().orc_bounds_batch(self.h, B, _ip(perms), _ip(l1), _ip(l2), pairs, ptr("cb"), ptr("ce"),
                               ptr("cbf"), ptr("lb1n"), ptr("lb2"), ptr("mk"))
        return out

    def neh(self) -> Tuple[int, np.ndarray]:
        p = np.arange(self.n, dtype=np.intc)
        c = lib().orc_neh(self.h, _ip(p))
        return c, p

    # ---- iterated greedy (heuristic/src/IG.cpp, fastinsertremove.cpp)
    def insert_makespans(self, seq, job: int) -> np.ndarray:
        s = np.ascontiguousarray(seq, dtype=np.intc)
        out = np.zeros(len(s) + 1, np.intc)
        lib().orc_insert_makespans(self.h, _ip(s), len(s), int(job), _ip(out))
        return out

    def remove_makespans(self, seq) -> np.ndarray:
        s = np.ascontiguousarray(seq, dtype=np.intc)
        out = np.zeros(len(s), np.intc)
        lib().orc_remove_makespans(self.h, _ip(s), len(s), _ip(out))
        return out

    def ig_thresholds(self) -> np.ndarray:
        thr = np.zeros(IG_THR, np.uint64)
        lib().orc_ig_thresholds(self.h, thr.ctypes.data_as(C.POINTER(C.c_uint64)))
        return thr

    def ig(self, state: int, iters: int) -> Tuple[int, np.ndarray]:
        """IG.cpp runIG from NEH with the seeded stream `state` (see
        orc_ig_chain_seed for the state of a device chain)."""
        thr = self.ig_thresholds()
        p = np.zeros(self.n, np.intc)
        c = lib().orc_ig(self.h, C.c_uint64(state), int(iters), thr.ctypes.data_as(C.POINTER(C.c_uint64)), _ip(p))
        return c, p


IG_THR = 256


def ig_visit_order(n: int) -> np.ndarray:
    vo = np.zeros(n, np.intc)
    lib().orc_ig_visit_order(n, _ip(vo))
    return vo


def ig_chain_seed(seed: int, inst: int, chain: int) -> int:
    return int(lib().orc_ig_chain_seed(C.c_uint64(seed), int(inst), int(chain)))


def ref_moves(ptm: np.ndarray, samples: int, seed: int):
    """The reference's IG visit order and fastInsertRemove insert/remove
    makespans on seeded partial sequences (oracle/_ref/ref_harness moves)."""
    import tempfile
    m, n = ptm.shape
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        f.write(f"{n} {m}\n" + "\n".join(" ".join(str(int(v)) for v in row) for row in ptm) + "\n")
        path = f.name
    try:
        out = subprocess.run([REF_HARNESS, "moves", path, str(samples), str(seed)], capture_output=True, text=True,
                             check=True).stdout.splitlines()
    finally:
        os.unlink(path)
    vo = [int(x) for x in out[0].split()[1:]]
    rows = []
    for line in out[1:]:
        head, seq, ins, rem = line.split("|")
        ln, job = (int(x) for x in head.split())
        rows.append({"job": job, "seq": [int(x) for x in seq.split()], "ins": [int(x) for x in ins.split()],
                     "rem": [int(x) for x in rem.split()]})
    return vo, rows


# ---------------------------------------------------------------- B&B

class _Cfg(C.Structure):
    _fields_ = [("bound_mode", C.c_int), ("branching", C.c_int), ("refine", C.c_int),
                ("johnson_pairs", C.c_int), ("find_all", C.c_int), ("single_node", C.c_int)]


class _Stats(C.Structure):
    _fields_ = [("decomposed", C.c_uint64), ("leaves", C.c_uint64),
                ("simple_bounds", C.c_uint64), ("johnson_bounds", C.c_uint64)]


@dataclass
class SolveResult:
    finished: bool
    best: int
    decomposed: int
    leaves: int
    simple_bounds: int
    johnson_bounds: int
    perm: List[int]
    wall_s: float = 0.0


class BB:
    """Single-explorer IVM B&B (sequentialbb semantics)."""

    def __init__(self, inst: Instance, ub: int, bound_mode=0, branching=1, refine=0,
                 johnson_pairs=0, find_all=False, single_node=True, root=None):
        self.inst = inst
        cfg = _Cfg(bound_mode, branching, refine, johnson_pairs, int(find_all), int(single_node))
        self.h = lib().orc_bb_new(inst.h, C.byref(cfg), ub, None)
        r = np.arange(inst.n, dtype=np.intc) if root is None else np.ascontiguousarray(root, np.intc)
        lib().orc_bb_set_root(self.h, _ip(r))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_bb_free(self.h)
            self.h = None

    def init_full(self) -> int:
        return lib().orc_bb_init_full(self.h)

    def init_interval(self, pos, end) -> int:
        p = np.ascontiguousarray(pos, np.intc)
        e = np.ascontiguousarray(end, np.intc)
        return lib().orc_bb_init_interval(self.h, _ip(p), _ip(e))

    def next(self) -> bool:
        return bool(lib().orc_bb_next(self.h))

    def run(self, max_nodes: int = 0) -> bool:
        return bool(lib().orc_bb_run(self.h, max_nodes))

    def stats(self) -> _Stats:
        st = _Stats()
        lib().orc_bb_stats(self.h, C.byref(st))
        return st

    def best(self) -> Tuple[int, np.ndarray]:
        p = np.zeros(self.inst.n, np.intc)
        c = lib().orc_bb_best(self.h, _ip(p))
        return c, p

    def line(self) -> int:
        return lib().orc_bb_line(self.h)

    def interval(self):
        p = np.zeros(self.inst.n, np.intc)
        e = np.zeros(self.inst.n, np.intc)
        lib().orc_bb_interval(self.h, _ip(p), _ip(e))
        return p, e

    def share_to(self, thief: "BB", margin: int = 10) -> int:
        return lib().orc_bb_share(self.h, thief.h, margin)

    def depth_histo(self) -> np.ndarray:
        h = np.zeros(self.inst.n + 1, np.uint64)
        lib().orc_bb_depth_histo(self.h, h.ctypes.data_as(C.POINTER(C.c_uint64)))
        return h


def solve(ptm: np.ndarray, ub: int, bound_mode=0, branching=1, refine=0, johnson_pairs=0,
          find_all=False, max_nodes=0, pos=None, end=None) -> SolveResult:
    inst = Instance(ptm)
    bb = BB(inst, ub, bound_mode, branching, refine, johnson_pairs, find_all)
    if pos is None:
        bb.init_full()
    else:
        bb.init_interval(pos, end)
    fin = bb.run(max_nodes)
    st = bb.stats()
    c, p = bb.best()
    return SolveResult(fin, c, st.decomposed, st.leaves, st.simple_bounds, st.johnson_bounds,
                       p.tolist())


# ---------------------------------------------------------------- reference harness

def random_partials(n: int, count: int, seed: int, edge_frac: float = 0.01):
    """Seeded random partial schedules (Fisher-Yates perm, l1 ~ U[-1, n-2],
    l2 ~ U[l1+1, n]) with a fraction of edge cases (empty front, empty back,
    l2-l1 in {1, 2})."""
    rng = np.random.default_rng(seed)
    perms = np.argsort(rng.random((count, n)), axis=1).astype(np.intc)
    l1 = rng.integers(-1, n - 1, size=count).astype(np.intc)
    l2 = (l1 + 1 + (rng.random(count) * (n - l1)).astype(np.intc)).astype(np.intc)
    l2 = np.minimum(l2, n).astype(np.intc)
    ne = max(4, int(count * edge_frac))
    kinds = rng.integers(0, 4, size=ne)
    idx = rng.choice(count, size=ne, replace=False)
    for i, k in zip(idx, kinds):
        if k == 0:
            l1[i] = -1
        elif k == 1:
            l2[i] = n
        elif k == 2:
            l2[i] = l1[i] + 1
        else:
            l1[i] = min(l1[i], n - 2)
            l2[i] = l1[i] + 2
    return perms, l1, l2


def ptm_text(ptm: np.ndarray) -> str:
    m, n = ptm.shape
    return f"{n} {m}\n" + "\n".join(" ".join(str(int(v)) for v in row) for row in ptm) + "\n"


def have_ref() -> bool:
    return os.path.exists(REF_HARNESS)


_tmp_seq = itertools.count()


def _write_tmp(ptm: np.ndarray, tmpdir: str) -> str:
    """A file of its own per call (callers run the harness from many threads)."""
    os.makedirs(tmpdir, exist_ok=True)
    path = os.path.join(tmpdir, f"ptm_{os.getpid()}_{threading.get_ident()}_{next(_tmp_seq)}.txt")
    with open(path, "w") as f:
        f.write(ptm_text(ptm))
    return path


def ref_solve(ptm: np.ndarray, ub: int, bound_mode=0, branching=1, refine=0, johnson_pairs=0,
              find_all=False, max_nodes=0, time_limit=0.0, pos=None, end=None,
              tmpdir="/tmp/pbb_ref") -> SolveResult:
    path = _write_tmp(ptm, tmpdir)
    args = [REF_HARNESS, "solve", path, str(ub), str(bound_mode), str(branching), str(refine),
            str(johnson_pairs), str(int(find_all)), str(max_nodes), str(time_limit)]
    if pos is not None:
        args += [str(int(v)) for v in pos] + [str(int(v)) for v in end]
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
    os.unlink(path)
    d = json.loads(out.strip().splitlines()[-1])
    return SolveResult(d["finished"], d["best"], d["decomposed"], d["leaves"],
                       d["simple_bounds"], d["johnson_bounds"], d["perm"], d["wall_s"])


def ref_bounds(ptm: np.ndarray, nsamples: int, seed: int, pairs: int = 0, tmpdir="/tmp/pbb_ref"):
    """Reference LB1 children / LB1' / LB2 / makespan on seeded random partial schedules."""
    path = _write_tmp(ptm, tmpdir)
    out = subprocess.run([REF_HARNESS, "bounds", path, str(nsamples), str(seed), str(pairs)],
                         check=True, capture_output=True, text=True).stdout
    os.unlink(path)
    m, n = ptm.shape
    rows = []
    for line in out.strip().splitlines():
        a, b, c, d = line.split("|")
        head = [int(x) for x in a.split()]
        tail = [int(x) for x in d.split()]
        rows.append((head[0], head[1], head[2:], [int(x) for x in b.split()],
                     [int(x) for x in c.split()], tail[0], tail[1], tail[2], tail[3]))
    return rows


def ref_neh(ptm: np.ndarray, tmpdir="/tmp/pbb_ref"):
    """The reference's fastNEH (initialSort + runNEH, identity start) with its
    uninitialised tail cell zeroed by the harness: (cost, order)."""
    return ref_neh_many([ptm], tmpdir)[0]


def ref_neh_many(ptms: Sequence[np.ndarray], tmpdir="/tmp/pbb_ref"):
    """ref_neh over many instances in one harness process."""
    os.makedirs(tmpdir, exist_ok=True)
    path = os.path.join(tmpdir, f"neh_{os.getpid()}_{len(ptms)}.txt")
    with open(path, "w") as f:
        for p in ptms:
            f.write(ptm_text(p))
    out = subprocess.run([REF_HARNESS, "neh", path], check=True, capture_output=True,
                         text=True).stdout
    os.unlink(path)
    res = []
    for line in out.strip().splitlines():
        v = [int(x) for x in line.split()]
        res.append((v[0], v[1:]))
    assert len(res) == len(ptms), (len(res), len(ptms))
    return res


def ref_taillard(ta_id: int) -> np.ndarray:
    out = subprocess.run([REF_HARNESS, "taillard", str(ta_id)], check=True, capture_output=True,
                         text=True).stdout
    v = [int(x) for x in out.split()]
    n, m = v[0], v[1]
    return np.array(v[2:2 + n * m], dtype=np.int32).reshape(m, n)
