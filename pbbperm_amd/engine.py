"""Host-side mirror of the reference's GPU worker engine and bound library.

* :class:`GpuBB` follows ``gpubb`` (gpu/include/gpubb.h:24-173) as driven by
  ``worker_gpu`` (distributed/src/worker_gpu.cpp:26-108): full-interval and
  factoradic-interval initialisation, ``next`` until all explorers are idle,
  interval export, incumbent in/out, interruption.
* :class:`Bounds` follows ``bound_abstract`` (bounds/include/bound_abstract.h:6-34),
  batched over device buffers.
* :func:`load_config` reads a gpuconfig.ini with the reference's keys
  (common/src/arguments.cpp:68-133).

Every call goes through the C ABI of ``libpbb_engine.so``; there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import Checkpoint, Config, PbbError, Stats, check, load

INT_MAX = 2**31 - 1


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_lib.IP)


def _i32(a: np.ndarray):
    return a.ctypes.data_as(_lib.I32P)


# --------------------------------------------------------------------- config

def default_config(**overrides) -> Config:
    cfg = Config()
    check(load().pbb_config_default(C.byref(cfg)))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, int(v))
    return cfg


def load_config(path: str) -> Tuple[Config, str, str]:
    """gpuconfig.ini -> (config, problem, instance)."""
    cfg = Config()
    prob = C.create_string_buffer(64)
    inst = C.create_string_buffer(64)
    check(load().pbb_config_load_ini(path.encode(), C.byref(cfg), prob, inst))
    return cfg, prob.value.decode(), inst.value.decode()


# ------------------------------------------------------------------ instances

def taillard(ta_id: int) -> np.ndarray:
    """Processing times (m x n, machine-major) of Taillard instance ta<id>."""
    buf = np.zeros(20 * 500, dtype=np.int32)
    n = C.c_int()
    m = C.c_int()
    check(load().pbb_instance_taillard(ta_id, C.byref(n), C.byref(m), _i32(buf)))
    return buf[: n.value * m.value].reshape(m.value, n.value).copy()


def synth(seed: int, n: int, m: int) -> np.ndarray:
    """Seeded synthetic instance, times uniform in [1, 99]."""
    buf = np.zeros(n * m, dtype=np.int32)
    check(load().pbb_instance_synth(seed, n, m, _i32(buf)))
    return buf.reshape(m, n)


def synth_shape(seed: int, nlo: int, nhi: int, mchoices: Sequence[int]) -> Tuple[int, int]:
    mc = np.asarray(mchoices, dtype=np.intc)
    n = C.c_int()
    m = C.c_int()
    check(load().pbb_instance_synth_shape(seed, nlo, nhi, _ip(mc), len(mc), C.byref(n), C.byref(m)))
    return n.value, m.value


def read_vrf(path: str) -> np.ndarray:
    buf = np.zeros(64 * 1000, dtype=np.int32)
    n = C.c_int()
    m = C.c_int()
    check(load().pbb_instance_read_vrf(path.encode(), C.byref(n), C.byref(m), _i32(buf), buf.size))
    return buf[: n.value * m.value].reshape(m.value, n.value).copy()


def neh(ptm: np.ndarray, start: Optional[Sequence[int]] = None) -> Tuple[int, np.ndarray]:
    ptm = np.ascontiguousarray(ptm, dtype=np.int32)
    m, n = ptm.shape
    perm = np.arange(n, dtype=np.intc) if start is None else np.array(start, dtype=np.intc)
    if perm.shape != (n,):
        raise ValueError(f"NEH start order of length {perm.size}, expected {n}")
    cost = C.c_int()
    check(load().pbb_neh(_i32(ptm), n, m, _ip(perm), C.byref(cost)))
    return cost.value, perm


def makespan(ptm: np.ndarray, perm: Sequence[int]) -> int:
    ptm = np.ascontiguousarray(ptm, dtype=np.int32)
    m, n = ptm.shape
    p = np.ascontiguousarray(perm, np.intc)
    if p.shape != (n,):
        raise ValueError(f"permutation of length {p.size}, expected {n}")
    cost = C.c_int()
    check(load().pbb_makespan_host(_i32(ptm), n, m, _ip(p), C.byref(cost)))
    return cost.value


# ------------------------------------------------------------------ bounds

class Bounds:
    """bound_fsp_weak / bound_fsp_strong for one instance, batched on the GPU.

    Inputs are int32 CUDA tensors (torch): perms [B, n], l1 [B], l2 [B].
    """

    def __init__(self, ptm: np.ndarray):
        ptm = np.ascontiguousarray(ptm, dtype=np.int32)
        self.m, self.n = ptm.shape
        self._h = C.c_void_p()
        check(load().pbb_bounds_create(_i32(ptm), self.n, self.m, C.byref(self._h)))

    def close(self):
        if self._h:
            load().pbb_bounds_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream():
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _check_inputs(self, perms, l1=None, l2=None):
        import torch
        for t in (perms, l1, l2):
            if t is None:
                continue
            if not (t.is_cuda and t.dtype == torch.int32 and t.is_contiguous()):
                raise ValueError("bound inputs must be contiguous int32 CUDA tensors")
        if perms.dim() != 2 or perms.shape[1] != self.n:
            raise ValueError(f"perms must be [B, {self.n}]")
        B = perms.shape[0]
        if l1 is not None and (l1.numel() != B or l2.numel() != B):
            raise ValueError("l1/l2 must have one entry per schedule")
        return B

    def lb1_children(self, perms, l1, l2, both: bool = True):
        import torch
        B = self._check_inputs(perms, l1, l2)
        cb = torch.empty((B, self.n), dtype=torch.int32, device=perms.device)
        ce = torch.empty((B, self.n), dtype=torch.int32, device=perms.device) if both else None
        check(load().pbb_lb1_children_batch(self._h, B, perms.data_ptr(), l1.data_ptr(), l2.data_ptr(),
                                            int(both), cb.data_ptr(), ce.data_ptr() if both else None,
                                            self._stream()))
        return cb, ce

    def lb1_node(self, perms, l1, l2):
        import torch
        B = self._check_inputs(perms, l1, l2)
        out = torch.empty(B, dtype=torch.int32, device=perms.device)
        check(load().pbb_lb1_node_batch(self._h, B, perms.data_ptr(), l1.data_ptr(), l2.data_ptr(),
                                        out.data_ptr(), self._stream()))
        return out

    def lb2_node(self, perms, l1, l2, pairs: int = 0):
        import torch
        B = self._check_inputs(perms, l1, l2)
        out = torch.empty(B, dtype=torch.int32, device=perms.device)
        check(load().pbb_lb2_node_batch(self._h, B, perms.data_ptr(), l1.data_ptr(), l2.data_ptr(),
                                        pairs, out.data_ptr(), self._stream()))
        return out

    def makespan(self, perms):
        import torch
        B = self._check_inputs(perms)
        out = torch.empty(B, dtype=torch.int32, device=perms.device)
        check(load().pbb_makespan_batch(self._h, B, perms.data_ptr(), out.data_ptr(), self._stream()))
        return out


# ------------------------------------------------------------------ engine

@dataclass
class SolveStats:
    decomposed: int
    leaves: int
    simple_bounds: int
    johnson_bounds: int
    explorers: int
    active: int
    kernel_ms: float
    depth_histo: List[int] = field(default_factory=list)


DEFAULT_TIME_BUDGET_S = 3600.0


class GpuBB:
    """The IVM branch-and-bound engine on one GPU (gpubb)."""

    def __init__(self, ptm: np.ndarray, config: Optional[Config] = None, ub: int = INT_MAX,
                 ub_perm: Optional[Sequence[int]] = None, root: Optional[Sequence[int]] = None):
        ptm = np.ascontiguousarray(ptm, dtype=np.int32)
        self.m, self.n = ptm.shape
        self.ptm = ptm
        self.config = config if config is not None else default_config()
        up = None if ub_perm is None else _lib.sized(ub_perm, np.intc, self.n, "ub_perm")
        r = None if root is None else _lib.sized(root, np.intc, self.n, "root order")
        self._h = C.c_void_p()
        check(load().pbb_create(_i32(ptm), self.n, self.m, C.byref(self.config), int(ub),
                                None if up is None else _ip(up), C.byref(self._h)))
        if r is not None:
            check(load().pbb_set_root(self._h, _ip(r)))

    def close(self):
        if self._h:
            load().pbb_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # gpubb::initFullInterval
    def init_full_interval(self):
        check(load().pbb_init_full(self._h))

    # gpubb::initFromFac
    def init_from_fac(self, ids: Sequence[int], pos: np.ndarray, end: np.ndarray):
        ids = np.ascontiguousarray(ids, np.intc)
        pos = np.ascontiguousarray(pos, np.intc).reshape(-1)
        end = np.ascontiguousarray(end, np.intc).reshape(-1)
        nb = len(ids)
        if pos.size != nb * self.n or end.size != nb * self.n:
            raise ValueError("pos/end must hold nb*n factoradic digits")
        check(load().pbb_init_intervals(self._h, nb, _ip(ids), _ip(pos), _ip(end)))

    # gpubb::next.  time_budget 0 means none; the default bounds a call that
    # nobody interrupts to an hour.
    def next(self, node_budget: int = 0, time_budget: float = DEFAULT_TIME_BUDGET_S) -> bool:
        done = C.c_int(0)
        check(load().pbb_run(self._h, int(node_budget), float(time_budget), C.byref(done)))
        return bool(done.value)

    # gpubb::getIntervals
    def get_intervals(self):
        E = self.stats().explorers
        pos = np.zeros(E * self.n, np.intc)
        end = np.zeros(E * self.n, np.intc)
        ids = np.zeros(E, np.intc)
        nb = C.c_int()
        dec = C.c_uint64()
        lv = C.c_uint64()
        check(load().pbb_get_intervals(self._h, _ip(pos), _ip(end), _ip(ids), C.byref(nb), E,
                                       C.byref(dec), C.byref(lv)))
        k = nb.value
        return (ids[:k].copy(), pos[: k * self.n].reshape(k, self.n).copy(),
                end[: k * self.n].reshape(k, self.n).copy(), dec.value, lv.value)

    def interrupt_exploration(self):
        check(load().pbb_interrupt(self._h))

    def best(self) -> Tuple[int, np.ndarray]:
        cost = C.c_int()
        perm = np.zeros(self.n, np.intc)
        check(load().pbb_get_incumbent(self._h, C.byref(cost), _ip(perm)))
        return cost.value, perm

    def set_best(self, cost: int, perm: Optional[Sequence[int]] = None):
        p = None if perm is None else _lib.sized(perm, np.intc, self.n, "incumbent permutation")
        check(load().pbb_set_incumbent(self._h, int(cost), None if p is None else _ip(p)))

    def stats(self) -> SolveStats:
        st = Stats()
        check(load().pbb_get_stats(self._h, C.byref(st)))
        return SolveStats(st.decomposed, st.leaves, st.simple_bounds, st.johnson_bounds, st.explorers,
                          st.active, st.kernel_ms, list(st.depth_histo)[: self.n + 1])

    def get_deep_subproblems(self, max_count: int = 64) -> np.ndarray:
        """gpubb::getDeepSubproblem: current schedules of the active explorers,
        lowest makespan first -> array [k, n]."""
        buf = np.zeros(max(1, max_count) * self.n, dtype=np.intc)
        k = C.c_int()
        check(load().pbb_get_deep_subproblems(self._h, _ip(buf), max_count, C.byref(k)))
        return buf[: k.value * self.n].reshape(k.value, self.n).copy()

    def save_checkpoint(self, path: str):
        """works::save: active intervals, incumbent, root order, counters."""
        check(load().pbb_checkpoint_save(self._h, path.encode()))

    def load_checkpoint(self, path: str) -> dict:
        """works::init_infile: restore root order, incumbent and intervals."""
        c = Checkpoint()
        check(load().pbb_checkpoint_load(self._h, path.encode(), C.byref(c)))
        return {"explored": c.explored, "simple_bounds": c.simple_bounds, "best_cost": c.best_cost, "nb": c.nb}

    def solve(self, node_budget: int = 0, time_budget: float = DEFAULT_TIME_BUDGET_S) -> bool:
        """init_full_interval + next until done (or budget). Returns True when finished."""
        self.init_full_interval()
        return self.next(node_budget, time_budget)


class CpuBB:
    """Host explorers on the engine's work units (the multicore worker of
    distributed/src/worker_mc.cpp): same interval, incumbent and statistics
    calls as GpuBB."""

    def __init__(self, ptm: np.ndarray, config: Optional[Config] = None, threads: int = 0, ub: int = INT_MAX,
                 ub_perm: Optional[Sequence[int]] = None, root: Optional[Sequence[int]] = None):
        ptm = np.ascontiguousarray(ptm, dtype=np.int32)
        self.m, self.n = ptm.shape
        self.config = config if config is not None else default_config()
        up = None if ub_perm is None else _lib.sized(ub_perm, np.intc, self.n, "ub_perm")
        r = None if root is None else _lib.sized(root, np.intc, self.n, "root order")
        self._h = C.c_void_p()
        check(load().pbb_cpu_create(_i32(ptm), self.n, self.m, C.byref(self.config), int(threads), int(ub),
                                    None if up is None else _ip(up), C.byref(self._h)))
        if r is not None:
            check(load().pbb_cpu_set_root(self._h, _ip(r)))

    def close(self):
        if self._h:
            load().pbb_cpu_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init_full_interval(self):
        check(load().pbb_cpu_init_full(self._h))

    def init_from_fac(self, ids: Sequence[int], pos: np.ndarray, end: np.ndarray):
        ids = np.ascontiguousarray(ids, np.intc)
        pos = np.ascontiguousarray(pos, np.intc).reshape(-1)
        end = np.ascontiguousarray(end, np.intc).reshape(-1)
        if pos.size != len(ids) * self.n or end.size != len(ids) * self.n:
            raise ValueError("pos/end must hold nb*n factoradic digits")
        check(load().pbb_cpu_init_intervals(self._h, len(ids), _ip(ids), _ip(pos), _ip(end)))

    def next(self, node_budget: int = 0, time_budget: float = DEFAULT_TIME_BUDGET_S) -> bool:
        done = C.c_int(0)
        check(load().pbb_cpu_run(self._h, int(node_budget), float(time_budget), C.byref(done)))
        return bool(done.value)

    def get_intervals(self):
        T = self.stats().explorers
        pos = np.zeros(T * self.n, np.intc)
        end = np.zeros(T * self.n, np.intc)
        ids = np.zeros(T, np.intc)
        nb = C.c_int()
        dec = C.c_uint64()
        lv = C.c_uint64()
        check(load().pbb_cpu_get_intervals(self._h, _ip(pos), _ip(end), _ip(ids), C.byref(nb), T, C.byref(dec),
                                           C.byref(lv)))
        k = nb.value
        return (ids[:k].copy(), pos[: k * self.n].reshape(k, self.n).copy(),
                end[: k * self.n].reshape(k, self.n).copy(), dec.value, lv.value)

    def interrupt_exploration(self):
        check(load().pbb_cpu_interrupt(self._h))

    def best(self) -> Tuple[int, np.ndarray]:
        cost = C.c_int()
        perm = np.zeros(self.n, np.intc)
        check(load().pbb_cpu_get_incumbent(self._h, C.byref(cost), _ip(perm)))
        return cost.value, perm

    def set_best(self, cost: int, perm: Optional[Sequence[int]] = None):
        p = None if perm is None else _lib.sized(perm, np.intc, self.n, "incumbent permutation")
        check(load().pbb_cpu_set_incumbent(self._h, int(cost), None if p is None else _ip(p)))

    def stats(self) -> SolveStats:
        st = Stats()
        check(load().pbb_cpu_get_stats(self._h, C.byref(st)))
        return SolveStats(st.decomposed, st.leaves, st.simple_bounds, st.johnson_bounds, st.explorers,
                          st.active, st.kernel_ms, list(st.depth_histo)[: self.n + 1])

    def solve(self, node_budget: int = 0, time_budget: float = DEFAULT_TIME_BUDGET_S) -> bool:
        self.init_full_interval()
        return self.next(node_budget, time_budget)


def hybrid_run(gpu: GpuBB, cpu: CpuBB, period: float = 0.05, time_budget: float = 0.0):
    """The engine and the host explorers on one search (pbb_hybrid_run):
    returns (all_done, gpu_decomposed, cpu_decomposed)."""
    g = C.c_uint64()
    c = C.c_uint64()
    done = C.c_int()
    check(load().pbb_hybrid_run(gpu._h, cpu._h, float(period), float(time_budget), C.byref(g), C.byref(c),
                                C.byref(done)))
    return bool(done.value), g.value, c.value


# ------------------------------------------------------- work-unit codec

def fact_to_dec(digits: Sequence[int]) -> str:
    """Decimal rank of a factoradic digit vector (fact_work::VectToBigint)."""
    d = np.ascontiguousarray(digits, dtype=np.intc)
    buf = C.create_string_buffer(128)
    check(load().pbb_fact_to_dec(len(d), _ip(d), buf, 128))
    return buf.value.decode()


def dec_to_fact(n: int, dec) -> np.ndarray:
    """Factoradic digits of a decimal rank (fact_work::BigintToVect)."""
    d = np.zeros(n, dtype=np.intc)
    check(load().pbb_dec_to_fact(n, str(dec).encode(), _ip(d)))
    return d


def checkpoint_format(n: int, ids, pos, end, root_cost: int, root_perm, simple_bounds: int, best_cost: int,
                      best_perm, explored: int = 0, work_id: int = 0, max_intervals: int = 0) -> str:
    """works::save text of one work unit holding the given intervals."""
    c = Checkpoint()
    c.n, c.root_cost, c.simple_bounds, c.best_cost = n, root_cost, simple_bounds, best_cost
    c.work_id, c.max_intervals, c.explored = work_id, max_intervals or max(1, len(ids)), explored
    for i in range(n):
        c.root_perm[i] = int(root_perm[i])
        c.best_perm[i] = int(best_perm[i])
    ids = np.ascontiguousarray(ids, dtype=np.intc).reshape(-1)
    pos = _lib.sized(pos, np.intc, ids.size * n, "pos digits").reshape(-1)
    end = _lib.sized(end, np.intc, ids.size * n, "end digits").reshape(-1)
    c.nb = len(ids)
    ln = C.c_int()
    check(load().pbb_checkpoint_format(C.byref(c), _ip(ids), _ip(pos), _ip(end), None, 0, C.byref(ln)))
    buf = C.create_string_buffer(ln.value + 1)
    check(load().pbb_checkpoint_format(C.byref(c), _ip(ids), _ip(pos), _ip(end), buf, ln.value + 1, C.byref(ln)))
    return buf.value.decode()


def checkpoint_parse(text: str, n: int, max_intervals: int = 1 << 16):
    """-> (header dict, ids, pos[nb, n], end[nb, n]) of a works::save text."""
    c = Checkpoint()
    ids = np.zeros(max_intervals, dtype=np.intc)
    pos = np.zeros(max_intervals * n, dtype=np.intc)
    end = np.zeros(max_intervals * n, dtype=np.intc)
    check(load().pbb_checkpoint_parse(text.encode(), n, C.byref(c), _ip(ids), _ip(pos), _ip(end), max_intervals))
    hdr = {"root_cost": c.root_cost, "root_perm": [c.root_perm[i] for i in range(n)],
           "simple_bounds": c.simple_bounds, "best_cost": c.best_cost,
           "best_perm": [c.best_perm[i] for i in range(n)], "work_id": c.work_id,
           "max_intervals": c.max_intervals, "explored": c.explored, "nb": c.nb}
    k = c.nb
    return hdr, ids[:k].copy(), pos[:k * n].reshape(k, n).copy(), end[:k * n].reshape(k, n).copy()


def explorers_for(n: int, m: int, config: Optional[Config] = None) -> int:
    cfg = config if config is not None else default_config()
    out = C.c_int()
    check(load().pbb_query_explorers(n, m, C.byref(cfg), C.byref(out)))
    return out.value
