"""Multi-instance pool: batched solves, serving loop, stream replay, autotuner.

The reference solves one instance per process (its strong-bound tables are
process-global, bounds/src/bound_fsp_strong.cpp:5-10).  The pool keeps many
instances resident in one launch of the gfx950 pool kernel, each solved with
the semantics of :class:`pbbperm_amd.GpuBB` (``initFullInterval`` + ``next``
until all explorers are idle, identity root order).  Every call goes through
the C ABI of ``libpbb_engine.so``.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import Config, PoolParams, Result, check, load
from .engine import default_config, synth, synth_shape

STATUS = {0: "pending", 1: "solved", 2: "aborted", 3: "invalid"}


@dataclass
class PoolResult:
    tag: int
    status: str
    best: int
    improved: bool
    decomposed: int
    leaves: int
    simple_bounds: int
    johnson_bounds: int
    admit_ms: float
    done_ms: float
    perm: List[int] = field(default_factory=list)


def _result(r: Result, n: int) -> PoolResult:
    return PoolResult(tag=r.tag, status=STATUS.get(r.status, str(r.status)), best=r.best,
                      improved=bool(r.improved), decomposed=r.decomposed, leaves=r.leaves,
                      simple_bounds=r.simple_bounds, johnson_bounds=r.johnson_bounds,
                      admit_ms=r.admit_ms, done_ms=r.done_ms,
                      perm=[r.perm[i] for i in range(n)] if r.improved else [])


def _pack(ptms: Sequence[np.ndarray]):
    ms = np.array([p.shape[0] for p in ptms], dtype=np.int32)
    ns = np.array([p.shape[1] for p in ptms], dtype=np.int32)
    flat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(p, dtype=np.int32).ravel()
                                                for p in ptms]) if len(ptms) else np.zeros(1, np.int32))
    return flat, ns, ms


def _ip(a):
    return a.ctypes.data_as(_lib.IP)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params(**overrides) -> PoolParams:
    prm = PoolParams()
    check(load().pbb_pool_params_default(C.byref(prm)))
    for k, v in overrides.items():
        if not hasattr(prm, k):
            raise KeyError(k)
        setattr(prm, k, float(v) if k in ("hbm_budget_gb", "request_budget_s") else int(v))
    return prm


class Pool:
    """A resident pool of instance slots on one GPU."""

    def __init__(self, config: Optional[Config] = None, params: Optional[PoolParams] = None, **kw):
        self._lib = load()
        self.config = config if config is not None else default_config()
        self.params = params if params is not None else default_params(**kw)
        h = C.c_void_p()
        check(self._lib.pbb_pool_create(C.byref(self.config), C.byref(self.params), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbb_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def retune(self, explorers_per_instance: int, steal_margin: int):
        check(self._lib.pbb_pool_retune(self._h, int(explorers_per_instance), int(steal_margin)))

    def info(self) -> Dict[str, int]:
        e, g, l = C.c_int(), C.c_int(), C.c_int()
        check(self._lib.pbb_pool_info(self._h, C.byref(e), C.byref(g), C.byref(l)))
        return {"explorers": e.value, "groups": g.value, "lds_bytes": l.value}

    def solve(self, ptms: Sequence[np.ndarray], ubs: Sequence[int], tags: Optional[Sequence[int]] = None,
              time_budget: float = 0.0) -> Tuple[List[PoolResult], float]:
        """Solve len(ptms) <= slots instances in one graph launch."""
        flat, ns, ms = _pack(ptms)
        ub = _lib.sized(ubs, np.int32, len(ptms), "ubs")
        tg = None if tags is None else _lib.sized(tags, np.int32, len(ptms), "tags")
        out = (Result * max(1, len(ptms)))()
        kms = C.c_double()
        check(self._lib.pbb_pool_solve(self._h, len(ptms), flat.ctypes.data_as(_lib.I32P), _ip(ns), _ip(ms),
                                       _ip(ub), None if tg is None else _ip(tg), float(time_budget), out,
                                       C.byref(kms)))
        return [_result(out[i], int(ns[i])) for i in range(len(ptms))], kms.value

    def batch_histograms(self, count: int):
        """Per-instance depth histograms of the last solve() batch:
        (decomposed [count, 64], refined [count, 64]) as uint64 arrays."""
        dec = np.zeros(max(1, count) * 64, dtype=np.uint64)
        ref = np.zeros(max(1, count) * 64, dtype=np.uint64)
        u64 = C.POINTER(C.c_uint64)
        check(self._lib.pbb_pool_batch_histograms(self._h, int(count), dec.ctypes.data_as(u64),
                                                  ref.ctypes.data_as(u64)))
        return dec[:count * 64].reshape(count, 64), ref[:count * 64].reshape(count, 64)

    # ------------------------------------------------------------ serving
    def serve_start(self, time_budget: float):
        check(self._lib.pbb_serve_start(self._h, float(time_budget)))

    def submit(self, ptm: np.ndarray, ub: int, tag: int) -> bool:
        """False when the request cannot be admitted now (no slot / HBM budget)."""
        p = np.ascontiguousarray(ptm, dtype=np.int32)
        rc = self._lib.pbb_serve_submit(self._h, p.ctypes.data_as(_lib.I32P), p.shape[1], p.shape[0],
                                        int(ub), int(tag))
        if rc == _lib.PBB_E_BUSY:
            return False
        check(rc)
        return True

    def poll(self, max_results: int = 1024) -> List[Tuple[int, Result]]:
        out = (Result * max_results)()
        k = C.c_int()
        check(self._lib.pbb_serve_poll(self._h, out, max_results, C.byref(k)))
        res = []
        for i in range(k.value):
            r = Result()
            C.memmove(C.byref(r), C.byref(out[i]), C.sizeof(Result))
            res.append(r)
        return res

    def serve_stop(self):
        check(self._lib.pbb_serve_stop(self._h))

    def footprint(self, n: int) -> int:
        b = C.c_uint64()
        check(self._lib.pbb_serve_footprint(self._h, int(n), C.byref(b)))
        return b.value

    def replay(self, ptms: Sequence[np.ndarray], ubs: Sequence[int], arrivals: Sequence[float],
               time_budget: float = 0.0):
        """Serve an arrival stream; returns (results, latency_ms, wall_s)."""
        flat, ns, ms = _pack(ptms)
        ub = _lib.sized(ubs, np.int32, len(ptms), "ubs")
        arr = _lib.sized(arrivals, np.float64, len(ptms), "arrivals")
        out = (Result * max(1, len(ptms)))()
        lat = np.zeros(max(1, len(ptms)), dtype=np.float64)
        wall = C.c_double()
        check(self._lib.pbb_serve_replay(self._h, len(ptms), flat.ctypes.data_as(_lib.I32P), _ip(ns), _ip(ms),
                                         _ip(ub), _dp(arr), float(time_budget), out, _dp(lat), C.byref(wall)))
        return [_result(out[i], int(ns[i])) for i in range(len(ptms))], lat[:len(ptms)], wall.value


UB_NEH = -1  # a request's ub: compute NEH on the device at admission


def neh_device(ptms: Sequence[np.ndarray]):
    """NEH of many instances on the device -> (orders list, costs array)."""
    flat, ns, ms = _pack(ptms)
    perms = np.zeros(64 * max(1, len(ptms)), dtype=np.intc)
    costs = np.zeros(max(1, len(ptms)), dtype=np.intc)
    check(load().pbb_neh_device(len(ptms), flat.ctypes.data_as(_lib.I32P), _ip(ns), _ip(ms), _ip(perms),
                                _ip(costs)))
    return [perms[64 * i:64 * i + int(ns[i])].copy() for i in range(len(ptms))], costs[:len(ptms)].copy()


def ig_device(ptms: Sequence[np.ndarray], chains: int = 64, iterations: int = 100, seed: int = 0,
              per_chain: bool = False):
    """Iterated greedy (IG.cpp runIG) on the device, `chains` seeded chains
    per instance -> (best orders list, best costs array); with per_chain also
    (chain costs [count, chains], chain orders [count, chains, n] lists)."""
    flat, ns, ms = _pack(ptms)
    k = max(1, len(ptms))
    perms = np.zeros(64 * k, dtype=np.intc)
    costs = np.zeros(k, dtype=np.intc)
    cc = np.zeros(k * chains, dtype=np.intc) if per_chain else None
    cp = np.zeros(k * chains * 64, dtype=np.intc) if per_chain else None
    check(load().pbb_ig_device(len(ptms), flat.ctypes.data_as(_lib.I32P), _ip(ns), _ip(ms), int(chains),
                               int(iterations), C.c_uint64(seed), _ip(perms), _ip(costs),
                               None if cc is None else _ip(cc), None if cp is None else _ip(cp)))
    best = [perms[64 * i:64 * i + int(ns[i])].copy() for i in range(len(ptms))]
    if not per_chain:
        return best, costs[:len(ptms)].copy()
    chain_perms = [[cp[(i * chains + c) * 64:(i * chains + c) * 64 + int(ns[i])].copy() for c in range(chains)]
                   for i in range(len(ptms))]
    return best, costs[:len(ptms)].copy(), cc[:len(ptms) * chains].reshape(len(ptms), chains), chain_perms


def replay_routed(pools: Sequence["Pool"], ptms: Sequence[np.ndarray], ubs: Sequence[int],
                  arrivals: Sequence[float], time_budget: float = 0.0):
    """Serve one arrival stream through several pools resident together: each
    request goes to the pool with the fewest machines that holds it.  The
    pools' serving kernels start in decreasing order of machines; size each
    pool's groups so the later ones stay resident beside the earlier ones.
    Returns (results, latency_ms, wall_s) in request order."""
    flat, ns, ms = _pack(ptms)
    ub = _lib.sized(ubs, np.int32, len(ptms), "ubs")
    arr = _lib.sized(arrivals, np.float64, len(ptms), "arrivals")
    hs = (C.c_void_p * len(pools))(*[p._h for p in pools])
    out = (Result * max(1, len(ptms)))()
    lat = np.zeros(max(1, len(ptms)), dtype=np.float64)
    wall = C.c_double()
    check(load().pbb_serve_replay_routed(hs, len(pools), len(ptms), flat.ctypes.data_as(_lib.I32P), _ip(ns),
                                         _ip(ms), _ip(ub), _dp(arr), float(time_budget), out, _dp(lat),
                                         C.byref(wall)))
    return [_result(out[i], int(ns[i])) for i in range(len(ptms))], lat[:len(ptms)], wall.value


def plain_replay(ptms: Sequence[np.ndarray], ubs: Sequence[int], arrivals: Sequence[float],
                 config: Optional[Config] = None, time_budget: float = 0.0, request_budget: float = 0.0):
    """The same stream on the one-instance engine, one launch per instance, FCFS
    (request_budget: per-instance time cap, as the pool's request_budget_s)."""
    cfg = config if config is not None else default_config()
    flat, ns, ms = _pack(ptms)
    ub = _lib.sized(ubs, np.int32, len(ptms), "ubs")
    arr = _lib.sized(arrivals, np.float64, len(ptms), "arrivals")
    out = (Result * max(1, len(ptms)))()
    lat = np.zeros(max(1, len(ptms)), dtype=np.float64)
    wall = C.c_double()
    check(load().pbb_plain_replay(C.byref(cfg), len(ptms), flat.ctypes.data_as(_lib.I32P), _ip(ns), _ip(ms),
                                  _ip(ub), _dp(arr), float(time_budget), float(request_budget), out, _dp(lat),
                                  C.byref(wall)))
    return [_result(out[i], int(ns[i])) for i in range(len(ptms))], lat[:len(ptms)], wall.value


# searched most-promising first, so the time cap tightens early
GRID_W = (4, 2, 1)
GRID_EPI = (256, 1024, 3072, 64)
GRID_MARGIN = (2, 4, 8)


OBJECTIVES = ("makespan", "replay_mean")


def autotune(ptms: Sequence[np.ndarray], ubs: Sequence[int], config: Optional[Config] = None,
             base: Optional[PoolParams] = None, time_budget: float = 120.0, progress=None,
             objective: str = "makespan", arrivals: Optional[Sequence[float]] = None):
    """Pick waves per workgroup (block size), explorers per instance and the
    steal margin for the given instances; every candidate of GRID_W x
    GRID_EPI x GRID_MARGIN is scored, each run capped at twice the best run so
    far plus 0.5 s (a capped run fails).
    objective "makespan": the instances solved as one batch, best of two
    graph launches after a warm one.  "replay_mean": the instances replayed
    as a request stream arriving at `arrivals` (s) through the resident
    serving loop, scored by the mean request latency (one replay per
    candidate, after one warm replay per block size).
    Returns (params, trials) with trials rows (W, explorers_per_instance,
    margin, score ms).  progress(row) is called after each trial."""
    if objective not in OBJECTIVES:
        raise ValueError(f"objective must be one of {OBJECTIVES}")
    if objective == "replay_mean" and (arrivals is None or len(arrivals) != len(ptms)):
        raise ValueError("replay_mean needs one arrival time per instance")
    cfg = config if config is not None else default_config()
    b = base if base is not None else default_params()
    trials = []
    best_run = float("inf")
    for W in GRID_W:
        prm = PoolParams()
        C.memmove(C.byref(prm), C.byref(b), C.sizeof(PoolParams))
        prm.waves_per_group = W
        prm.explorers_per_instance = max(GRID_EPI)
        prm.slots = max(b.slots, len(ptms))
        with Pool(cfg, params=prm) as pool:
            E = pool.info()["explorers"]
            if objective == "replay_mean":
                k = min(16, len(ptms))
                pool.replay(ptms[:k], ubs[:k], np.zeros(k), time_budget=time_budget)
            for epi in GRID_EPI:
                if epi < W or epi > E:
                    continue
                for mg in GRID_MARGIN:
                    pool.retune(epi, mg)
                    # a run may not last longer than twice the best so far (+0.5 s)
                    cap = min(time_budget, 2.0 * best_run / 1e3 + 0.5) if np.isfinite(best_run) else time_budget
                    ms = float("inf")
                    if objective == "makespan":
                        pool.solve(ptms, ubs, time_budget=cap)
                        for _ in range(2):  # best of two timed launches (the search order varies)
                            res, t = pool.solve(ptms, ubs, time_budget=cap)
                            if all(r.status == "solved" for r in res):
                                best_run = min(best_run, t)
                                ms = min(ms, t)
                    else:
                        # every request served (solved, or stopped at the
                        # request budget when base sets one) within the cap
                        res, lat, wall = pool.replay(ptms, ubs, arrivals, time_budget=cap)
                        stopped_ok = b.request_budget_s > 0
                        if all(l >= 0 and (r.status == "solved" or (stopped_ok and r.status == "aborted"))
                               for r, l in zip(res, lat)):
                            best_run = min(best_run, wall * 1e3)
                            ms = float(np.mean(lat))
                    row = (W, epi, mg, ms)
                    trials.append(row)
                    if progress:
                        progress(row)
    W, epi, mg, _ = min(trials, key=lambda r: r[3])
    best = PoolParams()
    C.memmove(C.byref(best), C.byref(b), C.sizeof(PoolParams))
    best.waves_per_group, best.explorers_per_instance, best.steal_margin = W, epi, mg
    return best, np.array(trials, dtype=np.float64)


def autotune_native(ptms: Sequence[np.ndarray], ubs: Sequence[int], config: Optional[Config] = None,
                    base: Optional[PoolParams] = None, max_trials: int = 64):
    """The same search run inside the library (pbb_pool_autotune)."""
    cfg = config if config is not None else default_config()
    b = base if base is not None else default_params()
    flat, ns, ms = _pack(ptms)
    ub = _lib.sized(ubs, np.int32, len(ptms), "ubs")
    best = PoolParams()
    trials = np.zeros(4 * max_trials, dtype=np.float64)
    nt = C.c_int()
    check(load().pbb_pool_autotune(C.byref(cfg), C.byref(b), len(ptms), flat.ctypes.data_as(_lib.I32P),
                                   _ip(ns), _ip(ms), _ip(ub), C.byref(best), _dp(trials), max_trials,
                                   C.byref(nt)))
    k = min(nt.value, max_trials)
    return best, trials[:4 * k].reshape(k, 4)


# ----------------------------------------------------------- workloads

def instance_stream(seeds: Sequence[int], nlo: int, nhi: int, mchoices: Sequence[int]) -> List[np.ndarray]:
    """Seeded instances of random shape: (n, m) from the shape stream of each
    seed, times from its PTM stream (the project's SplitMix64 generator)."""
    out = []
    for s in seeds:
        n, m = synth_shape(int(s), nlo, nhi, mchoices)
        out.append(synth(int(s), n, m))
    return out


def batch_stream(count: int = 512, seed0: int = 100) -> List[np.ndarray]:
    """The batched workload: seeds seed0.., n ~ U[10,22], m ~ U{5,10}."""
    return instance_stream(range(seed0, seed0 + count), 10, 22, (5, 10))


def serving_stream(count: int, seed0: int = 10000) -> List[np.ndarray]:
    """The serving workload: request i has seed seed0+i, n ~ U[12,30] and
    m ~ U{5,10,20} from its shape stream, times from its PTM stream."""
    return instance_stream(range(seed0, seed0 + count), 12, 30, (5, 10, 20))


def poisson_arrivals(count: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times (s) of a Poisson process of the given rate (1/s)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=count)
    t = np.cumsum(gaps)
    return t - t[0]


# ------------------------------------------------------ named workloads
# Bounding / branching of both multi-instance workloads: LB1 for every child,
# the Johnson LB2 over all machine pairs to refine the children kept in the
# chosen direction (boundingMode 2, the commented-out refinement of
# multicore/src/ivm_bound.cpp:171), MinMin branching (the reference's default
# adaptiveBranchingMode 3).  Both run the pool's iterated-greedy helper wave
# beside each instance's search (ig_iterations rounds at most; it stops when
# the search does).  Measured on the 512-instance batch (tuned 4 waves, 1024
# explorers per instance, margin 4): 180-194 ms and 66-69 M nodes without
# it, 115-128 ms / 35 M with 20 rounds, 104-119 ms / 28-30 M with 100 (300:
# 105-128 ms, 1000: 112-122 ms, margin 8); on the
# 600-request serving stream (single pool): p50 106 / 82 / 69 / 69 / 60 ms
# and 133 / 117 / 109 / 108 / 106 late at 0 / 20 / 100 / 300 / 1000 rounds.
_LB12 = dict(bound_mode=2, branching=3, refine=1, johnson_pairs=0, early_stop_johnson=0)

WORKLOADS = {
    "batched": dict(config=_LB12, max_jobs=22, machines=10, slots=512, ig_iterations=100,
                    tune=lambda: batch_stream(512, 1100), objective="makespan"),
    "serving": dict(config=_LB12, max_jobs=30, machines=20, slots=256, ig_iterations=1000,
                    tune=lambda: serving_stream(120, 30000), objective="replay_mean", rate=20.0,
                    request_budget_s=2.0),
}

# The serving stream split over two pools resident together (replay_routed):
# 5- and 10-machine requests on a 10-machine pool, 20-machine requests on a
# 20-machine pool, both for n <= 30.  The 20-machine pool starts first (its
# waves need the most registers).  Measured on the 600-request serving stream
# (20 req/s, 2 s budget per request, the stored serving tuning), groups of
# the 10-machine + 20-machine pools: 512+192 gave 165 late, p50 5.7 ms, p99
# 1.86 s; 384+256 156, 4.7 ms, 1.87 s; 256+320 145-149, 4.1-4.4 ms,
# 1.52-1.83 s; 192+352 146, 3.9 ms, 1.76 s; 128+384 137, 4.0 ms, 1.46 s;
# 128+416 133, 4.0 ms, 1.89 s.  Most of the work is in 20-machine requests,
# and a 10-machine pool of 128 groups still serves the rest within a few ms.
# Again with the IG helper wave (1000 rounds): 128+384 111-114 late, p50
# 4.1-4.2 ms; 256+320 133, 4.2 ms; 64+416 108, 7.5 ms.
SERVING_ROUTED = (dict(machines=10, max_jobs=30, groups=128),
                  dict(machines=20, max_jobs=30, groups=384))


def serving_routed_params(base: PoolParams) -> List[PoolParams]:
    """PoolParams of the routed serving pools, with base's tuned waves per
    group, explorers per instance and steal margin."""
    keep = dict(slots=base.slots, waves_per_group=base.waves_per_group,
                explorers_per_instance=base.explorers_per_instance, steal_margin=base.steal_margin,
                request_budget_s=base.request_budget_s, ig_iterations=base.ig_iterations)
    return [default_params(**keep, **shape) for shape in SERVING_ROUTED]


TUNED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")


def workload_config(name: str, **overrides) -> Config:
    kw = dict(WORKLOADS[name]["config"])
    kw.update(overrides)
    return default_config(**kw)


def workload_params(name: str, **overrides) -> PoolParams:
    w = WORKLOADS[name]
    kw = dict(max_jobs=w["max_jobs"], machines=w["machines"], slots=w["slots"], ig_iterations=w["ig_iterations"])
    kw.update(overrides)
    return default_params(**kw)


def tune_workload(name: str, progress=None):
    """Run the autotuner on the workload's tuning sample (disjoint seeds from
    the measured stream); returns (params, trials)."""
    from .engine import neh
    w = WORKLOADS[name]
    sample = w["tune"]()
    ubs = [neh(p)[0] for p in sample]
    arrivals = None
    base = workload_params(name)
    if w["objective"] == "replay_mean":
        # the serving leg's Poisson rate and per-request budget
        arrivals = poisson_arrivals(len(sample), w["rate"], seed=11)
        base.request_budget_s = w["request_budget_s"]
    return autotune(sample, ubs, config=workload_config(name), base=base, progress=progress,
                    objective=w["objective"], arrivals=arrivals)


def tuned_params(name: str, path: str = TUNED_PATH) -> Optional[PoolParams]:
    """The autotuner's stored choice for a workload, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f).get(name)
    if not d:
        return None
    return workload_params(name, waves_per_group=d["waves_per_group"],
                           explorers_per_instance=d["explorers_per_instance"], steal_margin=d["steal_margin"])
