"""ctypes binding of ``libpbb_engine.so`` (the C ABI of include/pbb_engine.h).

The library is the only compute path: there is no Python or CPU fallback.
If it is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbb_engine.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pbb_engine.h")

PBB_MAX_JOBS = 63
PBB_MAX_MACHINES = 20

ERRORS = {-1: "PBB_E_ARG", -2: "PBB_E_HIP", -3: "PBB_E_STATE", -4: "PBB_E_CAPACITY", -5: "PBB_E_NODEVICE",
          -6: "PBB_E_BUSY"}
PBB_E_BUSY = -6
RES_PENDING, RES_SOLVED, RES_ABORTED, RES_INVALID = 0, 1, 2, 3


class PbbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    """pbb_config: the [bb]/[gpu]/[initial]/[time] keys of gpuconfig.ini."""

    _fields_ = [
        ("bound_mode", C.c_int), ("branching", C.c_int), ("refine", C.c_int),
        ("johnson_pairs", C.c_int), ("early_stop_johnson", C.c_int), ("find_all", C.c_int),
        ("single_node", C.c_int), ("sorted_dfs", C.c_int), ("nb_ivms", C.c_int),
        ("steal_margin", C.c_int), ("init_mode", C.c_int), ("init_solution", C.c_int),
        ("checkpoint_s", C.c_int), ("balance_s", C.c_int), ("ramp_nodes", C.c_int),
        ("team_nodes", C.c_int), ("team_waves", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [
        ("decomposed", C.c_uint64), ("leaves", C.c_uint64), ("simple_bounds", C.c_uint64),
        ("johnson_bounds", C.c_uint64), ("explorers", C.c_int), ("active", C.c_int),
        ("kernel_ms", C.c_double), ("depth_histo", C.c_uint64 * (PBB_MAX_JOBS + 2)),
        ("refine_histo", C.c_uint64 * (PBB_MAX_JOBS + 2)),
    ]


class PoolParams(C.Structure):
    """pbb_pool_params: shape and tuning of a multi-instance pool."""

    _fields_ = [
        ("max_jobs", C.c_int), ("machines", C.c_int), ("slots", C.c_int),
        ("waves_per_group", C.c_int), ("explorers_per_instance", C.c_int),
        ("steal_margin", C.c_int), ("groups", C.c_int), ("hbm_budget_gb", C.c_double),
        ("request_budget_s", C.c_double), ("ig_iterations", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Result(C.Structure):
    """pbb_result: one solved (or aborted) instance of a pool."""

    _fields_ = [
        ("tag", C.c_int), ("status", C.c_int), ("best", C.c_int), ("improved", C.c_int),
        ("decomposed", C.c_uint64), ("leaves", C.c_uint64), ("simple_bounds", C.c_uint64),
        ("johnson_bounds", C.c_uint64), ("admit_ms", C.c_double), ("done_ms", C.c_double),
        ("perm", C.c_int * PBB_MAX_JOBS),
    ]


class Checkpoint(C.Structure):
    """pbb_checkpoint: the header fields of a works::save checkpoint."""

    _fields_ = [
        ("n", C.c_int), ("root_cost", C.c_int), ("root_perm", C.c_int * PBB_MAX_JOBS),
        ("simple_bounds", C.c_uint64), ("best_cost", C.c_int), ("best_perm", C.c_int * PBB_MAX_JOBS),
        ("work_id", C.c_int), ("max_intervals", C.c_int), ("explored", C.c_uint64), ("nb", C.c_int),
    ]


IP = C.POINTER(C.c_int)
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)
VP = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "pbb_config_default": (C.c_int, [C.POINTER(Config)]),
    "pbb_config_load_ini": (C.c_int, [C.c_char_p, C.POINTER(Config), C.c_char_p, C.c_char_p]),
    "pbb_instance_taillard": (C.c_int, [C.c_int, IP, IP, I32P]),
    "pbb_instance_synth": (C.c_int, [C.c_uint64, C.c_int, C.c_int, I32P]),
    "pbb_instance_synth_shape": (C.c_int, [C.c_uint64, C.c_int, C.c_int, IP, C.c_int, IP, IP]),
    "pbb_instance_read_vrf": (C.c_int, [C.c_char_p, IP, IP, I32P, C.c_int]),
    "pbb_neh": (C.c_int, [I32P, C.c_int, C.c_int, IP, IP]),
    "pbb_makespan_host": (C.c_int, [I32P, C.c_int, C.c_int, IP, IP]),
    "pbb_bounds_create": (C.c_int, [I32P, C.c_int, C.c_int, C.POINTER(VP)]),
    "pbb_bounds_destroy": (C.c_int, [VP]),
    "pbb_lb1_children_batch": (C.c_int, [VP, C.c_int, VP, VP, VP, C.c_int, VP, VP, VP]),
    "pbb_lb1_node_batch": (C.c_int, [VP, C.c_int, VP, VP, VP, VP, VP]),
    "pbb_lb2_node_batch": (C.c_int, [VP, C.c_int, VP, VP, VP, C.c_int, VP, VP]),
    "pbb_makespan_batch": (C.c_int, [VP, C.c_int, VP, VP, VP]),
    "pbb_create": (C.c_int, [I32P, C.c_int, C.c_int, C.POINTER(Config), C.c_int, IP, C.POINTER(VP)]),
    "pbb_destroy": (C.c_int, [VP]),
    "pbb_set_root": (C.c_int, [VP, IP]),
    "pbb_init_full": (C.c_int, [VP]),
    "pbb_init_intervals": (C.c_int, [VP, C.c_int, IP, IP, IP]),
    "pbb_run": (C.c_int, [VP, C.c_uint64, C.c_double, IP]),
    "pbb_get_intervals": (C.c_int, [VP, IP, IP, IP, IP, C.c_int, U64P, U64P]),
    "pbb_get_incumbent": (C.c_int, [VP, IP, IP]),
    "pbb_set_incumbent": (C.c_int, [VP, C.c_int, IP]),
    "pbb_get_stats": (C.c_int, [VP, C.POINTER(Stats)]),
    "pbb_interrupt": (C.c_int, [VP]),
    "pbb_get_deep_subproblems": (C.c_int, [VP, IP, C.c_int, IP]),
    "pbb_query_explorers": (C.c_int, [C.c_int, C.c_int, C.POINTER(Config), IP]),
    "pbb_pool_params_default": (C.c_int, [C.POINTER(PoolParams)]),
    "pbb_pool_create": (C.c_int, [C.POINTER(Config), C.POINTER(PoolParams), C.POINTER(VP)]),
    "pbb_pool_destroy": (C.c_int, [VP]),
    "pbb_pool_info": (C.c_int, [VP, IP, IP, IP]),
    "pbb_pool_retune": (C.c_int, [VP, C.c_int, C.c_int]),
    "pbb_pool_solve": (C.c_int, [VP, C.c_int, I32P, IP, IP, IP, IP, C.c_double, C.POINTER(Result),
                                 C.POINTER(C.c_double)]),
    "pbb_neh_device": (C.c_int, [C.c_int, I32P, IP, IP, IP, IP]),
    "pbb_neh_device_noalloc": (C.c_int, [C.c_int, VP, VP, VP, C.c_int, VP, VP, VP]),
    "pbb_ig_device": (C.c_int, [C.c_int, I32P, IP, IP, C.c_int, C.c_int, C.c_uint64, IP, IP, IP, IP]),
    "pbb_serve_start": (C.c_int, [VP, C.c_double]),
    "pbb_serve_submit": (C.c_int, [VP, I32P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "pbb_serve_poll": (C.c_int, [VP, C.POINTER(Result), C.c_int, IP]),
    "pbb_serve_stop": (C.c_int, [VP]),
    "pbb_serve_footprint": (C.c_int, [VP, C.c_int, U64P]),
    "pbb_serve_replay": (C.c_int, [VP, C.c_int, I32P, IP, IP, IP, C.POINTER(C.c_double), C.c_double,
                                   C.POINTER(Result), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pbb_serve_replay_routed": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, I32P, IP, IP, IP,
                                          C.POINTER(C.c_double), C.c_double, C.POINTER(Result),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pbb_plain_replay": (C.c_int, [C.POINTER(Config), C.c_int, I32P, IP, IP, IP, C.POINTER(C.c_double),
                                   C.c_double, C.c_double, C.POINTER(Result), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "pbb_pool_batch_histograms": (C.c_int, [VP, C.c_int, U64P, U64P]),
    "pbb_pool_autotune": (C.c_int, [C.POINTER(Config), C.POINTER(PoolParams), C.c_int, I32P, IP, IP, IP,
                                    C.POINTER(PoolParams), C.POINTER(C.c_double), C.c_int, IP]),
    "pbb_cpu_create": (C.c_int, [I32P, C.c_int, C.c_int, C.POINTER(Config), C.c_int, C.c_int, IP, C.POINTER(VP)]),
    "pbb_cpu_destroy": (C.c_int, [VP]),
    "pbb_cpu_set_root": (C.c_int, [VP, IP]),
    "pbb_cpu_init_intervals": (C.c_int, [VP, C.c_int, IP, IP, IP]),
    "pbb_cpu_init_full": (C.c_int, [VP]),
    "pbb_cpu_run": (C.c_int, [VP, C.c_uint64, C.c_double, IP]),
    "pbb_cpu_get_intervals": (C.c_int, [VP, IP, IP, IP, IP, C.c_int, U64P, U64P]),
    "pbb_cpu_get_incumbent": (C.c_int, [VP, IP, IP]),
    "pbb_cpu_set_incumbent": (C.c_int, [VP, C.c_int, IP]),
    "pbb_cpu_get_stats": (C.c_int, [VP, C.POINTER(Stats)]),
    "pbb_cpu_interrupt": (C.c_int, [VP]),
    "pbb_hybrid_run": (C.c_int, [VP, VP, C.c_double, C.c_double, U64P, U64P, IP]),
    "pbb_fact_to_dec": (C.c_int, [C.c_int, IP, C.c_char_p, C.c_int]),
    "pbb_dec_to_fact": (C.c_int, [C.c_int, C.c_char_p, IP]),
    "pbb_checkpoint_format": (C.c_int, [C.POINTER(Checkpoint), IP, IP, IP, C.c_char_p, C.c_int, IP]),
    "pbb_checkpoint_parse": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(Checkpoint), IP, IP, IP, C.c_int]),
    "pbb_checkpoint_save": (C.c_int, [VP, C.c_char_p]),
    "pbb_checkpoint_load": (C.c_int, [VP, C.c_char_p, C.POINTER(Checkpoint)]),
    "pbb_last_error": (C.c_char_p, []),
    "pbb_version": (C.c_char_p, []),
}

_lib = None


def load():
    """Load libpbb_engine.so (raises if it has not been built).

    torch is imported first: it carries its own HIP runtime
    (libamdhip64.so, SONAME libamdhip64.so.7), and loading it before the
    engine makes the engine's libamdhip64.so.7 dependency resolve to that
    same runtime instead of a second copy from /opt/rocm, so device
    pointers and streams are shared with torch.
    """
    global _lib
    if _lib is None:
        import torch  # noqa: F401  (HIP runtime shared with the engine)
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m pbbperm_amd.build` "
                "(there is no fallback path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def sized(values, dtype, size: int, what: str):
    """values as a contiguous array of exactly `size` elements: the C side
    reads that many, so a shorter array is refused here."""
    import numpy as np
    a = np.ascontiguousarray(values, dtype=dtype)
    if a.size != size:
        raise ValueError(f"{what}: {a.size} values, expected {size}")
    return a


def check(rc: int) -> int:
    if rc != 0:
        raise PbbError(rc, load().pbb_last_error().decode())
    return rc
