"""pbbperm_amd -- MI355X-native IVM branch-and-bound for the permutation
flow-shop problem (makespan), behind the reference's bound-library, GPU
worker and factoradic work-unit interfaces.

The compute path is ``libpbb_engine.so`` (hand-written HIP for gfx950 plus a
C ABI, include/pbb_engine.h); this package is its host-side mirror.
"""
from ._lib import Config, PbbError, Stats, load
from .engine import (INT_MAX, Bounds, CpuBB, GpuBB, SolveStats, checkpoint_format, checkpoint_parse,
                     dec_to_fact, default_config, explorers_for, fact_to_dec, hybrid_run, load_config, makespan, neh,
                     read_vrf, synth, synth_shape, taillard)

__all__ = [
    "Bounds", "Config", "CpuBB", "GpuBB", "INT_MAX", "PbbError", "SolveStats", "Stats", "checkpoint_format",
    "checkpoint_parse", "dec_to_fact", "default_config", "fact_to_dec", "hybrid_run",
    "explorers_for", "load", "load_config", "makespan", "neh", "read_vrf", "synth", "synth_shape",
    "taillard",
]
