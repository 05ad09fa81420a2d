"""Host-side checks that need no GPU: the C ABI library loads and exports
every entry point include/pbb_engine.h declares, the gpuconfig.ini surface,
the instance generators / readers, NEH and the factoradic codec agree with
the oracle and the reference's conventions."""
import ctypes as C
import os
import re
from math import factorial

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    with open(os.path.join(ROOT, "include", "pbb_engine.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*\*?\s*(pbb_\w+)\s*\(", text, re.M)))


def test_library_exports_every_header_symbol():
    import pbbperm_amd as P
    lib = P.load()
    syms = _header_symbols()
    assert len(syms) >= 28
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.pbb_version().decode().startswith("pbbperm_amd")


def test_python_bindings_cover_header():
    from pbbperm_amd import _lib
    assert set(_header_symbols()) == set(_lib.SIGNATURES)


def test_config_defaults_match_reference_defaults():
    import pbbperm_amd as P
    c = P.default_config()
    # common/src/arguments.cpp:135-175
    assert (c.bound_mode, c.branching, c.johnson_pairs, c.early_stop_johnson, c.nb_ivms) == (2, 3, 1, 1, 4096)
    assert c.sorted_dfs == 0 and c.find_all == 0


def test_ini_surface(tmp_path):
    import pbbperm_amd as P
    ini = tmp_path / "gpuconfig.ini"
    ini.write_text(
        "[problem]\nproblem = flowshop ;\ninstance = ta20 ;\n"
        "[initial]\nub = 0\n"
        "[bb]\nsingleNode = true\nsortedDFS = 0\nboundingMode = 2\n"
        "adaptiveBranchingMode = 1 ; //3\nJohnsonMode = 0  ; [0] full\nearlyStopJohnson = false\n"
        "findAll = false ;\n"
        "[gpu]\nnbIVMs = 16384\n[time]\nbalance = 5\n")
    cfg, prob, inst = P.load_config(str(ini))
    assert (prob, inst) == ("flowshop", "ta20")
    d = cfg.as_dict()
    assert d["bound_mode"] == 2 and d["branching"] == 1 and d["johnson_pairs"] == 0
    assert d["early_stop_johnson"] == 0 and d["single_node"] == 1 and d["nb_ivms"] == 16384
    assert d["init_mode"] == 0 and d["balance_s"] == 5
    # sortedDFS defaults to 1 when missing, as in the reference; the engine rejects it
    bad = tmp_path / "bad.ini"
    bad.write_text("[bb]\nboundingMode = 0\n")
    with pytest.raises(P.PbbError):
        P.load_config(str(bad))
    bad.write_text("[bb]\nsortedDFS = 0\nJohnsonMode = 3\n")
    with pytest.raises(P.PbbError):
        P.load_config(str(bad))
    # engine-specific keys: the team phase and the donation ramp (defaults
    # 8192 and 0), negative values refused
    assert d["team_nodes"] == 8192 and d["ramp_nodes"] == 0 and d["team_waves"] == 0
    bad.write_text("[bb]\nsortedDFS = 0\n[gpu]\nteamNodes = 0\nrampNodes = 32\nteamWaves = 4\n")
    cfg2, _, _ = P.load_config(str(bad))
    assert cfg2.team_nodes == 0 and cfg2.ramp_nodes == 32 and cfg2.team_waves == 4
    for key in ("teamNodes", "rampNodes", "teamWaves"):
        bad.write_text(f"[bb]\nsortedDFS = 0\n[gpu]\n{key} = -1\n")
        with pytest.raises(P.PbbError):
            P.load_config(str(bad))


def test_shipped_example_config_loads():
    import pbbperm_amd as P
    cfg, prob, inst = P.load_config(os.path.join(ROOT, "pbbperm_amd", "gpuconfig.ini"))
    assert prob == "flowshop" and cfg.sorted_dfs == 0


def test_generators_match_oracle(oracle_mod):
    import pbbperm_amd as P
    for i in (1, 11, 21, 31, 51, 81, 111):
        assert np.array_equal(P.taillard(i), oracle_mod.taillard(i))
    for s in range(5):
        assert np.array_equal(P.synth(s, 20, 7), oracle_mod.synth_ptm(s, 20, 7))
        assert P.synth_shape(100 + s, 10, 22, [5, 10]) == oracle_mod.synth_shape(100 + s, 10, 22, [5, 10])
    ptm = P.synth(0, 20, 5)
    assert ptm.min() >= 1 and ptm.max() <= 99


def test_neh_and_makespan_match_oracle(oracle_mod):
    import pbbperm_amd as P
    for s, n, m in [(0, 20, 5), (1, 20, 10), (2, 50, 20), (9, 7, 3)]:
        ptm = P.synth(s, n, m)
        inst = oracle_mod.Instance(ptm)
        c, perm = P.neh(ptm)
        oc, operm = inst.neh()
        assert c == oc and list(perm) == list(operm)
        assert P.makespan(ptm, perm) == inst.makespan(perm) == c


def test_vrf_reader(tmp_path):
    import pbbperm_amd as P
    ptm = P.synth(3, 9, 4)
    path = tmp_path / "VSYN_3_9x4_Gap.txt"
    lines = [f"{ptm.shape[1]} {ptm.shape[0]}"]
    for j in range(ptm.shape[1]):
        lines.append("  ".join(f"{k} {ptm[k, j]}" for k in range(ptm.shape[0])))
    path.write_text("\n".join(lines) + "\n")
    assert np.array_equal(P.read_vrf(str(path)), ptm)


def test_host_entry_points_reject_bad_inputs(tmp_path):
    """Inputs that would index out of range are refused: a VRF header whose
    n*m overflows 32 bits, job ids outside 0..n-1 or repeated, short arrays."""
    import pbbperm_amd as P
    from pbbperm_amd import _lib
    big = tmp_path / "big.txt"
    big.write_text("65536 65536\n0 1 1 1\n")
    with pytest.raises(P.PbbError):
        P.read_vrf(str(big))
    lib = P.load()
    n_, m_ = C.c_int(), C.c_int()
    buf = (C.c_int32 * 4)()
    assert lib.pbb_instance_read_vrf(str(big).encode(), C.byref(n_), C.byref(m_), buf, 4) == -4  # PBB_E_CAPACITY
    assert lib.pbb_instance_read_vrf(None, C.byref(n_), C.byref(m_), buf, 4) == -1
    ptm = P.synth(5, 6, 3)
    for bad in ([0, 0, 1, 2, 3, 4], [0, 1, 2, 3, 4, 6], [0, 1, 2, 3, 4, -1]):
        with pytest.raises(P.PbbError):
            P.makespan(ptm, bad)
        with pytest.raises(P.PbbError):
            P.neh(ptm, start=bad)
    with pytest.raises(ValueError):
        P.makespan(ptm, [0, 1, 2])
    with pytest.raises(ValueError):
        P.neh(ptm, start=[0, 1])
    # a valid start order other than the identity is still accepted
    c, perm = P.neh(ptm, start=[5, 4, 3, 2, 1, 0])
    assert sorted(perm.tolist()) == list(range(6)) and P.makespan(ptm, perm) == c


def test_wrappers_refuse_short_arrays():
    """The C side reads n (or count) values from every array it is given;
    the wrappers refuse arrays of another length before calling it."""
    import pbbperm_amd as P
    from pbbperm_amd import _lib
    assert _lib.sized([1, 2, 3], np.intc, 3, "x").dtype == np.intc
    with pytest.raises(ValueError):
        _lib.sized([1, 2], np.intc, 3, "x")
    ptm = P.synth(2, 6, 3)
    with pytest.raises(ValueError):
        P.GpuBB(ptm, ub=100, ub_perm=[0, 1, 2])
    with pytest.raises(ValueError):
        P.GpuBB(ptm, root=[0, 1, 2, 3, 4])
    with pytest.raises(ValueError):
        P.checkpoint_format(6, [0, 1], np.zeros(6, np.intc), np.zeros(12, np.intc), 10, range(6), 0, 10, range(6))


def test_factoradic_codec():
    from pbbperm_amd import factoradic as F
    for n in (3, 7, 20, 50):
        for v in (0, 1, factorial(n) - 1, factorial(n) // 3):
            assert F.to_int(F.to_digits(v, n)) == v
        p, e = F.full_interval(n)
        assert F.to_int(p) == 0 and F.to_int(e) == factorial(n) - 1
    pos, end = F.split(10, 7)
    assert F.to_int(pos[0]) == 0 and F.to_int(end[-1]) == factorial(10) - 1
    for k in range(1, 7):
        assert F.to_int(pos[k]) == F.to_int(end[k - 1]) + 1
    assert sum(F.length(p, e) for p, e in zip(pos, end)) == factorial(10)
    with pytest.raises(ValueError):
        F.to_int([3, 0, 0])


def test_routed_replay_rejects_bad_pool_lists():
    """Argument checks of the routed replay that run before any device work:
    an empty or oversized pool list, a null or repeated pool handle."""
    import pbbperm_amd as P
    from pbbperm_amd import _lib
    lib = P.load()
    ptm = np.ascontiguousarray(P.synth(0, 6, 3), dtype=np.int32)
    n = (C.c_int * 1)(6)
    m = (C.c_int * 1)(3)
    ub = (C.c_int * 1)(10**6)
    arr = (C.c_double * 1)(0.0)
    out = (_lib.Result * 1)()
    lat = (C.c_double * 1)()
    wall = C.c_double()
    fake = C.c_void_p(ptm.ctypes.data)   # never dereferenced: the lists below fail first

    def call(handles):
        hs = (C.c_void_p * max(1, len(handles)))(*handles)
        return lib.pbb_serve_replay_routed(hs, len(handles), 1, ptm.ctypes.data_as(_lib.I32P), n, m, ub, arr,
                                           1.0, out, lat, C.byref(wall))

    assert call([]) == -1                         # PBB_E_ARG
    assert call([None]) == -1
    assert "null pool" in lib.pbb_last_error().decode()
    assert call([fake, fake]) == -1
    assert "repeated" in lib.pbb_last_error().decode()
    assert call([fake] * 9) == -1


def test_serving_routed_params_keep_the_tuned_fields():
    """The routed serving pools: a 10-machine and a 20-machine pool, both for
    n <= 30, with the base parameters' tuned waves per group, explorers per
    instance, margin and request budget."""
    import pbbperm_amd.pool as PP
    base = PP.workload_params("serving", waves_per_group=2, explorers_per_instance=512, steal_margin=4,
                              request_budget_s=7.5)
    p10, p20 = PP.serving_routed_params(base)
    assert (p10.machines, p10.max_jobs, p20.machines, p20.max_jobs) == (10, 30, 20, 30)
    assert p10.request_budget_s == p20.request_budget_s == 7.5
    assert p10.groups > 0 and p20.groups > 0
    for q in (p10, p20):
        assert (q.waves_per_group, q.explorers_per_instance, q.steal_margin, q.slots) == (2, 512, 4, base.slots)
    # every request of the serving stream fits one of the two pools, and the
    # stream is not folded: n spans 12..30 at every machine count
    seen = {}
    for ptm in PP.serving_stream(600, 10000):
        m, n = ptm.shape
        assert any(m <= q.machines and n <= q.max_jobs for q in (p10, p20)), (n, m)
        seen.setdefault(m, set()).add(n)
    assert sorted(seen) == [5, 10, 20]
    for m, ns in seen.items():
        assert min(ns) == 12 and max(ns) == 30, (m, sorted(ns))


def test_engine_fails_loudly_without_device():
    import torch
    import pbbperm_amd as P
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(P.PbbError):
        P.GpuBB(P.synth(0, 10, 3))


def test_option_checks_before_any_device_work(tmp_path):
    """Option values the engine cannot honour are refused with PBB_E_ARG by
    the ini parser and by pbb_create before it touches a device: a negative
    incumbent (PBB_UB_NEH belongs to the pool entry points), an incumbent
    order that is not a permutation, a negative steal margin, and
    earlyStopJohnson with boundingMode 1 and the Johnson refinement (its
    early values would choose the branching direction,
    bounds/src/bound_fsp_strong.cpp:150-154, multicore/src/ivm_bound.cpp:668-676).
    The same option with boundingMode 2 only changes when an already-pruned
    child's bound stops, so it is accepted."""
    import pbbperm_amd as P
    from pbbperm_amd import _lib
    lib = P.load()
    ptm = np.ascontiguousarray(P.synth(1, 8, 3), dtype=np.int32)
    h = C.c_void_p()

    def create(cfg, ub, perm=None):
        p = None if perm is None else np.ascontiguousarray(perm, dtype=np.intc)
        return lib.pbb_create(ptm.ctypes.data_as(_lib.I32P), 8, 3, C.byref(cfg), ub,
                              None if p is None else p.ctypes.data_as(_lib.IP), C.byref(h))

    ok = P.default_config(bound_mode=0, early_stop_johnson=0)
    assert create(ok, -1) == -1 and "negative" in lib.pbb_last_error().decode()
    assert create(ok, 500, [0, 1, 2, 3, 4, 5, 6, 6]) == -1
    assert create(P.default_config(steal_margin=-1), 500) == -1
    bad = P.default_config(bound_mode=1, refine=1, early_stop_johnson=1)
    assert create(bad, 500) == -1 and "earlyStopJohnson" in lib.pbb_last_error().decode()
    ini = tmp_path / "gpuconfig.ini"
    ini.write_text("[bb]\nsortedDFS = 0\nboundingMode = 1\nrefinementBound = 1\nearlyStopJohnson = true\n")
    with pytest.raises(P.PbbError):
        P.load_config(str(ini))
    for text in ("boundingMode = 2\nrefinementBound = 1\nearlyStopJohnson = true\n",
                 "boundingMode = 1\nrefinementBound = 0\nearlyStopJohnson = true\n",
                 "boundingMode = 1\nrefinementBound = 1\nearlyStopJohnson = false\n"):
        ini.write_text("[bb]\nsortedDFS = 0\n" + text)
        P.load_config(str(ini))
