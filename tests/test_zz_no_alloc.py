"""The allocation-free NEH entry point (pbb_neh_device_noalloc) against the
allocating one (pbb_neh_device), through the C ABI, on two calls in a row
(the second smaller, so stale offsets from the first would show), plus its
refusals.  Run last: it shares the library's fixed device array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stream(P, seed0, count):
    rng = np.random.default_rng(seed0)
    ptms = []
    for i in range(count):
        n = int(rng.integers(3, 64))
        m = int(rng.integers(2, 21))
        ptms.append(P.synth(seed0 + i, n, m))
    return ptms


def _noalloc(lib, torch, ptms, max_m, stream=None):
    flat = np.concatenate([p.ravel() for p in ptms]).astype(np.int32)
    ns = np.array([p.shape[1] for p in ptms], dtype=np.int32)
    ms = np.array([p.shape[0] for p in ptms], dtype=np.int32)
    d_ptm = torch.from_numpy(flat).cuda()
    d_n = torch.from_numpy(ns).cuda()
    d_m = torch.from_numpy(ms).cuda()
    d_perm = torch.full((len(ptms) * 64,), 7, dtype=torch.int32, device="cuda")
    d_cost = torch.full((len(ptms),), 7, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream if stream is None else stream
    rc = lib.pbb_neh_device_noalloc(len(ptms), d_ptm.data_ptr(), d_n.data_ptr(), d_m.data_ptr(), max_m,
                                    d_perm.data_ptr(), d_cost.data_ptr(), s)
    torch.cuda.synchronize()
    assert rc == 0, lib.pbb_last_error().decode()
    perms = d_perm.cpu().numpy()
    costs = d_cost.cpu().numpy()
    return [perms[64 * i:64 * i + int(ns[i])].copy() for i in range(len(ptms))], costs


def test_noalloc_neh_matches_allocating_entry_point():
    import torch
    import pbbperm_amd as P
    import pbbperm_amd.pool as PP
    lib = P.load()
    for count, seed in ((700, 61000), (90, 62000)):
        ptms = _stream(P, seed, count)
        print(f"noalloc NEH: {count} instances, seed {seed}", flush=True)
        want_o, want_c = PP.neh_device(ptms)
        got_o, got_c = _noalloc(lib, torch, ptms, max(p.shape[0] for p in ptms))
        assert [int(c) for c in got_c] == [int(c) for c in want_c]
        for a, b in zip(got_o, want_o):
            assert list(a) == list(b)


def test_noalloc_neh_refusals():
    import torch
    import pbbperm_amd as P
    lib = P.load()
    dummy = torch.zeros(64, dtype=torch.int32, device="cuda")
    ptr = dummy.data_ptr()
    # more instances than the fixed array holds
    assert lib.pbb_neh_device_noalloc(32769, ptr, ptr, ptr, 20, ptr, ptr, None) == -4
    assert b"32768" in lib.pbb_last_error()
    assert lib.pbb_neh_device_noalloc(-1, ptr, ptr, ptr, 20, ptr, ptr, None) == -1
    assert lib.pbb_neh_device_noalloc(4, ptr, ptr, ptr, 21, ptr, ptr, None) == -1
    assert lib.pbb_neh_device_noalloc(4, None, ptr, ptr, 20, ptr, ptr, None) == -1
    assert lib.pbb_neh_device_noalloc(0, None, None, None, 20, None, None, None) == 0
    # shapes outside the kernel's range (m > max_m) give -1 on the device; the
    # others still read their own times (the refused ones count in the offsets)
    ptms = _stream(P, 63000, 12)
    small = sorted(p.shape[0] for p in ptms)[len(ptms) // 2]
    got_o, got_c = _noalloc(lib, torch, ptms, small)
    import pbbperm_amd.pool as PP
    want_o, want_c = PP.neh_device(ptms)
    refused = 0
    for p, o, c, wo, wc in zip(ptms, got_o, got_c, want_o, want_c):
        if p.shape[0] > small:
            refused += 1
            assert int(c) == -1 and all(v == -1 for v in o)
        else:
            assert int(c) == int(wc) and list(o) == list(wo)
    assert 0 < refused < len(ptms)
