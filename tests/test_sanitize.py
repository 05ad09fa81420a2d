"""The host parsers under the CPU sanitizers: the gpuconfig.ini reader, the
VRF instance reader, the checkpoint text parser and formatter, and the
factoradic <-> decimal codec, built with g++ -fsanitize=address,undefined
from the library's device-free host sources (pbb_io.cpp, pbb_checkpoint.cpp)
and driven by tests/sanitize/fuzz_parsers.cpp over seed files and seeded
mutations of them.  Any sanitizer report aborts the fuzzing program."""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "pbbperm_amd", "csrc")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is required")


def _seeds(d):
    import pbbperm_amd as P

    with open(os.path.join(ROOT, "pbbperm_amd", "gpuconfig.ini")) as f:
        ini = f.read()
    files = {
        "ini_shipped.txt": ini,
        "ini_edge.txt": "[bb]\nboundingMode = 1 ; inline\nbranchingMode=\n[gpu\nnbIVMs: 65536\n# c\n"
                        "[bb]\nfindAll = yes\nearlyStopJohnson = off\nsortedDFS = 0x0\n",
    }
    for n, m, seed in [(5, 3, 1), (20, 5, 2), (12, 20, 3)]:
        ptm = P.synth(seed, n, m)
        lines = [f"{n} {m}"]
        for j in range(n):
            lines.append(" ".join(f"{k} {ptm[k, j]}" for k in range(m)))
        files[f"vrf_{n}x{m}.txt"] = "\n".join(lines) + "\n"
    rng = np.random.default_rng(5)
    for n in (4, 20, 63):
        k = 3
        ids = np.arange(k)
        pos = np.array([[rng.integers(0, n - i) for i in range(n)] for _ in range(k)])
        end = np.array([[n - 1 - i for i in range(n)] for _ in range(k)])
        perm = rng.permutation(n)
        files[f"ckpt_{n}_a.txt"] = P.checkpoint_format(n, ids, pos, end, 1234, perm, 98765, 1200, perm,
                                                       explored=42, work_id=3, max_intervals=8)
        files[f"dec_{n}_a.txt"] = P.fact_to_dec(pos[0])
    files["dec_63_max.txt"] = P.fact_to_dec([63 - 1 - i for i in range(63)])
    for name, text in files.items():
        with open(os.path.join(d, name), "w") as f:
            f.write(text)


def test_parsers_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "fuzz_parsers")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           os.path.join(HERE, "sanitize", "fuzz_parsers.cpp"),
           os.path.join(CSRC, "pbb_io.cpp"), os.path.join(CSRC, "pbb_checkpoint.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    seeds = tmp_path / "seeds"
    work = tmp_path / "work"
    seeds.mkdir()
    work.mkdir()
    _seeds(str(seeds))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(seeds), str(work), "400"], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-6000:])
    assert "parser runs" in r.stdout
