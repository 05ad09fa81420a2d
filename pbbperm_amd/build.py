"""Build recipe for the in-tree gfx950 engine library ``libpbb_engine.so``.

The kernel sources (one-instance engine, multi-instance pool) are compiled
once per machine count M (2..20, -DPBB_M=M) in parallel, then linked with
the runtime-M dispatch and the host C ABI.
The library is built in-tree so it travels with the repository snapshot.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpbb_engine.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
MACHINES = range(2, 21)

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-mcode-object-version=5",
    # packed fp32 ops disabled for all device code (the path has no float)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return " ".join(cmd) + "\n" + r.stdout + r.stderr
    return None


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    units = []
    for m in MACHINES:
        for src in ("pbb_kernels", "pbb_pool", "pbb_ig"):
            obj = os.path.join(BUILD, f"{src}_m{m}.o")
            units.append([HIPCC, *FLAGS, f"-DPBB_M={m}", "-c", os.path.join(CSRC, src + ".hip"), "-o", obj])
    for src in ("pbb_dispatch.cpp", "pbb_host.cpp", "pbb_io.cpp", "pbb_pool.cpp", "pbb_checkpoint.cpp", "pbb_cpu.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        units.append([HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
    # compiled once: it defines the one device array of the allocation-free entry point
    units.append([HIPCC, *FLAGS, "-c", os.path.join(CSRC, "pbb_noalloc.hip"),
                  "-o", os.path.join(BUILD, "pbb_noalloc.o")])
    units.append([HIPCC, *FLAGS, "-c", os.path.join(CSRC, "pbb_setup.hip"),
                  "-o", os.path.join(BUILD, "pbb_setup.o")])
    with cf.ThreadPoolExecutor(jobs) as ex:
        errs = [e for e in ex.map(_run, units) if e]
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError("hipcc failed building libpbb_engine.so")
    objs = [u[-1] for u in units]
    tmp = LIB + ".tmp"
    err = _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    if err:
        sys.stderr.write(err)
        raise RuntimeError("link failed for libpbb_engine.so")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(objs)} objects", flush=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
