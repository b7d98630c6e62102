"""Build the HIP library of the package for gfx950 with hipcc.

The library is built in-tree (hermeshd_amd/libhermeshd_amd.so) so that it
travels with the source tree; there is no JIT path.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libhermeshd_amd.so")
ARCH = os.environ.get("HD_OFFLOAD_ARCH", "gfx950")

COMPILE_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
LINK_FLAGS = ["-shared", "-fPIC", f"--offload-arch={ARCH}"]
# per-unit scheduler choice, kept where it measured faster: the
# memory-clause strategy speeds the nb 10 / nQ 5 row-shared stage kernel by
# about 2 % and the nb 4 ensemble stage; it slowed the stage transposes, so
# the nb 4 ones sit in hd_k4_vjp.hip with the default, and the 13-moment
# units keep the default too
_MEM_CLAUSE = ["-mllvm", "-amdgpu-sched-strategy=max-memory-clause"]
UNIT_FLAGS = {"hd_k10_5.hip": _MEM_CLAUSE, "hd_k4.hip": _MEM_CLAUSE}


def sources():
    return [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]


def build(verbose: bool = True, jobs: int | None = None) -> str:
    """Compiles each .hip unit to an object in parallel, then links the library."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    procs, objs = [], []
    jobs = jobs or min(8, os.cpu_count() or 1)
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        objs.append(obj)
        cmd = [hipcc, *COMPILE_FLAGS, *UNIT_FLAGS.get(os.path.basename(src), []), *inc, "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        while len([p for p in procs if p.poll() is None]) >= jobs:
            procs[[p.poll() is None for p in procs].index(True)].wait()
        procs.append(subprocess.Popen(cmd))
    failed = [p.args[-1] for p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    cmd = [hipcc, *LINK_FLAGS, "-o", LIB + ".tmp", *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
