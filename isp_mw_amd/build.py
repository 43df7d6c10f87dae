"""Build libisp_bm3d.so (HIP kernels + C ABI) for gfx950 with hipcc, and the
`isp_mw` command-line front end on it (g++, zlib).

Run ``python -m isp_mw_amd.build`` (or ``__graft_entry__.build()``).  Both are
written in-tree to ``isp_mw_amd/lib/``.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libisp_bm3d.so")
OBJDIR = os.path.join(LIBDIR, "_obj")  # per-source objects (git-ignored *.o)
CLI = os.path.join(LIBDIR, "isp_mw")
CLI_SRC = os.path.join(HERE, "cli", "isp_mw.cpp")

SOURCES = ["bm3d_kernels.hip", "nlm_kernels.hip", "stack_kernels.hip", "bm3d_capi.cpp", "pilot.hip", "scratch.hip",
           "generic_kernels.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off (their forms have misbehaved on these GPUs)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compiles every source to an object in parallel, then links the library."""
    os.makedirs(OBJDIR, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]

    def compile_one(src: str) -> str:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *compile_flags, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed compiling {src}")
        return obj

    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libisp_bm3d.so")
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


def build_cli(verbose: bool = False) -> str:
    tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), CLI_SRC,
           "-L", LIBDIR, "-lisp_bm3d", "-lz", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building isp_mw")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose=True))
