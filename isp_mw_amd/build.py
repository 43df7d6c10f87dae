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
# host-side AddressSanitizer + UndefinedBehaviorSanitizer builds for the CPU
# tests (tests/test_sanitizers.py): the front end, and the library whose C-ABI
# host code (bm3d_capi.cpp) is instrumented; device code is never instrumented
CLI_SAN = os.path.join(LIBDIR, "isp_mw_san")
LIB_SAN = os.path.join(LIBDIR, "libisp_bm3d_san.so")
HOST_SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
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

    def compile_capi_san() -> str:
        obj = os.path.join(OBJDIR, "bm3d_capi_san.o")
        cmd = [hipcc(), *compile_flags, *HOST_SANITIZE, "-g", "-c", "-o", obj, os.path.join(CSRC, "bm3d_capi.cpp")]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed compiling the sanitized bm3d_capi.cpp")
        return obj

    jobs = max(1, min(len(SOURCES) + 1, int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        san = ex.submit(compile_capi_san)
        objs = list(ex.map(compile_one, SOURCES))
        san_obj = san.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libisp_bm3d.so")
    os.replace(tmp, LIB)
    # the same library with the sanitized host code; the sanitizer runtime
    # comes from the process (preloaded), so its symbols stay undefined here
    tmp = LIB_SAN + ".tmp"
    objs_san = [san_obj if o.endswith("bm3d_capi.o") else o for o in objs]
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs_san]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libisp_bm3d_san.so")
    os.replace(tmp, LIB_SAN)
    build_cli(verbose)
    build_cli(verbose, sanitize=True)
    return LIB


def build_cli(verbose: bool = False, sanitize: bool = False) -> str:
    out = CLI_SAN if sanitize else CLI
    tmp = out + ".tmp"
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined"] if sanitize else ["-O2"]
    cmd = ["g++", *flags, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), CLI_SRC,
           "-L", LIBDIR, "-lisp_bm3d", "-lz", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building isp_mw")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
