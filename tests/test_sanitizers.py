"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

Three builds made by the ordinary build (oracle/Makefile, isp_mw_amd/build.py):
  * oracle/_build/liborc_san.so - the FP64 oracle; the oracle's own pins
    (brute force, scipy, closed forms, golden fixtures) run against it;
  * isp_mw_amd/lib/isp_mw_san - the front end; its flag parsing and PNG /
    PNM / PFM codecs run over good and malformed files;
  * isp_mw_amd/lib/libisp_bm3d_san.so - the library with its C-ABI host code
    (bm3d_capi.cpp) instrumented; tests/host_harness.py drives what needs no
    GPU (presets, the aggregation bound, the host-only plan over many
    parameter sets and sizes, argument checks, context creation's error path).
Any sanitizer report aborts the child process and fails the test.
"""
import glob
import os
import struct
import subprocess
import sys
import zlib

import numpy as np
import pytest

from isp_mw_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:detect_odr_violation=0",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def gcc_runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def clang_asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/*/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def no_preload():
    # the sanitizer runtime must be the first library of the child; where the
    # environment already preloads something, these runs are left out
    if os.environ.get("LD_PRELOAD"):
        pytest.skip("LD_PRELOAD is already set in this environment")


def child_env(**extra):
    env = dict(os.environ)
    env.update(SAN_ENV)
    env.update(extra)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    return env


def test_oracle_under_sanitizers():
    no_preload()
    asan = gcc_runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc's libasan is not installed")
    so = os.path.join(ROOT, "oracle", "_build", "liborc_san.so")
    assert os.path.exists(so), "oracle/_build/liborc_san.so is built by make -C oracle"
    code = ("import sys, oracle, pytest\n"
            "assert oracle._lib()._name.endswith('liborc_san.so'), oracle._lib()._name\n"
            "sys.exit(pytest.main(['-q', '-x', '-p', 'no:cacheprovider', 'tests/test_oracle.py',"
            " 'tests/test_oracle_nlm.py', 'tests/test_oracle_stack.py', 'tests/test_golden.py']))\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=900,
                       env=child_env(LD_PRELOAD=asan, ISP_ORACLE_SANITIZED="1"))
    tail = (r.stdout + r.stderr)[-3000:]
    print(tail)
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def png(w, h, ct, bd, raw_rows, extra=b""):
    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    data = b"".join(b"\x00" + r for r in raw_rows)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, bd, ct, 0, 0, 0)) + extra +
            chunk(b"IDAT", zlib.compress(data)) + chunk(b"IEND", b""))


def test_cli_under_sanitizers(tmp_path):
    exe = B.CLI_SAN
    assert os.path.exists(exe), "isp_mw_san is built with the library"
    rng = np.random.default_rng(1)
    good = {
        "rgb8.png": png(5, 4, 2, 8, [rng.integers(0, 256, 15, dtype=np.uint8).tobytes() for _ in range(4)]),
        "grey16.png": png(3, 6, 0, 16, [rng.integers(0, 256, 6, dtype=np.uint8).tobytes() for _ in range(6)]),
        "rgba8.png": png(4, 2, 6, 8, [rng.integers(0, 256, 16, dtype=np.uint8).tobytes() for _ in range(2)]),
        "pal.png": png(2, 2, 3, 8, [bytes([0, 1]), bytes([1, 0])],
                       extra=struct.pack(">I", 6) + b"PLTE" + bytes(range(6)) +
                       struct.pack(">I", zlib.crc32(b"PLTE" + bytes(range(6))) & 0xFFFFFFFF)),
        "a.ppm": b"P6\n4 3\n255\n" + bytes(range(36)),
        "b.pgm": b"P5\n# c\n3 2\n65535\n" + bytes(12),
        "c.pfm": b"PF\n2 2\n-1.0\n" + np.zeros(12, "<f4").tobytes(),
    }
    bad = {
        "short_ihdr.png": b"\x89PNG\r\n\x1a\n" + struct.pack(">I", 4) + b"IHDR" + bytes(8),
        "trunc.png": png(5, 4, 2, 8, [bytes(15)] * 4)[:40],
        "neg.ppm": b"P6\n-4 4\n255\n" + bytes(48),
        "huge.pgm": b"P5\n99999999999 2\n255\n",
        "short.ppm": b"P6\n40 40\n255\n" + bytes(10),
        "junk.bin": bytes(range(256)),
        "empty.png": b"",
    }
    for name, data in {**good, **bad}.items():
        (tmp_path / name).write_bytes(data)
    runs = []
    for name in good:
        fmt = ".pfm" if name.endswith(".pfm") else ".ppm"   # float images are written as PFM
        runs.append((["--bm3d", str(tmp_path / name), "--copy", "-F", fmt], True))
    for name in bad:
        runs.append((["--bm3d", str(tmp_path / name), "--copy"], False))
    for flags in (["-P", "vn", "-S", "5"], ["-S", "1", "-S", "2", "-S", "3", "-S", "4"], ["-BS1", "11", "-GS2", "64"],
                  ["-P", "bogus"], ["-S"], ["--nope"], ["-TH1", "nan"]):
        runs.append((["--bm3d", "in.png", "--print-params", *flags], None))
    runs.append((["--nlm", "in.png", "--print-params", "-S", "9"], None))
    for args, ok in runs:
        r = subprocess.run([exe, *args], cwd=tmp_path, capture_output=True, text=True, timeout=120,
                           env=child_env())
        assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, (args, r.stderr[-2000:])
        assert r.returncode >= 0, (args, r.returncode)     # exited, not killed by a signal
        if ok is True:
            assert r.returncode == 0, (args, r.stderr)
        elif ok is False:
            assert r.returncode != 0, args


def test_library_host_code_under_sanitizers():
    no_preload()
    rt = clang_asan_runtime()
    if rt is None:
        pytest.skip("the ROCm LLVM AddressSanitizer runtime is not installed")
    assert os.path.exists(B.LIB_SAN), "libisp_bm3d_san.so is built with the library"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host_harness.py")], cwd=ROOT,
                       capture_output=True, text=True, timeout=600,
                       env=child_env(LD_PRELOAD=rt, ISP_BM3D_LIB=B.LIB_SAN,
                                     HARNESS_EXPECT_LIB="libisp_bm3d_san.so"))
    print((r.stdout + r.stderr)[-2000:])
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "host harness:" in r.stdout
