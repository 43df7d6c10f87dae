"""FP64 CPU oracle for ISP_MW's BM3D (test infrastructure only).

This package is the checker for the HIP path and the timed CPU baseline.  Only
tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
may import it.  The product package ``isp_mw_amd`` never imports it.

The C restatement lives in ``bm3d_oracle.c`` (built by ``make -C oracle`` into
``oracle/_build``); this module wraps it with ctypes and numpy.  Each wrapper
names the reference function it restates.

Parity pins (tests/test_oracle.py):
  * ``dct3d``      - scipy.fft.dctn type 2 / type 3 with norm=None, which are
                     FFTW's REDFT10 / REDFT01 (source/BM3D.cpp:14-15).
  * ``match_all``  - brute-force numpy search over the clipped window with a
                     lexsort, plus the invariants of include/Block.h:578-711.
  * ``kernel``     - an independent numpy formulation with orthonormal DCTs
                     and a uniform lambda*sigma threshold / sigma^2/2^zeta
                     Wiener noise, which the reference's unnormalised tables
                     (source/BM3D.cpp:26-70) reduce to; constant-image and
                     sigma->0 identities.
  * ``rgb2opp`` / ``opp2rgb`` - the OPP matrices and their exact inverse.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

_libs: dict[str, ctypes.CDLL] = {}


class StepParams(ctypes.Structure):
    """Mirror of BM3D_Para_Base (include/BM3D.h:15-45)."""

    _fields_ = [
        ("block_size", ctypes.c_int),
        ("block_step", ctypes.c_int),
        ("group_size", ctypes.c_int),
        ("bm_range", ctypes.c_int),
        ("bm_step", ctypes.c_int),
        ("thMSE", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("sigma", ctypes.c_double * 3),
    ]


def build() -> None:
    """Compile the oracle libraries (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class NlmParams(ctypes.Structure):
    """Mirror of NLMeans_Para (include/NLMeans.h:11-22)."""

    _fields_ = [
        ("correction", ctypes.c_int),
        ("sigma", ctypes.c_double),
        ("strength", ctypes.c_double),
        ("block_size", ctypes.c_int),
        ("block_step", ctypes.c_int),
        ("group_size", ctypes.c_int),
        ("bm_range", ctypes.c_int),
        ("bm_step", ctypes.c_int),
        ("thMSE", ctypes.c_double),
    ]


def nlm_params(correction=True, sigma=8.0, strength=None, BlockSize=8, BlockStep=5, GroupSize=16, BMrange=24,
               BMstep=3, thMSE=None) -> NlmParams:
    """NLMeans_Para defaults; strength and thMSE follow sigma as NLMeans_IO
    derives them (include/NLMeans.h:150-151) unless given."""
    if strength is None:
        strength = sigma * 5 if correction else sigma * 1.5
    if thMSE is None:
        thMSE = sigma * 50 if correction else sigma * 25
    return NlmParams(int(bool(correction)), float(sigma), float(strength), int(BlockSize), int(BlockStep),
                     int(GroupSize), int(BMrange), int(BMstep), float(thMSE))


def _lib(fast: bool = False) -> ctypes.CDLL:
    # ISP_ORACLE_SANITIZED=1: the ASan/UBSan build for both (the process must
    # preload the sanitizer runtime; tests/test_sanitizers.py)
    if os.environ.get("ISP_ORACLE_SANITIZED") == "1":
        name = "liborc_san.so"
    else:
        name = "liborc_fast.so" if fast else "liborc.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(_BUILD, name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    pp = ctypes.POINTER(dp)
    sp = ctypes.POINTER(StepParams)
    lib.orc_grid.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    lib.orc_set_variant.argtypes = [ctypes.c_int, ctypes.c_double]
    lib.orc_search_boundary.argtypes = [ctypes.c_int] * 4
    lib.orc_match_all.argtypes = [sp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int32), dp]
    lib.orc_kernel.argtypes = [sp, ctypes.c_int, ctypes.c_int, pp, pp, pp, ctypes.c_int, ctypes.c_int]
    ci = ctypes.c_int
    lib.orc_match_stack.argtypes = [sp, pp, ci, ci, ci, ci, ci, ctypes.POINTER(ci), ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ci)]
    lib.orc_kernel_stack.argtypes = [sp, ci, ci, ci, ci, pp, pp, pp, ci, ci]
    lib.orc_bm3d_stack.argtypes = [sp, sp, ci, ci, ci, ci, pp, pp, pp, ci, ci]
    lib.orc_bm3d_plane.argtypes = [sp, sp, dp, dp, dp, ctypes.c_int, ctypes.c_int]
    lib.orc_bm3d_rgb.argtypes = [sp, sp, ctypes.c_int, pp, pp, pp, ctypes.c_int, ctypes.c_int]
    lib.orc_rgb2opp.argtypes = [dp] * 6 + [ctypes.c_long]
    lib.orc_opp2rgb.argtypes = [dp] * 6 + [ctypes.c_long, ctypes.c_int]
    lib.orc_dct3d.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    ip = ctypes.POINTER(ctypes.c_int32)
    ipp = ctypes.POINTER(ip)
    lib.orc_range_to_float.argtypes = [dp, ip, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_range_to_int.argtypes = [ip, dp, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_bm3d_plane_int.argtypes = [sp, sp, ctypes.c_int, ip, ip, ip, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_bm3d_rgb_int.argtypes = [sp, sp, ctypes.c_int, ipp, ipp, ipp, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int]
    np_ = ctypes.POINTER(NlmParams)
    lib.orc_nlm_plane.argtypes = [np_, dp, dp, dp, ctypes.c_int, ctypes.c_int]
    lib.orc_nlm_rgb_int.argtypes = [np_, ipp, ipp, ipp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_nlm_plane_int.argtypes = [np_, ip, ip, ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_set_threads.restype = ctypes.c_int
    _libs[name] = lib
    return lib


def host_supports_fast() -> bool:
    """True when the host CPU has the AVX2/FMA/BMI2 set liborc_fast.so targets."""
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((ln for ln in f if ln.startswith("flags")), "")
    except OSError:
        return False
    return all(t in flags.split() for t in ("avx2", "fma", "bmi2", "movbe"))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def step_params(p) -> StepParams:
    """StepParams from any object with the BM3D_Para_Base field names."""
    sp = StepParams()
    sp.block_size = int(p.BlockSize)
    sp.block_step = int(p.BlockStep)
    sp.group_size = int(p.GroupSize)
    sp.bm_range = int(p.BMrange)
    sp.bm_step = int(p.BMstep)
    sp.thMSE = float(p.thMSE)
    sp.lambda_ = float(p.lambda_)
    sig = list(p.sigma) + [0.0] * (3 - len(p.sigma))
    for c in range(3):
        sp.sigma[c] = float(sig[c])
    return sp


def set_threads(n: int, fast: bool = False) -> int:
    return _lib(fast).orc_set_threads(int(n))


def set_variant(haar: bool = False, kaiser_beta: float = 0.0) -> None:
    """The configs' variant modes for later calls (no reference counterpart):
    Haar along the group axis with groups truncated to a power of two, and
    Kaiser-windowed aggregation.  set_variant() restores the reference."""
    for fast in (False, True):
        _lib(fast).orc_set_variant(int(bool(haar)), float(kaiser_beta))


def grid(n: int, block_size: int, block_step: int) -> np.ndarray:
    """Reference grid of one axis (source/BM3D.cpp:98-127)."""
    lib = _lib()
    cnt = lib.orc_grid(n, block_size, block_step, None, 0)
    out = (ctypes.c_int * max(cnt, 1))()
    lib.orc_grid(n, block_size, block_step, out, cnt)
    return np.array(out[:cnt], dtype=np.int32)


def search_boundary(pos: int, plane_boundary: int, search_range: int, search_step: int) -> int:
    """Block::_SearchBoundary (include/Block.h:716-745)."""
    return _lib().orc_search_boundary(pos, plane_boundary, search_range, search_step)


def match_all(p, ref: np.ndarray):
    """Groups for every grid point (source/BM3D.cpp:221-236, include/Block.h:658-703).

    Returns (counts[ngy, ngx], pos[ngy, ngx, G] as (y<<16)|x, keys[ngy, ngx, G]).
    """
    sp = step_params(p)
    ref = _f64(ref)
    h, w = ref.shape
    ngx = len(grid(w, sp.block_size, sp.block_step))
    ngy = len(grid(h, sp.block_size, sp.block_step))
    G = sp.group_size
    counts = np.zeros((ngy, ngx), dtype=np.int32)
    pos = np.zeros((ngy, ngx, G), dtype=np.int32)
    keys = np.zeros((ngy, ngx, G), dtype=np.float64)
    rc = _lib().orc_match_all(ctypes.byref(sp), _dptr(ref), w, h, w,
                              counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              pos.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dptr(keys))
    if rc != 0:
        raise ValueError("orc_match_all failed")
    return counts, pos, keys


def _planes(arrs):
    arrs = [_f64(a) for a in arrs]
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_dptr(a) for a in arrs])
    return arrs, ptrs


def kernel(p, src, ref, wiener: bool, fast: bool = False):
    """BM3D_Base::Kernel (source/BM3D.cpp:80-218) on 1 or 3 planes."""
    sp = step_params(p)
    src = [src] if np.ndim(src) == 2 else list(src)
    ref = [ref] if np.ndim(ref) == 2 else list(ref)
    n = len(src)
    h, w = np.shape(src[0])
    src_a, src_p = _planes(src)
    ref_a, ref_p = _planes(ref + [ref[0]] * (n - len(ref)))
    dst_a = [np.empty((h, w)) for _ in range(n)]
    dst_p = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(a) for a in dst_a])
    rc = _lib(fast).orc_kernel(ctypes.byref(sp), int(wiener), n, dst_p, src_p, ref_p, w, h)
    if rc < 0:
        raise ValueError(f"orc_kernel failed ({rc})")
    return dst_a[0] if n == 1 else np.stack(dst_a)


def bm3d_plane(basic, final, src, ref=None, fast: bool = False) -> np.ndarray:
    """Two-step BM3D on one plane (source/BM3D.cpp:464-472, intermediate sized like src)."""
    b, f = step_params(basic), step_params(final)
    src = _f64(src)
    h, w = src.shape
    refa = _f64(ref) if ref is not None else src
    dst = np.empty((h, w))
    rc = _lib(fast).orc_bm3d_plane(ctypes.byref(b), ctypes.byref(f), _dptr(dst), _dptr(src), _dptr(refa), w, h)
    if rc < 0:
        raise ValueError(f"orc_bm3d_plane failed ({rc})")
    return dst


def bm3d_rgb(basic, final, src, ref=None, mode: int = 3, fast: bool = False) -> np.ndarray:
    """Float-RGB Frame semantics of source/BM3D.cpp:282-313 / 501-527.  src: [3,H,W]."""
    b, f = step_params(basic), step_params(final)
    src_a, src_p = _planes(list(src))
    h, w = src_a[0].shape
    if ref is not None:
        ref_a, ref_p = _planes(list(ref))
    else:
        ref_p = None
    dst_a = [np.empty((h, w)) for _ in range(3)]
    dst_p = (ctypes.POINTER(ctypes.c_double) * 3)(*[_dptr(a) for a in dst_a])
    rc = _lib(fast).orc_bm3d_rgb(ctypes.byref(b), ctypes.byref(f), mode, dst_p, src_p, ref_p, w, h)
    if rc < 0:
        raise ValueError(f"orc_bm3d_rgb failed ({rc})")
    return np.stack(dst_a)


def rgb2opp(rgb: np.ndarray) -> np.ndarray:
    """MatrixConvert_RGB2YUV, OPP, float planes (include/Conversion.hpp:547-578)."""
    (r, g, b), _ = _planes(list(rgb))
    out = [np.empty_like(r) for _ in range(3)]
    _lib().orc_rgb2opp(*[_dptr(a) for a in out], _dptr(r), _dptr(g), _dptr(b), r.size)
    return np.stack(out)


def opp2rgb(yuv: np.ndarray, clip: bool = True) -> np.ndarray:
    """MatrixConvert_YUV2RGB, OPP, float planes (include/Conversion.hpp:739-762)."""
    (y, u, v), _ = _planes(list(yuv))
    out = [np.empty_like(y) for _ in range(3)]
    _lib().orc_opp2rgb(*[_dptr(a) for a in out], _dptr(y), _dptr(u), _dptr(v), y.size, int(clip))
    return np.stack(out)


def dct3d(a: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Unnormalised 3D REDFT10 (forward) / REDFT01 (inverse) on every axis."""
    out = _f64(a).copy()
    n0, n1, n2 = out.shape
    _lib().orc_dct3d(_dptr(out), n0, n1, n2, int(inverse))
    return out


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def range_to_float(a, floor, neutral, ceil) -> np.ndarray:
    """RangeConvert int -> float (include/Conversion.hpp:134-256)."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    out = np.empty(a.shape)
    _lib().orc_range_to_float(_dptr(out), _iptr(a), a.size, floor, neutral, ceil)
    return out


def range_to_int(a, floor, neutral, ceil) -> np.ndarray:
    """RangeConvert float -> int with clip (include/Conversion.hpp:241-256)."""
    a = _f64(a)
    out = np.empty(a.shape, np.int32)
    _lib().orc_range_to_int(_iptr(out), _dptr(a), a.size, floor, neutral, ceil)
    return out


def bm3d_plane_int(basic, final, src, ref=None, mode: int = 3, floor: int = 0, neutral: int = 0,
                   ceil: int = 255, fast: bool = False) -> np.ndarray:
    """Integer Plane path (source/BM3D.cpp:248-279, 475-498)."""
    b, f = step_params(basic), step_params(final)
    src = np.ascontiguousarray(src, dtype=np.int32)
    h, w = src.shape
    r = np.ascontiguousarray(ref, dtype=np.int32) if ref is not None else src
    dst = np.empty_like(src)
    rc = _lib(fast).orc_bm3d_plane_int(ctypes.byref(b), ctypes.byref(f), mode, _iptr(dst), _iptr(src), _iptr(r),
                                       w, h, floor, neutral, ceil)
    if rc < 0:
        raise ValueError(f"orc_bm3d_plane_int failed ({rc})")
    return dst


def bm3d_rgb_int(basic, final, src, ref=None, mode: int = 3, floor: int = 0, ceil: int = 255,
                 fast: bool = False) -> np.ndarray:
    """Integer Frame path (source/BM3D.cpp:282-313, 501-527).  src: int [3,H,W]."""
    b, f = step_params(basic), step_params(final)
    src = [np.ascontiguousarray(p, dtype=np.int32) for p in src]
    h, w = src[0].shape
    sp_ = (ctypes.POINTER(ctypes.c_int32) * 3)(*[_iptr(p) for p in src])
    if ref is not None:
        ref = [np.ascontiguousarray(p, dtype=np.int32) for p in ref]
        rp = (ctypes.POINTER(ctypes.c_int32) * 3)(*[_iptr(p) for p in ref])
    else:
        rp = None
    dst = [np.empty((h, w), np.int32) for _ in range(3)]
    dp_ = (ctypes.POINTER(ctypes.c_int32) * 3)(*[_iptr(p) for p in dst])
    rc = _lib(fast).orc_bm3d_rgb_int(ctypes.byref(b), ctypes.byref(f), mode, dp_, sp_, rp, w, h, floor, ceil)
    if rc < 0:
        raise ValueError(f"orc_bm3d_rgb_int failed ({rc})")
    return np.stack(dst)


def nlm_plane(p: NlmParams, src, ref=None) -> np.ndarray:
    """NLMeans::process_Plane_FL (source/NLMeans.cpp:150-228)."""
    src = _f64(src)
    h, w = src.shape
    r = _f64(ref) if ref is not None else src
    dst = np.empty_like(src)
    rc = _lib().orc_nlm_plane(ctypes.byref(p), _dptr(dst), _dptr(src), _dptr(r), w, h)
    if rc < 0:
        raise ValueError(f"orc_nlm_plane failed ({rc})")
    return dst


def nlm_plane_int(p: NlmParams, src, ref=None, floor: int = 0, neutral: int = 0, ceil: int = 255) -> np.ndarray:
    """NLMeans::process_Plane on an int plane (source/NLMeans.cpp:232-257)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    h, w = src.shape
    r = np.ascontiguousarray(ref, dtype=np.int32) if ref is not None else src
    dst = np.empty_like(src)
    rc = _lib().orc_nlm_plane_int(ctypes.byref(p), _iptr(dst), _iptr(src), _iptr(r), w, h, floor, neutral, ceil)
    if rc < 0:
        raise ValueError(f"orc_nlm_plane_int failed ({rc})")
    return dst


def nlm_rgb_int(p: NlmParams, src, ref=None, floor: int = 0, ceil: int = 255) -> np.ndarray:
    """NLMeans::process_Frame on an int RGB frame [3,H,W] (source/NLMeans.cpp:279-375)."""
    src = [np.ascontiguousarray(c, dtype=np.int32) for c in src]
    h, w = src[0].shape
    mk = lambda planes: (ctypes.POINTER(ctypes.c_int32) * 3)(*[_iptr(c) for c in planes])  # noqa: E731
    rr = [np.ascontiguousarray(c, dtype=np.int32) for c in ref] if ref is not None else None
    dst = [np.empty((h, w), np.int32) for _ in range(3)]
    rc = _lib().orc_nlm_rgb_int(ctypes.byref(p), mk(dst), mk(src), mk(rr) if rr is not None else None, w, h,
                                floor, ceil)
    if rc < 0:
        raise ValueError(f"orc_nlm_rgb_int failed ({rc})")
    return np.stack(dst)


def _stack_planes(frames):
    """[T, H, W] or [T, 3, H, W] -> (nplanes, flat list of planes frame-major)."""
    a = np.asarray(frames)
    if a.ndim == 3:
        return 1, [a[t] for t in range(a.shape[0])]
    if a.ndim == 4 and a.shape[1] == 3:
        return 3, [a[t, c] for t in range(a.shape[0]) for c in range(3)]
    raise ValueError("expected [T, H, W] or [T, 3, H, W]")


def match_stack(p, guides, radius: int, t: int):
    """Multi-frame groups of frame t's grid over the guide planes [T, H, W].
    Returns (counts[ngy, ngx], pos[ngy, ngx, G] as (y<<16)|x, frame[ngy, ngx, G])."""
    sp = step_params(p)
    g = np.asarray(guides)
    T, h, w = g.shape
    ga, gp = _planes([g[k] for k in range(T)])
    ngx = len(grid(w, sp.block_size, sp.block_step))
    ngy = len(grid(h, sp.block_size, sp.block_step))
    G = sp.group_size
    counts = np.zeros((ngy, ngx), dtype=np.int32)
    pos = np.zeros((ngy, ngx, G), dtype=np.int32)
    fz = np.zeros((ngy, ngx, G), dtype=np.int32)
    rc = _lib().orc_match_stack(ctypes.byref(sp), gp, T, radius, t, w, h,
                                counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                pos.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                fz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if rc != 0:
        raise ValueError("orc_match_stack failed")
    return counts, pos, fz


def bm3d_stack(basic, final, frames, radius: int, ref=None, mode: int = 3) -> np.ndarray:
    """BM3D of a stack of grey planes [T, H, W] or float RGB frames [T, 3, H, W]
    with multi-frame grouping of temporal radius `radius` (orc_bm3d_stack)."""
    b, f = step_params(basic), step_params(final)
    n, src = _stack_planes(frames)
    T = len(src) // n
    h, w = src[0].shape
    src_a, src_p = _planes(src)
    ref_p = None
    if ref is not None:
        n2, refl = _stack_planes(ref)
        assert n2 == n and len(refl) == len(src)
        ref_a, ref_p = _planes(refl)
    dst_a = [np.empty((h, w)) for _ in src]
    dst_p = (ctypes.POINTER(ctypes.c_double) * len(dst_a))(*[_dptr(a) for a in dst_a])
    rc = _lib().orc_bm3d_stack(ctypes.byref(b), ctypes.byref(f), mode, n, T, radius, dst_p, src_p, ref_p, w, h)
    if rc < 0:
        raise ValueError(f"orc_bm3d_stack failed ({rc})")
    out = np.stack(dst_a)
    return out.reshape((T, h, w) if n == 1 else (T, 3, h, w))


def kernel_stack(p, src, ref, radius: int, wiener: bool) -> np.ndarray:
    """One BM3D step over a stack (orc_kernel_stack): src / ref [T, H, W] or
    [T, 3, H, W]; matching on ref's plane 0 of each frame."""
    sp = step_params(p)
    n, sl = _stack_planes(src)
    _, rl = _stack_planes(ref)
    T = len(sl) // n
    h, w = sl[0].shape
    sa, spp = _planes(sl)
    ra, rpp = _planes(rl)
    dst_a = [np.empty((h, w)) for _ in sl]
    dst_p = (ctypes.POINTER(ctypes.c_double) * len(dst_a))(*[_dptr(a) for a in dst_a])
    rc = _lib().orc_kernel_stack(ctypes.byref(sp), int(wiener), n, T, radius, dst_p, spp, rpp, w, h)
    if rc < 0:
        raise ValueError(f"orc_kernel_stack failed ({rc})")
    out = np.stack(dst_a)
    return out.reshape((T, h, w) if n == 1 else (T, 3, h, w))
