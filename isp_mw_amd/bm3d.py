"""Python mirror of ISP_MW's BM3D interface over libisp_bm3d.so.

Class and field names follow include/BM3D.h so that code written against the
reference reads the same:

    para = BM3D_Para("np"); para.basic.sigma = [25.0] * 3; para.thMSE_Default()
    out = BM3D(para)(noisy)            # FilterIF2::operator()(src)
    BM3D_Basic(para.basic).process(dst, src, ref)

Planes are 2-D float32 arrays (Plane_FL, luma range [0,1]) or int32 arrays
(Plane: integer samples with Floor / Neutral / Ceil); frames are [3, H, W]
RGB arrays, float32 in [0,1] or int32 with Floor / Ceil (Frame).  numpy
arrays run through host memory; contiguous torch tensors on a HIP device run
in place on the current torch stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

PROFILES = ("fast", "lc", "np", "high", "vn")


class BM3D_Para_Base:
    """BM3D_Para_Base (include/BM3D.h:15-45)."""

    def __init__(self, profile: str = "fast"):
        self.profile = profile
        self.sigma = [10.0, 10.0, 10.0]
        self.BlockSize = 8
        self.BMrange = 16
        self.BMstep = 1
        if profile in ("fast", "lc"):
            self.BMrange = 9
        self.BlockStep = 0
        self.GroupSize = 0
        self.thMSE = 0.0
        self.lambda_ = 2.7

    def thMSE_Default(self) -> None:
        pass

    def to_c(self) -> L.StepParams:
        p = L.StepParams()
        p.profile = self.profile.encode()[:7]
        sig = list(self.sigma) + [self.sigma[-1]] * (3 - len(self.sigma))
        for c in range(3):
            p.sigma[c] = float(sig[c])
        p.block_size = int(self.BlockSize)
        p.block_step = int(self.BlockStep)
        p.group_size = int(self.GroupSize)
        p.bm_range = int(self.BMrange)
        p.bm_step = int(self.BMstep)
        p.thMSE = float(self.thMSE)
        p.lambda_ = float(self.lambda_)
        return p

    def __repr__(self) -> str:
        f = ("profile", "sigma", "BlockSize", "BlockStep", "GroupSize", "BMrange", "BMstep", "thMSE", "lambda_")
        return f"{type(self).__name__}(" + ", ".join(f"{k}={getattr(self, k)!r}" for k in f) + ")"


def _check_profile(profile: str, who: str) -> None:
    if profile not in PROFILES:
        raise ValueError(f'{who}: unrecognized profile "{profile}", should be "lc", "np", "vn" or "high"')


class BM3D_Basic_Para(BM3D_Para_Base):
    """BM3D_Basic_Para (include/BM3D.h:48-98)."""

    def __init__(self, profile: str = "fast"):
        super().__init__(profile)
        self.BlockStep = 4
        self.GroupSize = 16
        self.lambda_ = 2.7
        if profile == "fast":
            self.BlockStep = 8
        elif profile == "lc":
            self.BlockStep = 6
        elif profile == "high":
            self.BlockStep = 3
        elif profile == "vn":
            self.BlockStep = 4
            self.GroupSize = 32
            self.lambda_ = 2.8
        else:
            _check_profile(profile, "BM3D_Basic_Para")
        self.thMSE_Default()

    def thMSE_Default(self) -> None:
        self.thMSE = 400 + self.sigma[0] * 80
        if self.profile == "vn":
            self.thMSE = 1000 + self.sigma[0] * 150


class BM3D_Final_Para(BM3D_Para_Base):
    """BM3D_Final_Para (include/BM3D.h:101-151)."""

    def __init__(self, profile: str = "fast"):
        super().__init__(profile)
        self.BlockStep = 3
        self.GroupSize = 32
        if profile == "fast":
            self.BlockStep = 7
            self.GroupSize = 16
        elif profile == "lc":
            self.BlockStep = 5
            self.GroupSize = 16
        elif profile == "high":
            self.BlockStep = 2
        elif profile == "vn":
            self.BlockSize = 11
            self.BlockStep = 6
        else:
            _check_profile(profile, "BM3D_Final_Para")
        self.thMSE_Default()

    def thMSE_Default(self) -> None:
        self.thMSE = 200 + self.sigma[0] * 10
        if self.profile == "vn":
            self.thMSE = 400 + self.sigma[0] * 40


class BM3D_Para:
    """BM3D_Para (include/BM3D.h:154-168)."""

    def __init__(self, profile: str = "fast"):
        self.basic = BM3D_Basic_Para(profile)
        self.final = BM3D_Final_Para(profile)

    def thMSE_Default(self) -> None:
        self.basic.thMSE_Default()
        self.final.thMSE_Default()

    def set_sigma(self, sigma) -> "BM3D_Para":
        """What the CLI does with -S (include/BM3D.h:492-504): replicate the
        last value to 3 planes, copy to both steps, recompute thMSE."""
        s = list(sigma) if np.ndim(sigma) else [float(sigma)]
        s = s + [s[-1]] * (3 - len(s))
        self.basic.sigma = list(s[:3])
        self.final.sigma = list(s[:3])
        self.thMSE_Default()
        return self

    def to_c(self) -> L.Params:
        p = L.Params()
        p.basic = self.basic.to_c()
        p.final_ = self.final.to_c()
        return p


def _tuning_dict(t) -> dict:
    return {"match_tile": (t.match_tile[0], t.match_tile[1]), "filter_tile": (t.filter_tile[0], t.filter_tile[1]),
            "final_group_size": t.final_group_size, "ms_per_image": t.ms_per_image, "psnr_db": t.psnr_db}


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def io_args(dst, src, ref, floor, neutral, ceil, bit_depth):
    """Pointers, stream and memory kind of a FilterIF2::process call on numpy
    (host) or torch (device) arrays, and the integer quantisation."""
    if _is_torch(src):
        import torch
        if not (src.is_cuda and src.dtype in (torch.float32, torch.int32) and src.is_contiguous()):
            raise ValueError("torch inputs must be contiguous float32 or int32 tensors on a HIP device")
        is_int = src.dtype == torch.int32
        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        stream = ctypes.c_void_p(torch.cuda.current_stream(src.device).cuda_stream)
        mem = L.MEM_DEVICE
    else:
        is_int = isinstance(src, np.ndarray) and src.dtype == np.int32
        want = np.int32 if is_int else np.float32
        for a in (dst, src, ref):
            if a is not None and not (isinstance(a, np.ndarray) and a.dtype == want and a.flags.c_contiguous):
                raise ValueError("numpy inputs must be C-contiguous float32 (or all int32) arrays")
        ptr = lambda a: ctypes.c_void_p(a.ctypes.data) if a is not None else None  # noqa: E731
        stream = None
        mem = L.MEM_HOST
    if tuple(dst.shape) != tuple(src.shape) or (ref is not None and tuple(ref.shape) != tuple(src.shape)):
        raise ValueError("dst, src and ref must have the same shape")
    if is_int:
        floor = 0 if floor is None else int(floor)
        ceil = (1 << bit_depth) - 1 if ceil is None else int(ceil)
        neutral = floor if neutral is None else int(neutral)
    return ptr, stream, mem, is_int, (floor, neutral, ceil)


class _Filter:
    _mode = L.BOTH

    def __init__(self, params: L.Params, device: int = 0, precision: str = "fp32"):
        self._ctx = ctypes.c_void_p()
        lib = L.lib()
        L.check(lib.isp_bm3d_create(ctypes.byref(self._ctx), ctypes.byref(params), int(device)))
        self.device = device
        prec = {"fp32": L.PREC_FP32, "fp16": L.PREC_FP16, "bf16": L.PREC_BF16, "fp32fast": L.PREC_FP32_FAST,
                "fp32mfma": L.PREC_FP32_MFMA}[precision]
        L.check(lib.isp_bm3d_set_precision(self._ctx, prec), self._ctx)

    def set_variant(self, haar: bool = False, kaiser_beta: float = 0.0) -> None:
        """The configs' variant modes (no reference counterpart): the Haar
        transform along the group axis (groups truncated to a power of two),
        and Kaiser-windowed aggregation of shape kaiser_beta (0: off).  Runs
        the generic FP64 kernels; set_variant() restores the reference."""
        flags = (L.VARIANT_HAAR if haar else 0) | (L.VARIANT_KAISER if kaiser_beta else 0)
        L.check(L.lib().isp_bm3d_set_variant(self._ctx, flags, float(kaiser_beta)), self._ctx)

    def close(self) -> None:
        if getattr(self, "_ctx", None) and self._ctx.value:
            L.lib().isp_bm3d_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # FilterIF2::process(dst, src[, ref]) (include/Filter.h:173-209)
    def process(self, dst, src, ref=None, floor=None, neutral=None, ceil=None, bit_depth=8):
        """Integer samples use Floor / Neutral / Ceil (defaults: full range of
        bit_depth, luma, i.e. 0 / 0 / 2^bit_depth - 1)."""
        ptr, stream, mem, is_int, (floor, neutral, ceil) = io_args(dst, src, ref, floor, neutral, ceil, bit_depth)
        lib = L.lib()
        if src.ndim == 2:
            h, w = src.shape
            if is_int:
                rc = lib.isp_bm3d_plane_i32(self._ctx, ptr(dst), ptr(src), ptr(ref), w, h, w, floor, neutral, ceil,
                                            self._mode, mem, stream)
            else:
                rc = lib.isp_bm3d_plane_f32(self._ctx, ptr(dst), ptr(src), ptr(ref), w, h, w, self._mode, mem,
                                            stream)
        elif src.ndim == 3 and src.shape[0] == 3:
            _, h, w = src.shape
            d = [ptr(dst[c]) for c in range(3)]
            s_ = [ptr(src[c]) for c in range(3)]
            r = [ptr(ref[c]) for c in range(3)] if ref is not None else [None] * 3
            if is_int:
                rc = lib.isp_bm3d_rgb_i32(self._ctx, *d, *s_, *r, w, h, w, floor, ceil, self._mode, mem, stream)
            else:
                rc = lib.isp_bm3d_rgb_f32(self._ctx, *d, *s_, *r, w, h, w, self._mode, mem, stream)
        else:
            raise ValueError("expected a [H, W] plane or a [3, H, W] RGB frame")
        L.check(rc, self._ctx)
        return dst

    def process_stack(self, dst, src, radius: int = 1, ref=None):
        """Multi-frame grouping over a stack: src / dst / ref are [T, H, W]
        grey planes or [T, 3, H, W] float RGB frames (numpy: host memory,
        torch: device memory).  The groups of frame t take blocks from frames
        |z - t| <= radius; each filtered block returns to its own frame.
        radius 0 (or T = 1) equals filtering every frame on its own."""
        ptr, stream, mem, is_int, _ = io_args(dst, src, ref, None, None, None, 8)
        if is_int:
            raise ValueError("stacks are float frames")
        if src.ndim == 3:
            T, h, w = src.shape
            n = 1
            planes = lambda a: [a[t] for t in range(T)]  # noqa: E731
        elif src.ndim == 4 and src.shape[1] == 3:
            T, _, h, w = src.shape
            n = 3
            planes = lambda a: [a[t, c] for t in range(T) for c in range(3)]  # noqa: E731
        else:
            raise ValueError("expected [T, H, W] or [T, 3, H, W]")
        arr = ctypes.c_void_p * (T * n)
        d = arr(*[ptr(x) for x in planes(dst)])
        s_ = arr(*[ptr(x) for x in planes(src)])
        r = arr(*[ptr(x) for x in planes(ref)]) if ref is not None else None
        L.check(L.lib().isp_bm3d_stack_f32(self._ctx, d, s_, r, T, n, w, h, w, int(radius), self._mode, mem, stream),
                self._ctx)
        return dst

    # FilterIF2::operator()(src[, ref]) (include/Filter.h:203-216)
    def __call__(self, src, ref=None, **quant):
        if _is_torch(src):
            import torch
            dst = torch.empty_like(src)
        else:
            dst = np.empty_like(src)
        return self.process(dst, src, ref, **quant)

    def set_timing(self, enable: bool = True) -> None:
        """Bracket every kernel launch with hipEvents (per-stage device time)."""
        L.check(L.lib().isp_bm3d_set_timing(self._ctx, int(enable)), self._ctx)

    def timing(self, reset: bool = False) -> dict:
        """{stage: (total_ms, launches)} accumulated since the last reset."""
        n = L.NUM_STAGES
        names = (ctypes.c_char_p * n)()
        ms = (ctypes.c_double * n)()
        calls = (ctypes.c_long * n)()
        rc = L.lib().isp_bm3d_get_timing(self._ctx, names, ms, calls, n)
        if rc < 0:
            L.check(rc, self._ctx)
        out = {names[k].decode(): (ms[k], calls[k]) for k in range(n)}
        if reset:
            L.check(L.lib().isp_bm3d_reset_timing(self._ctx), self._ctx)
        return out

    def last_groups(self, final_step: bool, shape: tuple[int, int], with_pos: bool = False):
        """Groups of one step of the last call: counts [ngy, ngx] (and pos)."""
        ngy, ngx = shape
        counts = np.zeros((ngy, ngx), np.int32)
        G = self.group_size(final_step)
        pos = np.zeros((ngy, ngx, G), np.int32) if with_pos else None
        rc = L.lib().isp_bm3d_last_groups(self._ctx, int(final_step), ctypes.c_void_p(counts.ctypes.data),
                                          ctypes.c_void_p(pos.ctypes.data) if with_pos else None, ngy * ngx)
        if rc < 0:
            L.check(rc, self._ctx)
        return (counts, pos) if with_pos else counts

    # ---- tuning (tile shapes per size class; no reference counterpart)
    @staticmethod
    def size_class(width: int, height: int) -> int:
        return int(L.lib().isp_bm3d_size_class(int(width), int(height)))

    def autotune(self, width: int, height: int, sigma: float = 0.0, auto_group: bool = False,
                 rgb: bool = False) -> dict:
        """Measures tile shapes (and, with auto_group, a final GroupSize of 16)
        on seeded synthetic images (rgb: RGB frames) of this size; the choice
        applies to every later call and batch of the size class."""
        t = L.Tuning()
        flags = (L.BATCH_AUTO_GROUP if auto_group else 0) | (L.TUNE_RGB if rgb else 0)
        L.check(L.lib().isp_bm3d_autotune(self._ctx, int(width), int(height), float(sigma), flags, ctypes.byref(t)),
                self._ctx)
        return _tuning_dict(t)

    def set_tuning(self, size_class: int, tuning: dict | None) -> None:
        if tuning is None:
            L.check(L.lib().isp_bm3d_set_tuning(self._ctx, int(size_class), None), self._ctx)
            return
        t = L.Tuning()
        t.match_tile[0], t.match_tile[1] = tuning["match_tile"]
        t.filter_tile[0], t.filter_tile[1] = tuning["filter_tile"]
        t.final_group_size = int(tuning.get("final_group_size", 0))
        L.check(L.lib().isp_bm3d_set_tuning(self._ctx, int(size_class), ctypes.byref(t)), self._ctx)

    def get_tuning(self, size_class: int) -> dict | None:
        t = L.Tuning()
        rc = L.lib().isp_bm3d_get_tuning(self._ctx, int(size_class), ctypes.byref(t))
        if rc < 0:
            L.check(rc, self._ctx)
        return _tuning_dict(t) if rc == 1 else None

    def grid_shape(self, width: int, height: int, final_step: bool = False) -> tuple[int, int]:
        ngx, ngy = ctypes.c_int(), ctypes.c_int()
        L.check(L.lib().isp_bm3d_grid_size(self._ctx, width, height, int(final_step), ctypes.byref(ngx),
                                           ctypes.byref(ngy)), self._ctx)
        return ngy.value, ngx.value

    def match(self, guide: np.ndarray, final_step: bool = False):
        """Groups of one step on a host float32 plane: (counts, pos, keys) shaped
        [ngy, ngx], [ngy, ngx, G], [ngy, ngx, G]; pos packs (y << 16) | x."""
        guide = np.ascontiguousarray(guide, dtype=np.float32)
        h, w = guide.shape
        lib = L.lib()
        ngx, ngy = ctypes.c_int(), ctypes.c_int()
        L.check(lib.isp_bm3d_grid_size(self._ctx, w, h, int(final_step), ctypes.byref(ngx), ctypes.byref(ngy)),
                self._ctx)
        G = self.group_size(final_step)
        counts = np.zeros((ngy.value, ngx.value), np.int32)
        pos = np.zeros((ngy.value, ngx.value, G), np.int32)
        keys = np.zeros((ngy.value, ngx.value, G), np.float64)
        rc = lib.isp_bm3d_match(self._ctx, ctypes.c_void_p(guide.ctypes.data), w, h, w, int(final_step),
                                L.MEM_HOST, ctypes.c_void_p(counts.ctypes.data), ctypes.c_void_p(pos.ctypes.data),
                                ctypes.c_void_p(keys.ctypes.data))
        L.check(rc, self._ctx)
        return counts, pos, keys


class BM3D_Basic(_Filter):
    """BM3D_Basic (include/BM3D.h:287-309): the hard-threshold step."""

    _mode = L.BASIC

    def __init__(self, para: BM3D_Basic_Para | None = None, device: int = 0, precision: str = "fp32"):
        self.para = para or BM3D_Basic_Para()
        p = BM3D_Para(self.para.profile).to_c()
        p.basic = self.para.to_c()
        super().__init__(p, device, precision)

    def group_size(self, final_step: bool = False) -> int:
        return int(self.para.GroupSize)


class BM3D_Final(_Filter):
    """BM3D_Final (include/BM3D.h:312-334): the Wiener step guided by ref."""

    _mode = L.FINAL

    def __init__(self, para: BM3D_Final_Para | None = None, device: int = 0, precision: str = "fp32"):
        self.para = para or BM3D_Final_Para()
        p = BM3D_Para(self.para.profile).to_c()
        p.final_ = self.para.to_c()
        super().__init__(p, device, precision)

    def group_size(self, final_step: bool = True) -> int:
        return int(self.para.GroupSize)

    def match(self, guide: np.ndarray, final_step: bool = True):
        return super().match(guide, True)


class BM3D(_Filter):
    """BM3D (include/BM3D.h:337-357): basic step, then final step guided by it."""

    _mode = L.BOTH

    def __init__(self, para: BM3D_Para | None = None, device: int = 0, precision: str = "fp32"):
        self.para = para or BM3D_Para()
        super().__init__(self.para.to_c(), device, precision)

    def group_size(self, final_step: bool = False) -> int:
        return int((self.para.final if final_step else self.para.basic).GroupSize)


def c_params_default(profile: str) -> L.Params:
    """isp_bm3d_params_default from the C library (no device needed)."""
    p = L.Params()
    L.check(L.lib().isp_bm3d_params_default(ctypes.byref(p), profile.encode()))
    return p
