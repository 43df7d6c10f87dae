"""NL-means denoiser with ISP_MW's NLMeans interface (include/NLMeans.h) on
libisp_bm3d.so (include/isp_nlm.h): groups from the BM3D matching kernel,
weighted averages on the GPU."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .bm3d import _Filter, c_params_default, io_args


class NLMeans_Para:
    """NLMeans_Para (include/NLMeans.h:11-22).  strength and thMSE follow
    sigma and correction as NLMeans_IO derives them (:148-149) unless given."""

    def __init__(self, correction: bool = True, sigma: float = 8.0, strength: float | None = None,
                 BlockSize: int = 8, BlockStep: int = 5, GroupSize: int = 16, BMrange: int = 24, BMstep: int = 3,
                 thMSE: float | None = None):
        self.correction = bool(correction)
        self.sigma = sigma = float(sigma)
        self.strength = float(strength) if strength is not None else (sigma * 5 if correction else sigma * 1.5)
        self.BlockSize, self.BlockStep, self.GroupSize = int(BlockSize), int(BlockStep), int(GroupSize)
        self.BMrange, self.BMstep = int(BMrange), int(BMstep)
        self.thMSE = float(thMSE) if thMSE is not None else (sigma * 50 if correction else sigma * 25)

    def to_c(self) -> L.NlmParams:
        return L.NlmParams(int(self.correction), self.sigma, self.strength, self.BlockSize, self.BlockStep,
                           self.GroupSize, self.BMrange, self.BMstep, self.thMSE)

    def __repr__(self) -> str:
        return ("NLMeans_Para(" + ", ".join(f"{k}={v}" for k, v in vars(self).items()) + ")")


class NLMeans(_Filter):
    """NLMeans (include/NLMeans.h:25-59): FilterIF2 process / __call__ on a
    float plane [H, W], an int plane [H, W] or an int RGB frame [3, H, W];
    numpy arrays are host memory, torch tensors device memory.
    precision "fp32" forms the reference's groups exactly; "fp32fast", "fp16"
    and "bf16" match with FP32 distances."""

    def __init__(self, para: NLMeans_Para | None = None, device: int = 0, precision: str = "fp32"):
        super().__init__(c_params_default("fast"), device, precision)
        self.para = para if para is not None else NLMeans_Para()

    def process(self, dst, src, ref=None, floor=None, neutral=None, ceil=None, bit_depth=8):
        ptr, stream, mem, is_int, (floor, neutral, ceil) = io_args(dst, src, ref, floor, neutral, ceil, bit_depth)
        p = ctypes.byref(self.para.to_c())
        lib = L.lib()
        if src.ndim == 2:
            h, w = src.shape
            if is_int:
                rc = lib.isp_nlm_plane_i32(self._ctx, p, ptr(dst), ptr(src), ptr(ref), w, h, w, floor, neutral, ceil,
                                           mem, stream)
            else:
                rc = lib.isp_nlm_plane_f32(self._ctx, p, ptr(dst), ptr(src), ptr(ref), w, h, w, mem, stream)
        elif src.ndim == 3 and src.shape[0] == 3 and is_int:
            _, h, w = src.shape
            d = [ptr(dst[c]) for c in range(3)]
            s_ = [ptr(src[c]) for c in range(3)]
            r = [ptr(ref[c]) for c in range(3)] if ref is not None else [None] * 3
            rc = lib.isp_nlm_rgb_i32(self._ctx, p, *d, *s_, *r, w, h, w, floor, ceil, mem, stream)
        else:
            raise ValueError("expected a [H, W] plane or an int32 [3, H, W] RGB frame")
        L.check(rc, self._ctx)
        return dst


def denoise(src: np.ndarray, sigma: float = 8.0, correction: bool = True, **kw) -> np.ndarray:
    """One-shot NL-means of a host array with NLMeans_Para(correction, sigma, **kw)."""
    f = NLMeans(NLMeans_Para(correction, sigma, **kw))
    try:
        return f(src)
    finally:
        f.close()
