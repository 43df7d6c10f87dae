"""Batches of device images filtered together (include/isp_bm3d.h,
isp_bm3d_batch_*): one launch per stage for all images, replayed as one
hipGraph.  No reference counterpart: ISP_MW filters one image per call
(include/Filter.h:173-216); each image's result equals a single call on it.
"""
from __future__ import annotations

import ctypes

from . import _lib as L
from .bm3d import _Filter


class BM3DBatch:
    """Images are float32 torch tensors on the filter's device, [H, W] (grey)
    or [3, H, W] (RGB in [0, 1]), contiguous.  `sigmas[k]` > 0 gives image k
    its own sigma (0..255 units, both steps, thMSE from thMSE_Default)."""

    def __init__(self, filt: _Filter, dsts, srcs, sigmas=None, refs=None, auto_group: bool = False):
        """filt: BM3D (both steps), BM3D_Basic or BM3D_Final (refs are then the
        Wiener guides, as BM3D_Final::process(dst, src, ref))."""
        if not isinstance(filt, _Filter):
            raise TypeError("a batch runs a BM3D, BM3D_Basic or BM3D_Final filter")
        self._filt = filt
        self._flags = L.BATCH_AUTO_GROUP if auto_group else 0
        self._plan = ctypes.c_void_p()
        self._keep = None
        imgs, nplanes = self._images(dsts, srcs, sigmas, refs)
        lib = L.lib()
        L.check(lib.isp_bm3d_batch_create(filt._ctx, ctypes.byref(self._plan), imgs, len(dsts), nplanes,
                                          filt._mode, self._flags), filt._ctx)
        self.count, self.nplanes = len(dsts), nplanes

    def _images(self, dsts, srcs, sigmas, refs):
        if len(dsts) != len(srcs) or not srcs:
            raise ValueError("one dst per src, at least one image")
        n = len(srcs)
        sigmas = [0.0] * n if sigmas is None else [float(s) for s in sigmas]
        refs = [None] * n if refs is None else list(refs)
        nplanes = 3 if srcs[0].dim() == 3 else 1
        arr = (L.Image * n)()
        import torch
        for k in range(n):
            s, d, r = srcs[k], dsts[k], refs[k]
            for t in (s, d) + ((r,) if r is not None else ()):
                if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                    raise ValueError("batch images must be contiguous float32 device tensors")
                if tuple(t.shape) != tuple(s.shape):
                    raise ValueError(f"image {k}: dst/src/ref shapes differ")
            if (s.dim() == 3) != (nplanes == 3) or (nplanes == 3 and s.shape[0] != 3):
                raise ValueError("a batch holds grey [H, W] or RGB [3, H, W] images, not both")
            h, w = s.shape[-2:]
            planes = (lambda t, c: t[c] if nplanes == 3 else t)
            for c in range(nplanes):
                arr[k].dst[c] = planes(d, c).data_ptr()
                arr[k].src[c] = planes(s, c).data_ptr()
                arr[k].ref[c] = planes(r, c).data_ptr() if r is not None else None
            arr[k].width, arr[k].height, arr[k].stride = w, h, w
            arr[k].sigma = sigmas[k]
        self._keep = (list(dsts), list(srcs), refs)
        return arr, nplanes

    def run(self, graph: bool = True, stream=None) -> None:
        """Enqueue the whole batch on `stream` (default: the current torch stream)."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(self._filt.device)
        L.check(L.lib().isp_bm3d_batch_run(self._plan, ctypes.c_void_p(stream.cuda_stream), int(graph)),
                self._filt._ctx)

    def update(self, dsts, srcs, sigmas=None, refs=None) -> None:
        imgs, nplanes = self._images(dsts, srcs, sigmas, refs)
        if nplanes != self.nplanes or len(dsts) != self.count:
            raise ValueError("update keeps the number of images and planes")
        L.check(L.lib().isp_bm3d_batch_update(self._plan, imgs), self._filt._ctx)

    def info(self) -> dict:
        n, by, wg = ctypes.c_int(), ctypes.c_size_t(), ctypes.c_int()
        L.check(L.lib().isp_bm3d_batch_info(self._plan, ctypes.byref(n), ctypes.byref(by), ctypes.byref(wg)),
                self._filt._ctx)
        return {"launches": n.value, "device_bytes": by.value, "max_workgroups": wg.value}

    def close(self) -> None:
        if self._plan.value:
            L.lib().isp_bm3d_batch_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
