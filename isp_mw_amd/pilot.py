"""Learned pilot of the BM3D Wiener step (include/isp_pilot.h).

An 8-layer 3x3 residual CNN (48 channels, ReLU, BF16 weights and
activations, FP32 accumulation) that maps a noisy grey plane and its sigma to
an estimate of BM3D step 1; that estimate then guides the Wiener step
(BM3D_Final, source/BM3D.cpp:411-457) in place of BM3D_Basic's
(source/BM3D.cpp:339-385).  Plain step 1 is the fallback (``PilotBM3D`` uses
it when no trained network is given) and the reference the network is trained
on: training targets are GPU step-1 estimates of synthetic noisy crops.

Every convolution, gradient and optimiser update runs in the HIP kernels of
libisp_bm3d.so; torch provides device memory and the stream only.
"""
from __future__ import annotations

import ctypes
import math
import sys

import numpy as np

from . import _lib as L
from .bm3d import BM3D, BM3D_Basic, BM3D_Final, BM3D_Para

# (cin, cout, cin_pad, cout_pad, relu)
LAYERS = [(2, 48, 16, 48, True)] + [(48, 48, 48, 48, True)] * 6 + [(48, 1, 48, 16, False)]
# BM3D on a single plane scales its thresholds by normY = 1/sqrt(3) (include/BM3D.h:246-258);
# the pilot's targets and its two-step comparison use sigma * sqrt(3) so that
# the thresholds match noise of standard deviation sigma.
GREY_SIGMA_SCALE = math.sqrt(3.0)


def _plib():
    lib = L.lib()
    if not getattr(lib, "_pilot_typed", False):
        vp, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        f = ctypes.c_float
        lib.isp_pilot_kpad.argtypes = [i]
        lib.isp_pilot_conv3x3.argtypes = [vp, vp, vp, vp, vp, vp, i, i, i, i, i, i, vp]
        lib.isp_pilot_wgrad_partial_floats.argtypes = [i, i, i, i, i]
        lib.isp_pilot_wgrad_partial_floats.restype = sz
        lib.isp_pilot_conv3x3_wgrad.argtypes = [vp, vp, vp, vp, vp, vp, i, i, i, i, i, i, i, vp]
        lib.isp_pilot_pack.argtypes = [vp, vp, vp, i, i, i, i, vp]
        lib.isp_pilot_adam.argtypes = [vp, vp, vp, vp, ctypes.c_long, f, f, f, f, i, vp]
        lib.isp_pilot_input.argtypes = [vp, vp, vp, vp, i, i, i, vp]
        lib.isp_pilot_output.argtypes = [vp, vp, vp, vp, vp, vp, i, i, i, vp]
        lib._pilot_typed = True
    return lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise L.BM3DError(rc, f"{what} failed")


class PilotCNN:
    """The network, its FP32 master weights, Adam state and BF16 packed weights."""

    def __init__(self, seed: int = 0, device: int = 0):
        import torch
        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        lib = _plib()
        rng = np.random.default_rng(seed)
        sizes = []
        for cin, cout, _, _, relu in LAYERS:
            sizes += [cout * cin * 9, cout]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
        flat = np.zeros(self.offsets[-1], np.float32)
        for l, (cin, cout, _, _, relu) in enumerate(LAYERS):
            std = math.sqrt(2.0 / (cin * 9)) * (1.0 if relu else 0.1)
            o = self.offsets[2 * l]
            flat[o:o + cout * cin * 9] = rng.standard_normal(cout * cin * 9) * std
        self.params = torch.from_numpy(flat).to(self.dev)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.wf, self.wb = [], []
        for cin, cout, cinp, coutp, _ in LAYERS:
            self.wf.append(torch.empty(coutp * lib.isp_pilot_kpad(cinp), dtype=torch.int16, device=self.dev))
            self.wb.append(torch.empty(cinp * lib.isp_pilot_kpad(coutp), dtype=torch.int16, device=self.dev))
        self._bufs = {}
        self.pack()

    # views of the flat parameter / gradient vectors
    def w(self, l, t=None):
        t = self.params if t is None else t
        cin, cout = LAYERS[l][:2]
        o = self.offsets[2 * l]
        return t[o:o + cout * cin * 9]

    def b(self, l, t=None):
        t = self.params if t is None else t
        o = self.offsets[2 * l + 1]
        return t[o:o + LAYERS[l][1]]

    def bias_padded(self, l):
        cout, coutp = LAYERS[l][1], LAYERS[l][3]
        if cout == coutp:
            return self.b(l)
        out = self.torch.zeros(coutp, device=self.dev)
        out[:cout] = self.b(l)
        return out

    def pack(self):
        lib = _plib()
        s = self._stream()
        for l, (cin, cout, cinp, coutp, _) in enumerate(LAYERS):
            _check(lib.isp_pilot_pack(_ptr(self.w(l)), _ptr(self.wf[l]), _ptr(self.wb[l]), cin, cout, cinp, coutp, s),
                   "pack")

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def _buf(self, name, numel, dtype):
        t = self._bufs.get(name)
        if t is None or t.numel() < numel or t.dtype != dtype:
            t = self.torch.empty(numel, dtype=dtype, device=self.dev)
            self._bufs[name] = t
        return t[:numel]

    def forward(self, noisy, sigma, keep: bool = True, offset=None):
        """noisy [N,H,W] fp32 device, sigma [N] fp32 device (0..255), offset
        [N] fp32 device or None.  Returns (acts, res32): acts[0] the input,
        acts[l+1] layer l's output."""
        torch = self.torch
        lib = _plib()
        s = self._stream()
        N, H, W = noisy.shape
        P = N * H * W
        x = self._buf("act0", P * 16, torch.int16)
        _check(lib.isp_pilot_input(_ptr(noisy), _ptr(sigma), _ptr(offset), _ptr(x), N, H, W, s), "input")
        acts = [x]
        res32 = self._buf("res32", P * 16, torch.float32)
        cur = x
        for l, (cin, cout, cinp, coutp, relu) in enumerate(LAYERS):
            last = l == len(LAYERS) - 1
            name = f"act{l + 1}" if keep else f"ping{l % 2}"
            out = self._buf(name if not last else "res16", P * coutp, torch.int16)
            _check(lib.isp_pilot_conv3x3(_ptr(cur), None, _ptr(self.wf[l]), _ptr(self.bias_padded(l)), _ptr(out),
                                         _ptr(res32) if last else None, N, H, W, cinp, coutp, int(relu), s), "conv")
            if not last:
                acts.append(out)
            cur = out
        return acts, res32

    def loss_and_grads(self, noisy, sigma, target) -> float:
        """Mean-squared loss of (noisy - r) against target; gradients into self.grads."""
        torch = self.torch
        lib = _plib()
        s = self._stream()
        N, H, W = noisy.shape
        P = N * H * W
        acts, res32 = self.forward(noisy, sigma, keep=True)
        loss = torch.zeros(1, device=self.dev)
        g = self._buf("dres", P * 16, torch.int16)
        _check(lib.isp_pilot_output(_ptr(noisy), _ptr(res32), None, _ptr(target), _ptr(g), _ptr(loss), N, H, W, s),
               "output")
        part = self._buf("partial", max(lib.isp_pilot_wgrad_partial_floats(N, H, W, c[0], c[1]) for c in LAYERS),
                         torch.float32)
        for l in range(len(LAYERS) - 1, -1, -1):
            cin, cout, cinp, coutp, relu = LAYERS[l]
            mask = acts[l + 1] if relu else None
            _check(lib.isp_pilot_conv3x3_wgrad(_ptr(acts[l]), _ptr(g), _ptr(mask), _ptr(part),
                                               _ptr(self.w(l, self.grads)), _ptr(self.b(l, self.grads)), N, H, W,
                                               cinp, coutp, cin, cout, s), "wgrad")
            if l > 0:
                gn = self._buf(f"grad{l % 2}", P * cinp, torch.int16)
                _check(lib.isp_pilot_conv3x3(_ptr(g), _ptr(mask), _ptr(self.wb[l]), None, _ptr(gn), None, N, H, W,
                                             coutp, cinp, 0, s), "dinput")
                g = gn
        return loss

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.step += 1
        _check(_plib().isp_pilot_adam(_ptr(self.params), _ptr(self.grads), _ptr(self.m), _ptr(self.v),
                                      self.params.numel(), lr, beta1, beta2, eps, self.step, self._stream()), "adam")
        self.pack()

    def infer(self, noisy, sigma, offset=None):
        """noisy [H,W] or [N,H,W] fp32 device -> the estimate (same shape).
        sigma: one value or one per plane (0..255 units); offset: None or one
        value per plane added to the network input (0.5 for chroma centred on 0)."""
        torch = self.torch
        x = noisy if noisy.dim() == 3 else noisy[None]
        N, H, W = x.shape
        sg = torch.tensor(np.broadcast_to(np.asarray(sigma, np.float32), (N,)).copy(), device=self.dev)
        off = None if offset is None else torch.tensor(np.broadcast_to(np.asarray(offset, np.float32), (N,)).copy(),
                                                       device=self.dev)
        _, res32 = self.forward(x.contiguous(), sg, keep=False, offset=off)
        out = torch.empty_like(x)
        _check(_plib().isp_pilot_output(_ptr(x), _ptr(res32), _ptr(out), None, None, None, N, H, W,
                                        self._stream()), "output")
        return out if noisy.dim() == 3 else out[0]

    def state(self) -> dict:
        return {"params": self.params.cpu().numpy(), "m": self.m.cpu().numpy(), "v": self.v.cpu().numpy(),
                "step": self.step}

    def load_state(self, st: dict) -> None:
        torch = self.torch
        self.params.copy_(torch.from_numpy(np.asarray(st["params"], np.float32)))
        if "m" in st:
            self.m.copy_(torch.from_numpy(np.asarray(st["m"], np.float32)))
            self.v.copy_(torch.from_numpy(np.asarray(st["v"], np.float32)))
        self.step = int(st.get("step", 0))
        self.pack()

    def layer_params(self):
        """[(w [cout,cin,9], b [cout])] as numpy arrays (for the oracle)."""
        out = []
        for l, (cin, cout, _, _, _) in enumerate(LAYERS):
            out.append((self.w(l).cpu().numpy().astype(np.float64).reshape(cout, cin, 9),
                        self.b(l).cpu().numpy().astype(np.float64)))
        return out

    def layer_grads(self):
        out = []
        for l, (cin, cout, _, _, _) in enumerate(LAYERS):
            out.append((self.w(l, self.grads).cpu().numpy().astype(np.float64).reshape(cout, cin, 9),
                        self.b(l, self.grads).cpu().numpy().astype(np.float64)))
        return out


def make_training_set(n_crops: int, crop: int = 128, sigma_range=(10.0, 40.0), seed: int = 0,
                      source: int = 512, para: BM3D_Para | None = None, device: int = 0):
    """Random crops of seeded synthetic images (the 512x512 grey recipe) with
    noise of per-crop sigma, and their GPU BM3D step-1 estimates as targets.
    Returns device tensors noisy [n,crop,crop], sigma [n], target [n,crop,crop]."""
    import torch
    from . import synth
    from .batch import BM3DBatch
    rng = np.random.default_rng(seed)
    n_src = max(1, n_crops // 64)
    cleans = [synth.clean_grey(source, source, seed * 1000 + k) for k in range(n_src)]
    noisy = np.empty((n_crops, crop, crop), np.float32)
    sig = rng.uniform(*sigma_range, n_crops).astype(np.float32)
    for i in range(n_crops):
        c = cleans[i % n_src]
        y, x = rng.integers(0, source - crop + 1, 2)
        noisy[i] = c[y:y + crop, x:x + crop] + rng.standard_normal((crop, crop)).astype(np.float32) * (sig[i] / 255)
    p = para or default_para()
    f = BM3D_Basic(p.basic, device=device)
    nd = torch.from_numpy(noisy).to(f"cuda:{device}")
    tgt = torch.empty_like(nd)
    for s0 in range(0, n_crops, 4096):
        s1 = min(n_crops, s0 + 4096)
        b = BM3DBatch(f, [tgt[i] for i in range(s0, s1)], [nd[i] for i in range(s0, s1)],
                      sigmas=[float(v) * GREY_SIGMA_SCALE for v in sig[s0:s1]])
        b.run()
        torch.cuda.synchronize()
        b.close()
    return nd, torch.from_numpy(sig).to(nd.device), tgt


def default_para(sigma: float = 25.0) -> BM3D_Para:
    p = BM3D_Para("np")
    for st in (p.basic, p.final):
        st.BlockStep, st.BMrange, st.BMstep = 3, 12, 1
    p.basic.GroupSize, p.final.GroupSize = 16, 32
    p.set_sigma(sigma * GREY_SIGMA_SCALE)
    return p


def train(net: PilotCNN, noisy, sigma, target, steps: int, batch: int = 32, lr: float = 1e-3, seed: int = 0,
          log_every: int = 0) -> list[float]:
    """Adam on random mini-batches; returns the loss of every step."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    n = noisy.shape[0]
    losses = []
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(noisy.device)
        lr_t = lr * (0.5 * (1 + math.cos(math.pi * it / max(1, steps))))  # cosine decay
        loss = net.loss_and_grads(noisy[idx].contiguous(), sigma[idx].contiguous(), target[idx].contiguous())
        net.adam(lr_t)
        losses.append(loss)
        if log_every and (it + 1) % log_every == 0:
            print(f"pilot step {it + 1}/{steps} loss {float(loss.item()):.3e}", file=sys.stderr, flush=True)
    return [float(v.item()) for v in losses]


# OPP norms of Y, U, V (include/BM3D.h:241-248): the noise of an OPP plane is
# sigma * norm for RGB noise sigma
OPP_NORMS = (1.0 / math.sqrt(3.0), 1.0 / math.sqrt(2.0), math.sqrt(6.0) / 4.0)


def rgb_guide(net: PilotCNN, rgb, sigma: float):
    """The pilot's estimate of an RGB frame [3,H,W] (device fp32, noise sigma
    per channel in 0..255 units): the network runs on the OPP planes (Y, U, V
    with their own noise levels, chroma offset by 0.5), converted back to RGB
    without clip, so that BM3D_Final's own OPP conversion recovers them."""
    import torch
    lib = L.lib()
    _, H, W = rgb.shape
    s = ctypes.c_void_p(torch.cuda.current_stream(rgb.device).cuda_stream)
    yuv = torch.empty_like(rgb)
    _check(lib.isp_bm3d_rgb_to_opp(_ptr(yuv[0]), _ptr(yuv[1]), _ptr(yuv[2]), _ptr(rgb[0]), _ptr(rgb[1]), _ptr(rgb[2]),
                                   W, H, W, s), "rgb_to_opp")
    est = net.infer(yuv, [sigma * n for n in OPP_NORMS], [0.0, 0.5, 0.5])
    out = torch.empty_like(rgb)
    _check(lib.isp_bm3d_opp_to_rgb(_ptr(out[0]), _ptr(out[1]), _ptr(out[2]), _ptr(est[0]), _ptr(est[1]), _ptr(est[2]),
                                   W, H, W, 0, s), "opp_to_rgb")
    return out


class PilotBM3D:
    """Two-step BM3D on a grey plane [H,W] or an RGB frame [3,H,W] with the
    pilot's estimate as the Wiener guide.  Without a network it runs plain
    step 1 (the fallback).  para: the BM3D parameters of the plain two-step
    form (for grey planes with sigma scaled by GREY_SIGMA_SCALE, see
    default_para)."""

    def __init__(self, para: BM3D_Para, net: PilotCNN | None = None, device: int = 0, precision: str = "fp32",
                 rgb: bool = False):
        self.net = net
        self.rgb = rgb
        self.final = BM3D_Final(para.final, device=device, precision=precision)
        self.basic = BM3D_Basic(para.basic, device=device, precision=precision)
        self.noise_sigma = float(para.basic.sigma[0]) / (1.0 if rgb else GREY_SIGMA_SCALE)

    def guide(self, noisy):
        if self.net is None:
            return self.basic(noisy)
        if noisy.dim() == 3:
            return rgb_guide(self.net, noisy, self.noise_sigma)
        return self.net.infer(noisy, self.noise_sigma)

    def __call__(self, noisy):
        import torch
        g = self.guide(noisy)
        out = torch.empty_like(noisy)
        return self.final.process(out, noisy, g)


__all__ = ["PilotCNN", "PilotBM3D", "make_training_set", "train", "default_para", "rgb_guide", "LAYERS",
           "GREY_SIGMA_SCALE", "OPP_NORMS", "BM3D"]
