"""Learned pilot of the BM3D Wiener step (include/isp_pilot.h).

An 8-layer 3x3 residual CNN (48 channels, ReLU, BF16 weights and
activations, FP32 accumulation) that maps a noisy image and its sigma to an
estimate of BM3D step 1; that estimate then guides the Wiener step
(BM3D_Final, source/BM3D.cpp:411-457) in place of BM3D_Basic's
(source/BM3D.cpp:339-385).  A grey network takes one plane; an RGB network
takes the three OPP planes of a frame (Y, U, V: the planes BM3D filters,
include/Conversion.hpp:547-578) at once, so its chroma estimates see the luma
structure, as BM3D's chroma groups do (they are matched on Y).  Plain step 1 is
the fallback (``PilotBM3D`` uses it when no trained network is given) and the
reference the network is trained on: training targets are GPU step-1
estimates of synthetic noisy crops.

Every convolution, gradient and optimiser update runs in the HIP kernels of
libisp_bm3d.so; torch provides device memory and the stream only.
"""
from __future__ import annotations

import ctypes
import math
import os
import sys

import numpy as np

from . import _lib as L
from .bm3d import BM3D, BM3D_Basic, BM3D_Final, BM3D_Para


def layers(channels: int = 1):
    """(cin, cout, cin_pad, cout_pad, relu) of the 8 layers of a network on
    images of `channels` planes (1: grey, 3: OPP Y, U, V): the planes and a
    sigma map in, one residual per plane out."""
    return [(channels + 1, 48, 16, 48, True)] + [(48, 48, 48, 48, True)] * 6 + [(48, channels, 48, 16, False)]


LAYERS = layers(1)
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
        lib.isp_pilot_input_planes.argtypes = [vp, i, vp, vp, vp, i, i, i, vp]
        lib.isp_pilot_output_planes.argtypes = [vp, vp, vp, vp, vp, vp, i, i, i, i, vp]
        lib.isp_pilot_conv3x3_residual.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, i, i, vp]
        lib._pilot_typed = True
    return lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise L.BM3DError(rc, f"{what} failed")


class PilotCNN:
    """The network, its FP32 master weights, Adam state and BF16 packed weights.
    channels: planes per image (1: grey planes, 3: the OPP planes of RGB frames)."""

    def __init__(self, seed: int = 0, device: int = 0, channels: int = 1):
        import torch
        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        if channels not in (1, 3):
            raise ValueError("channels must be 1 (grey) or 3 (OPP planes of RGB frames)")
        self.channels = channels
        self.layers = layers(channels)
        lib = _plib()
        rng = np.random.default_rng(seed)
        sizes = []
        for cin, cout, _, _, relu in self.layers:
            sizes += [cout * cin * 9, cout]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
        flat = np.zeros(self.offsets[-1], np.float32)
        for l, (cin, cout, _, _, relu) in enumerate(self.layers):
            std = math.sqrt(2.0 / (cin * 9)) * (1.0 if relu else 0.1)
            o = self.offsets[2 * l]
            flat[o:o + cout * cin * 9] = rng.standard_normal(cout * cin * 9) * std
        self.params = torch.from_numpy(flat).to(self.dev)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.wf, self.wb = [], []
        for cin, cout, cinp, coutp, _ in self.layers:
            self.wf.append(torch.empty(coutp * lib.isp_pilot_kpad(cinp), dtype=torch.int16, device=self.dev))
            self.wb.append(torch.empty(cinp * lib.isp_pilot_kpad(coutp), dtype=torch.int16, device=self.dev))
        self._bufs = {}
        self.pack()

    # views of the flat parameter / gradient vectors
    def w(self, l, t=None):
        t = self.params if t is None else t
        cin, cout = self.layers[l][:2]
        o = self.offsets[2 * l]
        return t[o:o + cout * cin * 9]

    def b(self, l, t=None):
        t = self.params if t is None else t
        o = self.offsets[2 * l + 1]
        return t[o:o + self.layers[l][1]]

    def bias_padded(self, l):
        cout, coutp = self.layers[l][1], self.layers[l][3]
        if cout == coutp:
            return self.b(l)
        out = self.torch.zeros(coutp, device=self.dev)
        out[:cout] = self.b(l)
        return out

    def pack(self):
        lib = _plib()
        s = self._stream()
        for l, (cin, cout, cinp, coutp, _) in enumerate(self.layers):
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

    def _shape(self, noisy):
        """(N, H, W) of a batch: [N,H,W] planes (grey) or [N,3,H,W] frames."""
        if self.channels == 1:
            return tuple(noisy.shape)
        N, C, H, W = noisy.shape
        if C != self.channels:
            raise ValueError(f"expected {self.channels} planes per image, got {C}")
        return N, H, W

    def forward(self, noisy, sigma, keep: bool = True, offset=None, est=None):
        """noisy [N,H,W] (grey) or [N,3,H,W] (OPP frames) fp32 device, sigma [N]
        fp32 device (0..255), offset one fp32 value per plane ([N] or [N,3]) on
        the device, or None.  Returns (acts, res32): acts[0] the input,
        acts[l+1] layer l's output.  With est (a tensor shaped like noisy),
        the last layer writes the estimate noisy - residual into it directly
        (isp_pilot_conv3x3_residual) and res32 is None."""
        torch = self.torch
        lib = _plib()
        s = self._stream()
        N, H, W = self._shape(noisy)
        P = N * H * W
        x = self._buf("act0", P * 16, torch.int16)
        _check(lib.isp_pilot_input_planes(_ptr(noisy), self.channels, _ptr(sigma), _ptr(offset), _ptr(x), N, H, W,
                                          s), "input")
        acts = [x]
        res32 = None if est is not None else self._buf("res32", P * 16, torch.float32)
        cur = x
        for l, (cin, cout, cinp, coutp, relu) in enumerate(self.layers):
            last = l == len(self.layers) - 1
            if last and est is not None:
                _check(lib.isp_pilot_conv3x3_residual(_ptr(cur), _ptr(self.wf[l]), _ptr(self.bias_padded(l)),
                                                      _ptr(noisy), _ptr(est), self.channels, N, H, W, cinp, coutp,
                                                      s), "residual")
                return acts, None
            name = f"act{l + 1}" if keep else f"ping{l % 2}"
            out = self._buf(name if not last else "res16", P * coutp, torch.int16)
            _check(lib.isp_pilot_conv3x3(_ptr(cur), None, _ptr(self.wf[l]), _ptr(self.bias_padded(l)), _ptr(out),
                                         _ptr(res32) if last else None, N, H, W, cinp, coutp, int(relu), s), "conv")
            if not last:
                acts.append(out)
            cur = out
        return acts, res32

    def loss_and_grads(self, noisy, sigma, target, offset=None) -> float:
        """Mean-squared loss of (noisy - r) against target; gradients into self.grads."""
        torch = self.torch
        lib = _plib()
        s = self._stream()
        N, H, W = self._shape(noisy)
        P = N * H * W
        acts, res32 = self.forward(noisy, sigma, keep=True, offset=offset)
        loss = torch.zeros(1, device=self.dev)
        g = self._buf("dres", P * 16, torch.int16)
        _check(lib.isp_pilot_output_planes(_ptr(noisy), _ptr(res32), None, _ptr(target), _ptr(g), _ptr(loss),
                                           self.channels, N, H, W, s), "output")
        part = self._buf("partial", max(lib.isp_pilot_wgrad_partial_floats(N, H, W, c[0], c[1])
                                        for c in self.layers), torch.float32)
        for l in range(len(self.layers) - 1, -1, -1):
            cin, cout, cinp, coutp, relu = self.layers[l]
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
        """Grey network: noisy [H,W] or [N,H,W]; RGB network: [3,H,W] or
        [N,3,H,W] OPP planes (fp32 device) -> the estimate (same shape).
        sigma: one value, or one per image (0..255 units); offset: None or the
        value added to each plane of the network input (one per plane, or one
        per plane of an image; 0.5 for chroma centred on 0)."""
        torch = self.torch
        one = noisy.dim() == (2 if self.channels == 1 else 3)
        x = noisy[None] if one else noisy
        N, H, W = self._shape(x)
        C = self.channels
        sg = torch.tensor(np.broadcast_to(np.asarray(sigma, np.float32), (N,)).copy(), device=self.dev)
        off = None
        if offset is not None:
            o = np.asarray(offset, np.float32)
            o = np.broadcast_to(o.reshape(-1), (N * C,)) if o.size == N * C or o.size == 1 else \
                np.broadcast_to(o.reshape(1, C), (N, C)).reshape(-1)
            off = torch.tensor(np.ascontiguousarray(o), device=self.dev)
        x = x.contiguous()
        out = torch.empty_like(x)
        self.forward(x, sg, keep=False, offset=off, est=out)
        return out[0] if one else out

    def state(self) -> dict:
        return {"params": self.params.cpu().numpy(), "m": self.m.cpu().numpy(), "v": self.v.cpu().numpy(),
                "step": self.step, "channels": self.channels}

    def load_state(self, st: dict) -> None:
        torch = self.torch
        if int(st.get("channels", 1)) != self.channels:
            raise ValueError(f"state of a {int(st.get('channels', 1))}-plane network, this one takes {self.channels}")
        self.params.copy_(torch.from_numpy(np.asarray(st["params"], np.float32)))
        if "m" in st:
            self.m.copy_(torch.from_numpy(np.asarray(st["m"], np.float32)))
            self.v.copy_(torch.from_numpy(np.asarray(st["v"], np.float32)))
        self.step = int(st.get("step", 0))
        self.pack()

    def save(self, path, optimizer: bool = True, meta: dict | None = None) -> None:
        """Weights, and with optimizer the Adam state (moments and step
        count), to an .npz file, so that training can resume from it (numpy
        arrays only, no pickles).  meta: a JSON-able record of how the weights
        were made, kept as a string."""
        import json
        st = self.state()
        arrays = dict(params=st["params"], step=np.int64(st["step"]), channels=np.int64(self.channels))
        if optimizer:
            arrays.update(m=st["m"], v=st["v"])
        if meta is not None:
            arrays["meta"] = np.array(json.dumps(meta, sort_keys=True))
        np.savez(path, **arrays)

    @staticmethod
    def meta_of(path) -> dict:
        """The meta record save() stored with a checkpoint ({} when none)."""
        import json
        with np.load(path, allow_pickle=False) as z:
            return json.loads(str(z["meta"])) if "meta" in z.files else {}

    @classmethod
    def load(cls, path, device: int = 0) -> "PilotCNN":
        """A network restored from save(): the same inference, and, when the
        file holds the Adam state, training continues exactly where it
        stopped (otherwise it restarts Adam from zero moments)."""
        with np.load(path, allow_pickle=False) as z:
            st = {k: z[k] for k in z.files if k != "meta"}
        net = cls(device=device, channels=int(st["channels"]))
        state = {"params": st["params"], "step": int(st["step"]), "channels": int(st["channels"])}
        if "m" in st:
            state.update(m=st["m"], v=st["v"])
        else:
            state["step"] = 0
        net.load_state(state)
        return net

    def layer_params(self):
        """[(w [cout,cin,9], b [cout])] as numpy arrays (for the oracle)."""
        out = []
        for l, (cin, cout, _, _, _) in enumerate(self.layers):
            out.append((self.w(l).cpu().numpy().astype(np.float64).reshape(cout, cin, 9),
                        self.b(l).cpu().numpy().astype(np.float64)))
        return out

    def layer_grads(self):
        out = []
        for l, (cin, cout, _, _, _) in enumerate(self.layers):
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


def dihedral(t, k: int):
    """Crops [N,H,W] or [N,C,H,W] (square) under the dihedral transform k in
    0..7: a transpose when k & 4, then k & 3 quarter turns."""
    import torch
    if k & 4:
        t = t.transpose(-2, -1)
    return torch.rot90(t, k & 3, (-2, -1)) if k & 3 else t


def train(net: PilotCNN, noisy, sigma, target, steps: int, batch: int = 32, lr: float = 1e-3, seed: int = 0,
          log_every: int = 0, augment: bool = True) -> list[float]:
    """Adam on random mini-batches; returns the loss of every step.  With
    augment, each mini-batch (noisy crops and their step-1 targets alike) is
    taken under one of the 8 dihedral transforms of the square crops, drawn
    from the same seeded generator."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    n = noisy.shape[0]
    if augment and noisy.shape[-1] != noisy.shape[-2]:
        raise ValueError("dihedral augmentation needs square crops")
    losses = []
    # an RGB network's inputs are OPP planes: chroma offset by 0.5, as in rgb_guide
    off = None if net.channels == 1 else torch.tensor(list(OPP_OFFSETS) * batch, device=noisy.device)
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(noisy.device)
        a, t = noisy[idx], target[idx]
        if augment:
            k = int(torch.randint(0, 8, (1,), generator=g))
            a, t = dihedral(a, k), dihedral(t, k)
        lr_t = lr * (0.5 * (1 + math.cos(math.pi * it / max(1, steps))))  # cosine decay
        loss = net.loss_and_grads(a.contiguous(), sigma[idx].contiguous(), t.contiguous(), offset=off)
        net.adam(lr_t)
        losses.append(loss)
        if log_every and (it + 1) % log_every == 0:
            print(f"pilot step {it + 1}/{steps} loss {float(loss.item()):.3e}", file=sys.stderr, flush=True)
    return [float(v.item()) for v in losses]


# OPP norms of Y, U, V (include/BM3D.h:241-248): the noise of an OPP plane is
# sigma * norm for RGB noise sigma
OPP_NORMS = (1.0 / math.sqrt(3.0), 1.0 / math.sqrt(2.0), math.sqrt(6.0) / 4.0)
# offsets added to the OPP planes at the network input (chroma is centred on 0)
OPP_OFFSETS = (0.0, 0.5, 0.5)


def to_opp(rgb):
    """OPP planes of device RGB frames [3,H,W] or [N,3,H,W] (no clip)."""
    import torch
    lib = L.lib()
    frames = rgb[None] if rgb.dim() == 3 else rgb
    H, W = frames.shape[-2:]
    s = ctypes.c_void_p(torch.cuda.current_stream(rgb.device).cuda_stream)
    yuv = torch.empty_like(frames)
    for f, o in zip(frames, yuv):
        _check(lib.isp_bm3d_rgb_to_opp(_ptr(o[0]), _ptr(o[1]), _ptr(o[2]), _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), W, H,
                                       W, s), "rgb_to_opp")
    return yuv[0] if rgb.dim() == 3 else yuv


def from_opp(yuv, out=None):
    """RGB of device OPP planes [3,H,W] or [N,3,H,W], without clip (BM3D_Final's
    own OPP conversion of the guide then recovers the planes); into `out`
    (same shape) when given."""
    import torch
    lib = L.lib()
    frames = yuv[None] if yuv.dim() == 3 else yuv
    H, W = frames.shape[-2:]
    s = ctypes.c_void_p(torch.cuda.current_stream(yuv.device).cuda_stream)
    if out is not None and (tuple(out.shape) != tuple(yuv.shape) or not out.is_contiguous()):
        raise ValueError("out must be a contiguous tensor shaped like the planes")
    out = torch.empty_like(frames) if out is None else out.reshape(frames.shape)
    for f, o in zip(frames, out):
        _check(lib.isp_bm3d_opp_to_rgb(_ptr(o[0]), _ptr(o[1]), _ptr(o[2]), _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), W, H,
                                       W, 0, s), "opp_to_rgb")
    return out[0] if yuv.dim() == 3 else out


def rgb_guide(net: PilotCNN, rgb, sigma: float, out=None):
    """The pilot's estimate of RGB frames [3,H,W] or [N,3,H,W] (device fp32,
    noise sigma per channel in 0..255 units), as RGB without clip, into `out`
    when given.  An RGB network (channels 3) estimates the three OPP planes of
    a frame together; a grey network runs on each OPP plane alone, with that
    plane's noise level."""
    yuv = to_opp(rgb)
    if net.channels == 3:
        est = net.infer(yuv, sigma, OPP_OFFSETS)
    else:
        planes = yuv.reshape(-1, *yuv.shape[-2:])
        nfr = planes.shape[0] // 3
        est = net.infer(planes, [sigma * n for n in OPP_NORMS] * nfr, list(OPP_OFFSETS) * nfr).reshape(yuv.shape)
    return from_opp(est, out)


def rgb_para(sigma: float = 25.0, bm_range: int = 12, group_sizes=(16, 32)) -> BM3D_Para:
    """Two-step parameters of the RGB forms: profile np with BlockStep 3, BMstep 1."""
    p = BM3D_Para("np")
    for st in (p.basic, p.final):
        st.BlockStep, st.BMrange, st.BMstep = 3, bm_range, 1
    p.basic.GroupSize, p.final.GroupSize = group_sizes
    p.set_sigma(sigma)
    return p


def make_rgb_training_set(n_crops: int, crop: int = 128, sigma_range=(10.0, 40.0), seed: int = 0,
                          source: int = 512, para: BM3D_Para | None = None, device: int = 0):
    """Random crops of seeded synthetic RGB frames (the mixed-channel recipe)
    with AWGN of per-crop sigma on every channel, as OPP planes, and the OPP
    planes of their GPU BM3D step-1 estimates (RGB frames, per-image sigma) as
    targets.  Returns device tensors noisy [n,3,crop,crop] (OPP), sigma [n],
    target [n,3,crop,crop] (OPP)."""
    import torch
    from . import synth
    from .batch import BM3DBatch
    dev = f"cuda:{device}"
    rng = np.random.default_rng(seed)
    n_src = max(1, n_crops // 256)
    cleans = torch.from_numpy(np.stack([synth.clean_rgb(source, source, seed * 1000 + k) for k in range(n_src)])
                              .astype(np.float32)).to(dev)
    sig = rng.uniform(*sigma_range, n_crops).astype(np.float32)
    ys = rng.integers(0, source - crop + 1, n_crops)
    xs = rng.integers(0, source - crop + 1, n_crops)
    gen = torch.Generator(device=dev).manual_seed(seed + 7919)
    rgb = torch.empty((n_crops, 3, crop, crop), dtype=torch.float32, device=dev)
    for i in range(n_crops):
        rgb[i] = cleans[i % n_src, :, ys[i]:ys[i] + crop, xs[i]:xs[i] + crop]
    rgb += torch.randn(rgb.shape, generator=gen, device=dev) * torch.from_numpy(sig / 255.0).to(dev)[:, None, None,
                                                                                                   None]
    p = para or rgb_para()
    f = BM3D_Basic(p.basic, device=device)
    est = torch.empty_like(rgb)
    for s0 in range(0, n_crops, 2048):
        s1 = min(n_crops, s0 + 2048)
        b = BM3DBatch(f, [est[i] for i in range(s0, s1)], [rgb[i] for i in range(s0, s1)],
                      sigmas=[float(v) for v in sig[s0:s1]])
        b.run()
        torch.cuda.synchronize()
        b.close()
    return to_opp(rgb), torch.from_numpy(sig).to(dev), to_opp(est)


# the shipped networks (trained by ``python -m isp_mw_amd.train_pilot KIND``)
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
WEIGHT_FILES = {"grey": "pilot_grey.npz", "rgb": "pilot_rgb_r19.npz"}


def weights_path(kind: str) -> str:
    """Path of the shipped weights of `kind`: "grey" (grey planes, the grey
    parameters of default_para) or "rgb" (OPP planes of RGB frames, step-1
    targets at BMrange 19, GroupSize 32)."""
    return os.path.join(WEIGHTS_DIR, WEIGHT_FILES[kind])


def pretrained(kind: str = "grey", device: int = 0) -> PilotCNN:
    """The shipped network of `kind`; PilotCNN.meta_of(weights_path(kind))
    gives the recipe it was trained with and its final training loss."""
    return PilotCNN.load(weights_path(kind), device=device)


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
        if self.rgb:
            return rgb_guide(self.net, noisy, self.noise_sigma)
        return self.net.infer(noisy, self.noise_sigma)

    def __call__(self, noisy):
        import torch
        g = self.guide(noisy)
        out = torch.empty_like(noisy)
        return self.final.process(out, noisy, g)


__all__ = ["PilotCNN", "PilotBM3D", "make_training_set", "make_rgb_training_set", "train", "dihedral",
           "pretrained", "weights_path", "default_para",
           "rgb_para", "rgb_guide", "to_opp", "from_opp", "layers", "LAYERS", "GREY_SIGMA_SCALE", "OPP_NORMS",
           "OPP_OFFSETS", "BM3D"]
