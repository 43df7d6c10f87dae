"""FP64 numpy oracle of the learned pilot network (test infrastructure only).

Only tests/ and ``__graft_entry__.smoke()`` may import it; the product
package never does.  It restates the pilot's arithmetic in plain numpy: 3x3
cross-correlation with zero padding, ReLU, the residual output
``out = noisy - r``, the mean-squared loss and the analytic backward pass.
The pilot has no counterpart in ISP_MW (it stands in for BM3D_Basic's estimate,
source/BM3D.cpp:339-385, as the guide of BM3D_Final).

With ``bf16=True`` every tensor the GPU stores in BF16 (inputs, weights,
activations, gradients of activations) is rounded to BF16 (round to nearest
even) at the same place, so that the remaining difference to the GPU is FP32
versus FP64 accumulation.

Pins (tests/test_oracle.py): ``conv3x3`` against scipy.signal.correlate2d;
``loss_and_grads`` against central finite differences of ``loss``.
"""
from __future__ import annotations

import numpy as np


def layers(channels: int = 1):
    """(cin, cout, relu) of the 8 layers of a network on images of `channels`
    planes: the planes + a sigma map in, one residual per plane out."""
    return [(channels + 1, 48, True)] + [(48, 48, True)] * 6 + [(48, channels, False)]


LAYERS = layers(1)


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3x3(x: np.ndarray, w: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """x [N,H,W,Cin], w [Cout,Cin,9] (tap = ky*3+kx) ->
    out[n,y,x,co] = b[co] + sum w[co,ci,ky*3+kx] * x[n,y+ky-1,x+kx-1,ci]."""
    N, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((N, H, W, w.shape[0]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        out += np.einsum("nhwc,oc->nhwo", xp[:, ky:ky + H, kx:kx + W, :], w[:, :, tap])
    if b is not None:
        out += b
    return out


def conv3x3_wgrad(x: np.ndarray, dz: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """dw[co,ci,tap] = sum dz[n,y,x,co] * x[n,y+ky-1,x+kx-1,ci]; db[co] = sum dz."""
    N, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros((dz.shape[-1], x.shape[-1], 9))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dw[:, :, tap] = np.einsum("nhwo,nhwc->oc", dz, xp[:, ky:ky + H, kx:kx + W, :])
    return dw, dz.sum(axis=(0, 1, 2))


def conv3x3_dinput(dz: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Gradient with respect to the input of conv3x3: the correlation of dz
    with the flipped, transposed kernel."""
    wf = np.transpose(w, (1, 0, 2))[:, :, ::-1]
    return conv3x3(dz, np.ascontiguousarray(wf))


def planes(noisy) -> np.ndarray:
    """[N,H,W] (one plane per image) or [N,C,H,W] -> [N,C,H,W] float64."""
    a = np.asarray(noisy, np.float64)
    return a[:, None] if a.ndim == 3 else a


def net_input(noisy: np.ndarray, sigma: np.ndarray, offset=None) -> np.ndarray:
    """Noisy planes [N,H,W] or [N,C,H,W] (plus offset [N,C] per plane, or
    None) and sigma per image (0..255) -> [N,H,W,C+1]."""
    x = planes(noisy)
    if offset is not None:
        x = x + np.asarray(offset, np.float64).reshape(x.shape[0], x.shape[1])[:, :, None, None]
    N, C, H, W = x.shape
    s = np.broadcast_to((np.asarray(sigma, np.float64) / 255.0)[:, None, None, None], (N, 1, H, W))
    return np.moveaxis(np.concatenate([x, s], axis=1), 1, -1)


def forward(params, noisy, sigma, bf16: bool = False, offset=None):
    """params: list of (w [cout,cin,9], b [cout]).  Returns (out, acts, r)
    with acts[0] the input and acts[l+1] the output of layer l; out and r
    have the shape of noisy."""
    rnd = bf16_round if bf16 else (lambda a: a)
    C = params[-1][0].shape[0]
    a = rnd(net_input(noisy, sigma, offset))
    acts = [a]
    r = None
    for (w, b), (_, _, relu) in zip(params, layers(C)):
        z = conv3x3(a, rnd(w), b)
        if relu:
            a = rnd(np.maximum(z, 0.0))
            acts.append(a)
        else:
            r = np.moveaxis(z, -1, 1)  # [N,C,H,W]
    r = r.reshape(np.shape(noisy))
    out = np.asarray(noisy, np.float64) - r
    return out, acts, r


def loss(params, noisy, sigma, target, bf16: bool = False, offset=None) -> float:
    out, _, _ = forward(params, noisy, sigma, bf16, offset)
    return float(np.mean((out - target) ** 2))


def loss_and_grads(params, noisy, sigma, target, bf16: bool = False, offset=None):
    """Loss and [(dw, db)] per layer (backward of forward())."""
    rnd = bf16_round if bf16 else (lambda a: a)
    C = params[-1][0].shape[0]
    out, acts, _ = forward(params, noisy, sigma, bf16, offset)
    d = out - target
    L = float(np.mean(d ** 2))
    g = rnd(np.moveaxis(planes(-2.0 * d / d.size), 1, -1))     # d loss / d r, [N,H,W,C]
    grads = [None] * len(params)
    net = layers(C)
    for l in range(len(params) - 1, -1, -1):
        w, _ = params[l]
        _, _, relu = net[l]
        dz = g * (acts[l + 1] > 0) if relu else g
        grads[l] = conv3x3_wgrad(acts[l], dz)
        if l > 0:
            g = rnd(conv3x3_dinput(dz, rnd(w)))
    return L, grads
