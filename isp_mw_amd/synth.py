"""Seeded synthetic test images for the BM3D benchmark workloads.

Grey recipe (one plane, from one seed; numpy PCG64 draws in this order):
  1. gx, gy ~ U[-0.3, 0.3]; img = 0.25 + gx (x/W - 1/2) + gy (y/H - 1/2)
  2. 20 rectangles painted in turn: x0 ~ U[0, W), y0 ~ U[0, H),
     w ~ U[W/32, W/4], h ~ U[H/32, H/4], value ~ U[0, 1]
  3. 4 gratings a sin(2 pi (x cos t + y sin t) / T + phi) added, with
     a ~ U[0.05, 0.2], T ~ U[4, 32], t ~ U[0, pi), phi ~ U[0, 2 pi)
  4. clip to [0, 1]
  5. AWGN with standard deviation sigma/255 from a second generator seeded
     seed + 7919; the noisy image is NOT clipped (Plane_FL input is never
     clipped by the reference).
RGB recipe: channels c = 0, 1, 2 from seeds seed + c (steps 1-4), mixed by a
random positive 3x3 matrix (entries U[0.1, 1] from seed + 104729, rows
normalised to sum 1), then noise per channel from seed + 7919.
"""
from __future__ import annotations

import numpy as np


def clean_grey(h: int, w: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    gx, gy = rng.uniform(-0.3, 0.3, 2)
    img = 0.25 + gx * (x / w - 0.5) + gy * (y / h - 0.5)
    for _ in range(20):
        x0 = int(rng.uniform(0, w))
        y0 = int(rng.uniform(0, h))
        rw = max(1, int(rng.uniform(w / 32, w / 4)))
        rh = max(1, int(rng.uniform(h / 32, h / 4)))
        img[y0:y0 + rh, x0:x0 + rw] = rng.uniform(0, 1)
    for _ in range(4):
        a = rng.uniform(0.05, 0.2)
        T = rng.uniform(4, 32)
        t = rng.uniform(0, np.pi)
        phi = rng.uniform(0, 2 * np.pi)
        img += a * np.sin(2 * np.pi * (x * np.cos(t) + y * np.sin(t)) / T + phi)
    return np.clip(img, 0.0, 1.0)


def noisy(clean: np.ndarray, sigma: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed + 7919)
    return (clean + rng.standard_normal(clean.shape) * (sigma / 255.0)).astype(np.float32)


def grey_pair(h: int, w: int, sigma: float, seed: int):
    """(clean float64, noisy float32) single-channel pair."""
    c = clean_grey(h, w, seed)
    return c, noisy(c, sigma, seed)


def clean_rgb(h: int, w: int, seed: int) -> np.ndarray:
    chans = np.stack([clean_grey(h, w, seed + c) for c in range(3)])
    m = np.random.default_rng(seed + 104729).uniform(0.1, 1.0, (3, 3))
    m /= m.sum(axis=1, keepdims=True)
    return np.einsum("ij,jhw->ihw", m, chans)


def rgb_pair(h: int, w: int, sigma: float, seed: int):
    """(clean float64 [3,H,W], noisy float32 [3,H,W]) RGB pair."""
    c = clean_rgb(h, w, seed)
    return c, noisy(c, sigma, seed)


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)
