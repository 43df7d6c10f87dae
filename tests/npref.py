"""Independent numpy formulation of the BM3D collaborative filter, for
pinning the oracle (test code only).

It does NOT restate the reference line by line.  It uses orthonormal DCTs
(scipy.fft.dctn norm="ortho") and the closed forms that the reference's
unnormalised FFTW tables reduce to (source/BM3D.cpp:26-70, 339-457):
  step 1: zero an orthonormal coefficient iff |O| <= lambda*sigma_c; weight
          1/retained (1 if none);
  step 2: W = O_r^2 / (O_r^2 + sigma_c^2 / 2^zeta), zeta = number of zero
          indices (z, u, v); weight 1/sum(W^2) (1 if 0);
  num += weight * block, den += weight over every block of the group.
Groups come from a brute-force window search (brute_groups).
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sf

NORM = np.array([np.sqrt(1 / 3), np.sqrt(1 / 2), np.sqrt(3 / 8)])
B = 8


def grid(n, s, b=B):
    last = n - b
    out, j = [], 0
    while True:
        if j >= last + s:
            break
        if j > last:
            j = last
        out.append(j)
        j += s
    return out


def brute_groups(p, ref, norm=NORM[0]):
    """All candidates of the clipped window (BMstep grid through the reference
    position), SSE <= thSSE and != 0, sorted by (key, scan order), self first,
    truncated to GroupSize.  Returns {(j, i): [(key, y, x), ...]}.  thMSE is
    scaled by `norm` (BM3D's luma OPP norm; NL-means passes 1).  The block
    size is p.BlockSize (8 when p has none)."""
    B = int(getattr(p, "BlockSize", 8))
    ref = np.asarray(ref, np.float64)
    h, w = ref.shape
    st = p.BMstep
    R = p.BMrange // st * st
    thMSE = p.thMSE * norm
    mse2sse = B * B / 255.0 ** 2
    th = thMSE * mse2sse
    out = {}
    for j in grid(h, p.BlockStep, B):
        for i in grid(w, p.BlockStep, B):
            rb = ref[j:j + B, i:i + B]
            if p.GroupSize == 1:
                out[(j, i)] = [(0.0, j, i)]
                continue
            ys = [y for y in range(j - R, j + R + 1, st) if 0 <= y <= h - B]
            xs = [x for x in range(i - R, i + R + 1, st) if 0 <= x <= w - B]
            cands = []
            for y in ys:
                for x in xs:
                    if (y, x) == (j, i):
                        continue
                    d = float(np.sum((rb - ref[y:y + B, x:x + B]) ** 2))
                    if d <= th and d != 0:
                        cands.append((d / mse2sse, len(cands), y, x))
            cands.sort(key=lambda t: (t[0], t[1]))
            g = [(0.0, j, i)] + [(k, y, x) for k, _, y, x in cands]
            out[(j, i)] = g[:p.GroupSize] if p.GroupSize > 0 else g
    return out


def haar_basis(n):
    """Orthonormal Haar matrix of a power-of-two n from its basis functions:
    row 0 the constant 1/sqrt(n); row 2^j + m (0 <= m < 2^j) supported on
    [m n / 2^j, (m + 1) n / 2^j), +sqrt(2^j / n) on the first half and
    -sqrt(2^j / n) on the second."""
    h = np.zeros((n, n))
    h[0] = 1.0 / np.sqrt(n)
    k = 1
    j = 0
    while k < n:
        for m in range(2 ** j):
            L = n // 2 ** j
            a = m * L
            h[k, a:a + L // 2] = np.sqrt(2 ** j / n)
            h[k, a + L // 2:a + L] = -np.sqrt(2 ** j / n)
            k += 1
        j += 1
    return h


def filter_plane(p, src, ref, groups, c, wiener, haar=False, kaiser_beta=0.0):
    """Aggregated num/den of one plane for the given groups.  haar: the group
    axis by the Haar matrix, groups truncated to a power of two; kaiser_beta:
    blocks aggregated with numpy's Kaiser window of that shape."""
    B = int(getattr(p, "BlockSize", 8))
    win = np.outer(np.kaiser(B, kaiser_beta), np.kaiser(B, kaiser_beta)) if kaiser_beta else np.ones((B, B))
    src = np.asarray(src, np.float64)
    ref = np.asarray(ref, np.float64)
    h, w = src.shape
    sigma = p.sigma[c] / 255.0 * NORM[c]
    num = np.zeros((h, w))
    den = np.zeros((h, w))
    for g in groups.values():
        if haar:
            g = g[:2 ** int(np.floor(np.log2(len(g))))]
        n = len(g)
        grp = np.stack([src[y:y + B, x:x + B] for _, y, x in g])
        if haar:
            O = np.einsum("kz,zuv->kuv", haar_basis(n), sf.dctn(grp, type=2, norm="ortho", axes=(1, 2)))
        else:
            O = sf.dctn(grp, type=2, norm="ortho")
        if not wiener:
            keep = np.abs(O) > p.lambda_ * sigma
            O = np.where(keep, O, 0.0)
            r = int(keep.sum())
            wgt = 1.0 if r < 1 else 1.0 / r
        else:
            gr = np.stack([ref[y:y + B, x:x + B] for _, y, x in g])
            if haar:
                Or = np.einsum("kz,zuv->kuv", haar_basis(n), sf.dctn(gr, type=2, norm="ortho", axes=(1, 2)))
            else:
                Or = sf.dctn(gr, type=2, norm="ortho")
            zeta = np.zeros((n, B, B))
            zeta[0] += 1
            zeta[:, 0, :] += 1
            zeta[:, :, 0] += 1
            Wc = Or ** 2 / (Or ** 2 + sigma ** 2 / 2.0 ** zeta)
            O = O * Wc
            l2 = float((Wc ** 2).sum())
            wgt = 1.0 if l2 <= 0 else 1.0 / l2
        if haar:
            blk = sf.idctn(np.einsum("kz,kuv->zuv", haar_basis(n), O), type=2, norm="ortho", axes=(1, 2))
        else:
            blk = sf.idctn(O, type=2, norm="ortho")
        for z, (_, y, x) in enumerate(g):
            num[y:y + B, x:x + B] += wgt * win * blk[z]
            den[y:y + B, x:x + B] += wgt * win
    return num, den


def kernel(p, src_planes, ref_planes, wiener, haar=False, kaiser_beta=0.0):
    """BM3D_Base::Kernel semantics: match on ref_planes[0], filter each plane
    with sigma > 0, others copied from src."""
    groups = brute_groups(p, ref_planes[0])
    out = []
    for c, s in enumerate(src_planes):
        if p.sigma[c] > 0:
            num, den = filter_plane(p, s, ref_planes[c], groups, c, wiener, haar, kaiser_beta)
            out.append(num / den)
        else:
            out.append(np.asarray(s, np.float64).copy())
    return out


def nlm(p, src_planes, ref, value_range=1.0):
    """NL-means from its definition (source/NLMeans.cpp:6-375) on the groups
    of brute_groups with the unscaled thMSE: per group, weights
    w_k = exp(-key_k / strength^2) and the weighted moments
    EX = sum w X / sum w, VarX = sum w X^2 / sum w - EX^2 as array reductions;
    the block estimate is EX (no correction) or, with correction,
    X - VarN/VarX (X - EX) where VarX > VarN (VarN = (sigma * range/255)^2)
    and EX elsewhere; a group of one returns the source block.  Every grid
    block adds its estimate with weight 1.  Returns (num, den) per plane."""
    groups = brute_groups(p, ref, norm=1.0)
    h, w = np.asarray(ref).shape
    src_planes = [np.asarray(s_, np.float64) for s_ in src_planes]
    nums = [np.zeros((h, w)) for _ in src_planes]
    den = np.zeros((h, w))
    varN = (p.sigma * value_range / 255.0) ** 2
    for (j, i), g in groups.items():
        wts = np.exp(-np.array([k for k, _, _ in g]) / p.strength ** 2)
        for c, s_ in enumerate(src_planes):
            X = s_[j:j + B, i:i + B]
            if len(g) < 2:
                est = X
            else:
                blocks = np.stack([s_[y:y + B, x:x + B] for _, y, x in g])
                EX = np.tensordot(wts, blocks, 1) / wts.sum()
                if p.correction:
                    VarX = np.tensordot(wts, blocks ** 2, 1) / wts.sum() - EX ** 2
                    est = np.where(VarX > varN, X - varN / np.where(VarX > 0, VarX, 1) * (X - EX), EX)
                else:
                    est = EX
            nums[c][j:j + B, i:i + B] += est
        den[j:j + B, i:i + B] += 1
    return nums, den


def brute_groups_stack(p, guides, radius, t, norm=NORM[0]):
    """Multi-frame groups of frame t over guides [T, H, W]: every window
    position of every frame z with |z - t| <= radius, frames visited in the
    order t, t-1, t+1, t-2, ..., row-major within a frame; the same
    acceptance test and (key, scan order) sort as brute_groups.  Returns
    {(j, i): [(key, z, y, x), ...]}."""
    guides = np.asarray(guides, np.float64)
    T, h, w = guides.shape
    st = p.BMstep
    R = p.BMrange // st * st
    th = p.thMSE * norm * (B * B / 255.0 ** 2)
    mse2sse = B * B / 255.0 ** 2
    order = [t] + [z for d in range(1, radius + 1) for z in (t - d, t + d) if 0 <= z < T]
    out = {}
    for j in grid(h, p.BlockStep):
        for i in grid(w, p.BlockStep):
            rb = guides[t, j:j + B, i:i + B]
            if p.GroupSize == 1:
                out[(j, i)] = [(0.0, t, j, i)]
                continue
            ys = [y for y in range(j - R, j + R + 1, st) if 0 <= y <= h - B]
            xs = [x for x in range(i - R, i + R + 1, st) if 0 <= x <= w - B]
            cands = []
            for z in order:
                # all window blocks of frame z at once: SSE by array reduction
                blocks = np.stack([guides[z, y:y + B, x:x + B] for y in ys for x in xs])
                d = ((blocks - rb) ** 2).reshape(len(ys) * len(xs), -1).sum(1)
                for k, (y, x) in enumerate([(y, x) for y in ys for x in xs]):
                    if z == t and (y, x) == (j, i):
                        continue
                    if d[k] <= th and d[k] != 0:
                        cands.append((d[k] / mse2sse, len(cands), z, y, x))
            cands.sort(key=lambda c: (c[0], c[1]))
            g = [(0.0, t, j, i)] + [(k, z, y, x) for k, _, z, y, x in cands]
            out[(j, i)] = g[:p.GroupSize] if p.GroupSize > 0 else g
    return out


def stack_step(p, src, ref, radius, wiener):
    """One step of multi-frame BM3D on grey stacks src / ref [T, H, W] with the
    orthonormal closed forms of filter_plane; every member aggregated into its
    own frame.  Returns the [T, H, W] estimate."""
    src = np.asarray(src, np.float64)
    ref = np.asarray(ref, np.float64)
    T, h, w = src.shape
    sigma = p.sigma[0] / 255.0 * NORM[0]
    num = np.zeros((T, h, w))
    den = np.zeros((T, h, w))
    for t in range(T):
        for g in brute_groups_stack(p, ref, radius, t).values():
            n = len(g)
            O = sf.dctn(np.stack([src[z, y:y + B, x:x + B] for _, z, y, x in g]), type=2, norm="ortho")
            if not wiener:
                keep = np.abs(O) > p.lambda_ * sigma
                O = np.where(keep, O, 0.0)
                r = int(keep.sum())
                wgt = 1.0 if r < 1 else 1.0 / r
            else:
                Or = sf.dctn(np.stack([ref[z, y:y + B, x:x + B] for _, z, y, x in g]), type=2, norm="ortho")
                zeta = np.zeros((n, B, B))
                zeta[0] += 1
                zeta[:, 0, :] += 1
                zeta[:, :, 0] += 1
                Wc = Or ** 2 / (Or ** 2 + sigma ** 2 / 2.0 ** zeta)
                O = O * Wc
                l2 = float((Wc ** 2).sum())
                wgt = 1.0 if l2 <= 0 else 1.0 / l2
            blk = sf.idctn(O, type=2, norm="ortho")
            for m, (_, z, y, x) in enumerate(g):
                num[z, y:y + B, x:x + B] += wgt * blk[m]
                den[z, y:y + B, x:x + B] += wgt
    return num / den
