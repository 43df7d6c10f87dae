"""Column-walk matching (match_walk_kernel) against the list kernel
(match_kernel) in the FP32-distance forms.

Both sum a candidate's distance the same way (each row's 8 fused terms from 0,
then the row sums in row order), so they must give identical groups: counts,
positions, order and keys, bit for bit, and therefore identical outputs.  The
environment variable ISP_BM3D_MATCH_WALK=0 at context creation keeps every
search on the list kernel.  The walk takes BlockStep 3..7 (smaller
steps stay on the list kernel, so both runs agree trivially); windows beyond
128 lane strips are walked in chunks, each block keeping its best between
chunks.  Cases cover
block steps 1..7, BMsteps 1..3, ranges, group sizes 2..32, grids that end in
clamped positions, windows clipped on every side, columns of 4 to 32 blocks
per tile, quantised images full of equal distances, RGB frames, batches,
multi-frame stacks and NL-means."""
import numpy as np
import pytest

from isp_mw_amd import BM3D, BM3DBatch, NLMeans, NLMeans_Para, synth
from tests.gpu_common import para

pytestmark = pytest.mark.gpu


def with_walk(monkeypatch, walk, fn):
    if walk:
        monkeypatch.delenv("ISP_BM3D_MATCH_WALK", raising=False)
    else:
        monkeypatch.setenv("ISP_BM3D_MATCH_WALK", "0")
    try:
        return fn()
    finally:
        monkeypatch.delenv("ISP_BM3D_MATCH_WALK", raising=False)


def quantised(h, w, levels, seed):
    """Synthetic image on `levels` grey levels, no noise: many equal distances
    and identical blocks (distance 0, never candidates)."""
    c = synth.clean_grey(h, w, seed)
    return (np.floor(c * levels) / levels).astype(np.float32)


def assert_same_groups(a, b, label):
    (c0, p0, k0), (c1, p1, k1) = a, b
    np.testing.assert_array_equal(c0, c1, err_msg=label)
    live = np.arange(p0.shape[-1])[None, None, :] < c0[..., None]
    np.testing.assert_array_equal(np.where(live, p0, -1), np.where(live, p1, -1), err_msg=label)
    np.testing.assert_array_equal(np.where(live, k0, -1.0), np.where(live, k1, -1.0), err_msg=label)
    print(f"{label}: {c0.size} groups, mean size {c0.mean():.2f}, identical", flush=True)


# (h, w, BlockStep, BMrange, BMstep, G1, G2)
CASES = [
    (64, 80, 3, 12, 1, 16, 32),
    (61, 77, 3, 12, 1, 16, 32),    # clamped last grid row and column
    (50, 70, 1, 6, 1, 16, 16),
    (57, 66, 2, 8, 2, 8, 16),
    (45, 93, 5, 9, 3, 16, 32),
    (70, 60, 6, 9, 1, 32, 32),
    (20, 24, 3, 12, 1, 16, 32),    # window larger than the image
    (33, 41, 4, 10, 1, 2, 4),      # K = 1 and 3
    (75, 66, 7, 8, 2, 16, 16),
    (100, 37, 3, 12, 1, 16, 32),   # columns of 30 blocks: two tiles, the second clamped
    (40, 40, 4, 16, 1, 16, 32),    # 231 strips: walked in two chunks
    (70, 90, 3, 19, 1, 16, 32),    # 312 strips: three chunks
    (61, 83, 4, 19, 1, 32, 32),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}_s{c[2]}_R{c[3]}_st{c[4]}_G{c[5]}" for c in CASES])
@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_groups_identical(gpu, monkeypatch, case, prec):
    h, w, step, R, bmstep, G1, G2 = case
    p = para(step=step, R=R, bmstep=bmstep, G1=G1, G2=G2)
    _, n = synth.grey_pair(h, w, 25.0, h + w)
    print(f"case {case} {prec}", flush=True)

    def go():
        f = BM3D(p, precision=prec)
        g = [f.match(n, final_step=s) for s in (False, True)]
        return g, f(n)

    (g0, out0) = with_walk(monkeypatch, False, go)
    (g1, out1) = with_walk(monkeypatch, True, go)
    for s in (0, 1):
        assert_same_groups(g0[s], g1[s], f"{case} step {s + 1}")
    np.testing.assert_array_equal(out0, out1)


@pytest.mark.parametrize("tile", [(8, 4), (2, 32), (8, 8), (4, 16)])
def test_tile_shapes_do_not_change_groups(gpu, monkeypatch, tile):
    # walks of 4, 32, 8 and 16 blocks per column against the list kernel
    p = para()
    _, n = synth.grey_pair(90, 70, 25.0, 8)
    print(f"match tile {tile}", flush=True)

    def go():
        f = BM3D(p, precision="fp16")
        f.set_tuning(f.size_class(70, 90), {"match_tile": tile, "filter_tile": (8, 8), "final_group_size": 0})
        return [f.match(n, final_step=s) for s in (False, True)]

    g0 = with_walk(monkeypatch, False, go)
    g1 = with_walk(monkeypatch, True, go)
    for s in (0, 1):
        assert_same_groups(g0[s], g1[s], f"tile {tile} step {s + 1}")


@pytest.mark.parametrize("R", [12, 19])
@pytest.mark.parametrize("levels", [4, 16, 255])
def test_equal_distances(gpu, monkeypatch, levels, R):
    # quantised, noise-free images: long runs of equal keys exercise the
    # selection's index order on ties (at R 19 also across chunks)
    p = para(R=R, G1=16, G2=32)
    img = quantised(72, 96, levels, levels)

    def go():
        f = BM3D(p, precision="fp16")
        return [f.match(img, final_step=s) for s in (False, True)]

    g0 = with_walk(monkeypatch, False, go)
    g1 = with_walk(monkeypatch, True, go)
    for s in (0, 1):
        assert_same_groups(g0[s], g1[s], f"{levels} levels step {s + 1}")


def test_rgb_frame(gpu, monkeypatch):
    p = para()
    _, c = synth.rgb_pair(67, 85, 25.0, 5)
    shape = BM3D(p).grid_shape(85, 67)

    def go():
        f = BM3D(p, precision="fp16")
        out = f(c)
        return out, [f.last_groups(s, shape, with_pos=True) for s in (False, True)]

    out0, g0 = with_walk(monkeypatch, False, go)
    out1, g1 = with_walk(monkeypatch, True, go)
    np.testing.assert_array_equal(out0, out1)
    for (c0, p0), (c1, p1) in zip(g0, g1):
        # slots past a group's count hold whatever the buffer held
        np.testing.assert_array_equal(c0, c1)
        live = np.arange(p0.shape[-1])[None, None, :] < c0[..., None]
        np.testing.assert_array_equal(np.where(live, p0, -1), np.where(live, p1, -1))


def test_batch(gpu, monkeypatch):
    import torch
    sizes = [(64, 80), (40, 56), (33, 47)]
    imgs = [synth.grey_pair(h, w, 25.0, 7 + k)[1] for k, (h, w) in enumerate(sizes)]

    def go():
        f = BM3D(para(), precision="fp16")
        srcs = [torch.from_numpy(a).to(gpu) for a in imgs]
        dsts = [torch.empty_like(t) for t in srcs]
        BM3DBatch(f, dsts, srcs).run()
        torch.cuda.synchronize()
        return [d.cpu().numpy() for d in dsts]

    a = with_walk(monkeypatch, False, go)
    b = with_walk(monkeypatch, True, go)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_stack(gpu, monkeypatch):
    # the reference frame's own list takes the walk; other frames' lists the
    # cross-frame list kernel; their keys compare alike in the merge
    rng = np.random.default_rng(3)
    base = synth.clean_grey(56, 72, 3)
    g = np.stack([np.roll(base, t, axis=1) for t in range(3)])
    g = (g + rng.normal(0, 25 / 255.0, g.shape)).astype(np.float32)

    def go():
        return BM3D(para(R=8), precision="fp16").process_stack(np.empty_like(g), g, radius=1)

    np.testing.assert_array_equal(with_walk(monkeypatch, False, go), with_walk(monkeypatch, True, go))


@pytest.mark.parametrize("kw", [dict(), dict(BlockStep=3, BMstep=1, BMrange=8, GroupSize=8),
                                dict(BlockStep=2, BMstep=2, BMrange=12, GroupSize=32)])
def test_nlm(gpu, monkeypatch, kw):
    para_ = NLMeans_Para(sigma=12.0, **kw)
    _, n = synth.rgb_pair(60, 84, 12.0, 11)
    src = np.clip(np.rint(n * 255), 0, 255).astype(np.int32)

    def go():
        return NLMeans(para_, precision="fp16")(src, floor=0, ceil=255)

    np.testing.assert_array_equal(with_walk(monkeypatch, False, go), with_walk(monkeypatch, True, go))
