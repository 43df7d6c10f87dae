"""Pins of the oracle's multi-frame grouping (orc_match_stack / orc_kernel_stack
/ orc_bm3d_stack: Pos3 groups over a frame stack, include/Block.h:843-846,
927-936, 1141-1208): the single-frame special cases against the pinned
single-frame oracle, an independent numpy formulation (tests/npref.py
brute_groups_stack / stack_step), and closed forms.  CPU only."""
import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D_Para, synth
from tests import npref


def para(sigma=25.0, R=8, step=3, G1=16, G2=32):
    p = BM3D_Para("np")
    for st in (p.basic, p.final):
        st.BlockStep, st.BMrange = step, R
    p.basic.GroupSize, p.final.GroupSize = G1, G2
    return p.set_sigma(sigma)


def static_stack(T, h, w, sigma, seed, shift=(0, 0)):
    """One clean scene seen T times (frame t moved by t*shift), independent noise."""
    big = synth.clean_grey(h + 8 * T, w + 8 * T, seed)
    rng = np.random.default_rng(seed)
    cl = np.stack([big[4 + t * shift[0]:4 + t * shift[0] + h, 4 + t * shift[1]:4 + t * shift[1] + w]
                   for t in range(T)])
    return cl, cl + rng.normal(0, sigma / 255.0, cl.shape)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_one_frame_is_the_single_frame_filter(mode):
    p = para()
    cl, n = static_stack(1, 40, 48, 25.0, 3)
    ref = cl if mode == 2 else None
    got = oracle.bm3d_stack(p.basic, p.final, n, radius=2, ref=ref, mode=mode)
    if mode == 3:
        want = oracle.bm3d_plane(p.basic, p.final, n[0])
    else:
        st = p.basic if mode == 1 else p.final
        want = oracle.kernel(st, n[0], cl[0] if mode == 2 else n[0], wiener=mode == 2)
    np.testing.assert_array_equal(got[0], want)
    rgb = np.stack([n[0], 0.8 * n[0] + 0.1, 1 - n[0]])[None]
    np.testing.assert_array_equal(oracle.bm3d_stack(p.basic, p.final, rgb, radius=1)[0],
                                  oracle.bm3d_rgb(p.basic, p.final, rgb[0]))


def test_radius_zero_is_per_frame():
    p = para()
    _, n = static_stack(3, 32, 40, 25.0, 5)
    got = oracle.bm3d_stack(p.basic, p.final, n, radius=0)
    for t in range(3):
        np.testing.assert_array_equal(got[t], oracle.bm3d_plane(p.basic, p.final, n[t]))


@pytest.mark.parametrize("radius,t", [(1, 0), (1, 1), (2, 2), (2, 0)])
def test_groups_against_numpy_restatement(radius, t):
    p = para(R=6)
    _, n = static_stack(4, 30, 34, 25.0, 7, shift=(1, 2))
    cnt, pos, fz = oracle.match_stack(p.basic, n, radius, t)
    want = npref.brute_groups_stack(p.basic, n, radius, t)
    gy, gx = npref.grid(30, 3), npref.grid(34, 3)
    multi = 0
    for a, j in enumerate(gy):
        for b, i in enumerate(gx):
            g = want[(j, i)]
            assert cnt[a, b] == len(g)
            got = [(int(fz[a, b, k]), int(pos[a, b, k]) >> 16, int(pos[a, b, k]) & 0xFFFF) for k in range(len(g))]
            assert got == [(z, y, x) for _, z, y, x in g]
            multi += any(z != t for _, z, _, _ in g)
    assert multi > 0.5 * len(gy) * len(gx)      # most groups take members from other frames


@pytest.mark.parametrize("wiener", [False, True])
def test_step_against_numpy_restatement(wiener):
    p = para(R=6)
    cl, n = static_stack(3, 30, 34, 25.0, 9, shift=(0, 1))
    st = p.final if wiener else p.basic
    ref = cl + 0.01 * np.sin(np.arange(34))[None, None, :] if wiener else n
    got = oracle.kernel_stack(st, n, ref, radius=1, wiener=wiener)
    want = npref.stack_step(st, n, ref, radius=1, wiener=wiener)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)


def test_a_translated_copy_is_the_best_match():
    # frame 1 = frame 0 moved by (2, 3) and offset by 1e-3: the copy of each
    # reference block (distance 64e-6, far below any noisy candidate) is its
    # second member wherever the copy lies inside the window
    p = para(R=8, step=4)
    rng = np.random.default_rng(1)
    big = synth.clean_grey(60, 70, 2) + rng.normal(0, 25 / 255.0, (60, 70))
    f0 = big[4:44, 4:52]
    f1 = big[2:42, 1:49] + 1e-3            # f1[y, x] = f0[y - 2, x - 3] + 1e-3 where both exist
    stack = np.stack([f0, f1])
    cnt, pos, fz = oracle.match_stack(p.basic, stack, 1, 0)
    checked = 0
    for a, j in enumerate(npref.grid(40, 4)):
        for b, i in enumerate(npref.grid(48, 4)):
            y, x = j + 2, i + 3
            if y > 40 - 8 or x > 48 - 8 or cnt[a, b] < 2:
                continue
            assert (fz[a, b, 1], pos[a, b, 1] >> 16, pos[a, b, 1] & 0xFFFF) == (1, y, x)
            checked += 1
    assert checked > 40


def test_temporal_grouping_helps_a_static_scene():
    p = para()
    cl, n = static_stack(3, 48, 64, 25.0, 3)
    r0 = oracle.bm3d_stack(p.basic, p.final, n, radius=0)
    r1 = oracle.bm3d_stack(p.basic, p.final, n, radius=1)
    gain = np.mean([synth.psnr(r1[t], cl[t]) - synth.psnr(r0[t], cl[t]) for t in range(3)])
    assert gain > 0.1
