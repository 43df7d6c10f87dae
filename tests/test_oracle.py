"""Pins the FP64 oracle (oracle/bm3d_oracle.c) against what fixes it
independently of its own code: library DCTs, brute-force matching, an
orthonormal numpy formulation of the filter, closed forms and identities."""
import numpy as np
import pytest
import scipy.fft as sf

import oracle
from isp_mw_amd import BM3D_Para, synth
from tests import npref


def para(profile="np", sigma=25.0, step=3, R=12, G1=16, G2=32, bmstep=1):
    p = BM3D_Para(profile)
    for st in (p.basic, p.final):
        st.BlockStep, st.BMrange, st.BMstep = step, R, bmstep
    p.basic.GroupSize, p.final.GroupSize = G1, G2
    p.set_sigma(sigma)
    return p


# --- transforms: FFTW REDFT10 / REDFT01 == scipy DCT-II / DCT-III, norm=None
@pytest.mark.parametrize("n0", list(range(1, 33)))
def test_dct3d_matches_scipy(n0):
    a = np.random.default_rng(n0).standard_normal((n0, 8, 8))
    f = oracle.dct3d(a)
    np.testing.assert_allclose(f, sf.dctn(a, type=2, norm=None), rtol=0, atol=1e-12 * np.abs(f).max())
    i = oracle.dct3d(a, inverse=True)
    np.testing.assert_allclose(i, sf.dctn(a, type=3, norm=None), rtol=0, atol=1e-12 * np.abs(i).max())
    # REDFT01(REDFT10(x)) = finalAMP * x with finalAMP = 2N*2B*2B (source/BM3D.cpp:26)
    rt = oracle.dct3d(f, inverse=True)
    np.testing.assert_allclose(rt, 8 * n0 * 64 * a, rtol=0, atol=1e-10 * n0 * 512)


def test_dct3d_odd_shapes():
    a = np.random.default_rng(7).standard_normal((5, 12, 7))
    np.testing.assert_allclose(oracle.dct3d(a), sf.dctn(a, type=2), atol=1e-11)


def test_dct_known_answer():
    # REDFT10 of a unit impulse at j: Y_k = 2 cos(pi (2j+1) k / (2n)) (long double)
    n = 6
    a = np.zeros((n, 1, 1))
    a[2, 0, 0] = 1.0
    y = oracle.dct3d(a)[:, 0, 0] / 4.0  # the two length-1 axes each multiply by 2
    k = np.arange(n, dtype=np.longdouble)
    ref = 2 * np.cos(np.pi * (2 * 2 + 1) * k / (2 * n))
    np.testing.assert_allclose(y, ref.astype(np.float64), atol=1e-15)


# --- grid and search window
@pytest.mark.parametrize("n,s", [(512, 3), (1920, 3), (1080, 3), (37, 4), (8, 3), (9, 8), (100, 7), (64, 8)])
def test_grid(n, s):
    g = oracle.grid(n, 8, s)
    assert list(g) == npref.grid(n, s)
    last = n - 8
    assert g[0] == 0 and g[-1] == last and np.all(np.diff(g) > 0) and np.all(np.diff(g) <= s)
    np.testing.assert_array_equal(g, np.minimum(np.arange(len(g)) * s, last))


def test_grid_sizes_of_workloads():
    assert len(oracle.grid(512, 8, 3)) == 169
    assert len(oracle.grid(1920, 8, 3)) * len(oracle.grid(1080, 8, 3)) == 229401


def test_search_boundary_brute_force():
    for step in (1, 2, 3):
        for R in (4, 9, 12):
            Ra = R // step * step
            for bound in (0, 30):
                for pos in range(0, 31):
                    lo = oracle.search_boundary(pos, 0, R, step)
                    hi = oracle.search_boundary(pos, bound, R, step)
                    reach = [q for q in range(pos - Ra, pos + Ra + 1, step) if 0 <= q]
                    assert lo == min(reach) if pos > 0 else lo == 0
                    if pos < bound:
                        assert hi == max(q for q in range(pos - Ra, pos + Ra + 1, step) if q <= bound)
                    elif pos == bound:
                        assert hi == bound


# --- block matching vs brute force
CASES = [((24, 20), dict(step=3, R=4)), ((37, 41), dict(step=3, R=6)), ((40, 32), dict(step=4, R=9, G1=8)),
         ((33, 29), dict(step=2, R=6, bmstep=2)), ((30, 26), dict(step=3, R=5, G1=1))]


@pytest.mark.parametrize("shape,kw", CASES)
def test_match_vs_brute_force(shape, kw):
    h, w = shape
    _, n = synth.grey_pair(h, w, 25.0, h * 100 + w)
    p = para(**kw)
    counts, pos, keys = oracle.match_all(p.basic, n)
    bf = npref.brute_groups(p.basic, n)
    gy, gx = npref.grid(h, p.basic.BlockStep), npref.grid(w, p.basic.BlockStep)
    assert counts.shape == (len(gy), len(gx))
    for a, j in enumerate(gy):
        for b, i in enumerate(gx):
            g = bf[(j, i)]
            N = counts[a, b]
            assert N == len(g)
            got = [(int(v) >> 16, int(v) & 0xFFFF) for v in pos[a, b, :N]]
            assert got == [(y, x) for _, y, x in g]
            np.testing.assert_allclose(keys[a, b, :N], [k for k, _, _ in g], rtol=1e-12, atol=1e-12)


# BlockSize other than 8 (profile vn's final step: 11) and GroupSize above 32
BLOCK_CASES = [((40, 37), 11, dict(step=6, R=8, G1=32, G2=32)),
               ((31, 33), 5, dict(step=2, R=6, G1=48, G2=48)),
               ((38, 36), 16, dict(step=5, R=5, G1=12, G2=12)),
               ((30, 29), 8, dict(step=1, R=4, G1=64, G2=64, sigma=50.0))]


def block_para(B, kw):
    p = para(**kw)
    p.basic.BlockSize = p.final.BlockSize = B
    return p


def smooth_pair(h, w, seed):
    """A smooth image with weak noise (10/255): most window candidates pass
    the default threshold, so groups fill up to GroupSize."""
    y, x = np.mgrid[0:h, 0:w]
    clean = 0.4 + 0.1 * np.sin(x / 7.0) * np.cos(y / 9.0) + 0.002 * x
    n = clean + np.random.default_rng(seed).standard_normal((h, w)) * (10 / 255)
    return clean, n.astype(np.float32)


@pytest.mark.parametrize("shape,B,kw", BLOCK_CASES)
def test_match_vs_brute_force_any_block_size(shape, B, kw):
    h, w = shape
    _, n = smooth_pair(h, w, h * 7 + w)
    p = block_para(B, kw)
    counts, pos, keys = oracle.match_all(p.basic, n)
    bf = npref.brute_groups(p.basic, n)
    gy, gx = npref.grid(h, p.basic.BlockStep, B), npref.grid(w, p.basic.BlockStep, B)
    assert list(gy) == list(oracle.grid(h, B, p.basic.BlockStep))
    assert counts.shape == (len(gy), len(gx))
    assert counts.max() == p.basic.GroupSize  # full groups are exercised
    for a, j in enumerate(gy):
        for b, i in enumerate(gx):
            g = bf[(j, i)]
            N = counts[a, b]
            assert N == len(g)
            got = [(int(v) >> 16, int(v) & 0xFFFF) for v in pos[a, b, :N]]
            assert got == [(y, x) for _, y, x in g]
            np.testing.assert_allclose(keys[a, b, :N], [k for k, _, _ in g], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("wiener", [False, True])
@pytest.mark.parametrize("shape,B,kw", BLOCK_CASES)
def test_kernel_vs_orthonormal_numpy_any_block_size(shape, B, kw, wiener):
    h, w = shape
    clean, n = smooth_pair(h, w, h + 3 * w)
    p = block_para(B, kw)
    st = p.final if wiener else p.basic
    ref = (clean + 0.02 * np.sin(np.arange(w) / 3.0)).astype(np.float32) if wiener else n
    got = oracle.kernel(st, n, ref, wiener=wiener)
    want = npref.kernel(st, [n], [ref], wiener)[0]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)


def test_match_invariants():
    _, n = synth.grey_pair(64, 72, 25.0, 3)
    p = para()
    counts, pos, keys = oracle.match_all(p.final, n)
    gy, gx = oracle.grid(64, 8, 3), oracle.grid(72, 8, 3)
    thkey = p.final.thMSE * np.sqrt(1 / 3)
    for a, j in enumerate(gy):
        for b, i in enumerate(gx):
            N = counts[a, b]
            assert 1 <= N <= p.final.GroupSize
            assert pos[a, b, 0] == (j << 16) | i and keys[a, b, 0] == 0
            k = keys[a, b, 1:N]
            assert np.all(k > 0) and np.all(k <= thkey * (1 + 1e-12)) and np.all(np.diff(k) >= 0)
            ys, xs = pos[a, b, :N] >> 16, pos[a, b, :N] & 0xFFFF
            assert np.all(np.abs(ys - j) <= 12) and np.all(np.abs(xs - i) <= 12)
            assert np.all(ys <= 64 - 8) and np.all(xs <= 72 - 8)
            assert len(set(zip(ys.tolist(), xs.tolist()))) == N


# --- collaborative filter and aggregation vs the orthonormal numpy formulation
@pytest.mark.parametrize("wiener", [False, True])
@pytest.mark.parametrize("shape,kw", [((24, 32), dict(step=3, R=4, G1=8, G2=8)),
                                      ((29, 27), dict(step=2, R=5, G1=16, G2=16, sigma=40.0))])
def test_kernel_vs_orthonormal_numpy(shape, kw, wiener):
    h, w = shape
    clean, n = synth.grey_pair(h, w, kw.get("sigma", 25.0), 5)
    p = para(**kw)
    st = p.final if wiener else p.basic
    ref = clean.astype(np.float32) if wiener else n  # Wiener guide: a clean-ish estimate
    got = oracle.kernel(st, n, ref, wiener=wiener)
    want = npref.kernel(st, [n], [ref], wiener)[0]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)


def test_kernel_three_planes_vs_numpy():
    c, n = synth.rgb_pair(24, 28, 30.0, 9)
    p = para(step=3, R=4, G1=8, G2=8)
    yuv = oracle.rgb2opp(n)
    got = oracle.kernel(p.basic, list(yuv), list(yuv), wiener=False)
    want = npref.kernel(p.basic, list(yuv), list(yuv), False)
    for c_ in range(3):
        np.testing.assert_allclose(got[c_], want[c_], atol=1e-10)


def test_constant_image_closed_form():
    # A constant image matches nothing (every SSE is 0, include/Block.h:608),
    # so every group is the lone reference block (N = 1).  Step 1 keeps the
    # DC (8c > lambda*sigma) and reproduces c exactly; step 2 scales it by the
    # Wiener factor of the DC, W = 64c^2 / (64c^2 + sigma^2 / 2^3).
    p = para()
    c = 0.4
    src = np.full((40, 48), c)
    counts, _, _ = oracle.match_all(p.final, src)
    assert counts.max() == 1
    np.testing.assert_allclose(oracle.kernel(p.basic, src, src, wiener=False), c, atol=1e-15)
    s = 25 / 255 * np.sqrt(1 / 3)
    W = 64 * c * c / (64 * c * c + s * s / 8)
    np.testing.assert_allclose(oracle.kernel(p.final, src, src, wiener=True), c * W, rtol=1e-13)
    # Two steps: the basic estimate carries ~1e-17 rounding, so step-2 SSEs
    # are no longer exactly 0 and groups can fill; the Wiener DC factor then
    # lies between the N = 1 value and 1.
    two = oracle.bm3d_plane(p.basic, p.final, src)
    assert np.ptp(two) < 1e-13 and c * W - 1e-15 <= two.min() and two.max() <= c
    rgb = np.stack([np.full((40, 48), v) for v in (0.2, 0.5, 0.7)])
    out = oracle.bm3d_rgb(p.basic, p.final, rgb)
    assert np.ptp(out, axis=(1, 2)).max() < 1e-13  # stays spatially constant
    np.testing.assert_allclose(out[:, 0, 0], [0.2, 0.5, 0.7], atol=1e-4)


def test_tiny_sigma_is_identity():
    _, n = synth.grey_pair(40, 40, 25.0, 2)
    p = para(sigma=1e-9)
    np.testing.assert_allclose(oracle.kernel(p.basic, n, n, wiener=False), n, atol=1e-12)
    np.testing.assert_allclose(oracle.kernel(p.final, n, n, wiener=True), n, atol=1e-9)


def test_degenerate_params_copy_through():
    _, n = synth.grey_pair(32, 32, 25.0, 2)
    p = para(sigma=0.0)
    np.testing.assert_array_equal(oracle.kernel(p.basic, n, n, wiener=False), n.astype(np.float64))
    p = para()
    p.basic.GroupSize = 0
    np.testing.assert_array_equal(oracle.kernel(p.basic, n, n, wiener=False), n.astype(np.float64))


def test_denoises():
    c, n = synth.grey_pair(96, 96, 25.0, 4)
    p = para(sigma=25.0 / np.sqrt(1 / 3))  # effective sigma 25 after the normY scaling
    out = oracle.bm3d_plane(p.basic, p.final, n)
    assert synth.psnr(out, c) > synth.psnr(n, c) + 5


# --- OPP colour
def test_opp_matrices_and_inverse():
    rgb = np.random.default_rng(1).uniform(0, 1, (3, 5, 7))
    yuv = oracle.rgb2opp(rgb)
    r, g, b = rgb
    np.testing.assert_allclose(yuv[0], (r + g + b) / 3, atol=1e-15)
    np.testing.assert_allclose(yuv[1], (r - b) / 2, atol=1e-15)
    np.testing.assert_allclose(yuv[2], (r - 2 * g + b) / 4, atol=1e-15)
    np.testing.assert_allclose(oracle.opp2rgb(yuv, clip=False), rgb, atol=1e-14)
    out = oracle.opp2rgb(np.stack([np.full((2, 2), 1.2), np.zeros((2, 2)), np.zeros((2, 2))]), clip=True)
    assert out.max() == 1.0


def test_threads_agree():
    c, n = synth.rgb_pair(60, 70, 25.0, 8)
    p = para()
    oracle.set_threads(1)
    a = oracle.bm3d_rgb(p.basic, p.final, n)
    oracle.set_threads(4)
    b = oracle.bm3d_rgb(p.basic, p.final, n)
    oracle.set_threads(1)
    np.testing.assert_allclose(a, b, atol=1e-12)


# --- integer quantisation (RangeConvert, include/Conversion.hpp:134-256)
QUANT = [(0, 0, 255), (16, 16, 235), (0, 0, 1023), (0, 128, 255), (16, 128, 240), (0, 32768, 65535)]


@pytest.mark.parametrize("q", QUANT, ids=[str(q) for q in QUANT])
def test_range_convert_closed_forms(q):
    floor, neutral, ceil = q
    lv = np.arange(floor, ceil + 1, dtype=np.int32)
    f = oracle.range_to_float(lv, *q)
    chroma = floor < neutral
    pc_chroma = chroma and (floor + ceil) % 2 == 1
    zero = neutral if chroma else floor
    # the zero level maps to 0, one level is one 1/(Ceil-Floor) step
    assert f[zero - floor] == 0.0
    np.testing.assert_allclose(np.diff(f[1:-1] if pc_chroma else f), 1.0 / (ceil - floor), rtol=0, atol=1e-15)
    if pc_chroma:
        # an odd PC range has no centre level: Floor lies below -0.5 and is clipped
        assert f[0] == -0.5 and f[-1] < 0.5
    else:
        lo = -0.5 if chroma else 0.0
        np.testing.assert_allclose([f[0], f[-1]], [lo, lo + 1.0], atol=1e-15)
    # every level survives the round trip, the clipped PC-chroma Floor included
    # (the 0.499999 offset: with 0.5 it would come back as Floor + 1)
    np.testing.assert_array_equal(oracle.range_to_int(f, *q), lv)


def test_range_to_int_rounding_and_clip():
    q = (16, 16, 235)
    step = 1.0 / 219
    k = np.arange(10)
    # halfway values round up, just below halfway rounds down
    np.testing.assert_array_equal(oracle.range_to_int((k + 0.5) * step, *q), 16 + k + 1)
    np.testing.assert_array_equal(oracle.range_to_int((k + 0.49) * step, *q), 16 + k)
    np.testing.assert_array_equal(oracle.range_to_int(np.array([-1.0, -1e-9, 1.0 + 1e-9, 7.0]), *q),
                                  [16, 16, 235, 235])
    q = (0, 128, 255)
    np.testing.assert_array_equal(oracle.range_to_int(np.array([-0.9, 0.0, 0.9]), *q), [0, 128, 255])


def _quant(x, floor, ceil):
    return np.clip(np.floor(x * (ceil - floor) + floor + 0.5), floor, ceil).astype(np.int32)


@pytest.mark.parametrize("mode", [1, 3])
def test_plane_int_agrees_with_float_path(mode):
    # same data as a float plane: at most one level apart (the conversions round
    # v*(1/range) where the float input was v/range), on very few pixels
    clean, n = synth.grey_pair(48, 56, 25.0, 21)
    src = _quant(np.clip(n, 0, 1), 0, 255)
    p = para()
    got = oracle.bm3d_plane_int(p.basic, p.final, src, mode=mode)
    x = src.astype(np.float64) / 255
    ref = oracle.kernel(p.basic, x, x, wiener=False) if mode == 1 else oracle.bm3d_plane(p.basic, p.final, x)
    want = _quant(ref, 0, 255)
    d = np.abs(got.astype(np.int64) - want)
    assert d.max() <= 1 and (d > 0).mean() < 0.01
    assert synth.psnr(got / 255.0, clean) > synth.psnr(src / 255.0, clean) + 1


def test_plane_int_chroma_shift():
    # a chroma plane is the luma plane moved by Neutral: the filter sees the
    # same data minus a constant, so outputs agree up to that shift except where
    # a group's DC coefficient falls under the hard threshold
    _, n = synth.grey_pair(40, 48, 10.0, 5)
    src = _quant(0.5 + 0.2 * (np.clip(n, 0, 1) - 0.5), 0, 255)
    p = para(sigma=10.0)
    lum = oracle.bm3d_plane_int(p.basic, p.final, src, mode=3, floor=16, neutral=16, ceil=240)
    chr_ = oracle.bm3d_plane_int(p.basic, p.final, src, mode=3, floor=16, neutral=128, ceil=240)
    d = np.abs(lum.astype(np.int64) - chr_)
    assert np.median(d) == 0 and (d > 1).mean() < 0.05


def test_rgb_int_agrees_with_float_path():
    clean, n = synth.rgb_pair(40, 48, 25.0, 22)
    src = _quant(np.clip(n, 0, 1), 0, 255)
    p = para()
    got = oracle.bm3d_rgb_int(p.basic, p.final, src, mode=3)
    want = _quant(oracle.bm3d_rgb(p.basic, p.final, src.astype(np.float64) / 255, mode=3), 0, 255)
    d = np.abs(got.astype(np.int64) - want)
    assert d.max() <= 1 and (d > 0).mean() < 0.01


def test_int_degenerate_copy_through():
    src = _quant(synth.grey_pair(32, 32, 25.0, 2)[1].clip(0, 1), 16, 235)
    p = para(sigma=0.0)
    for mode in (1, 2):
        np.testing.assert_array_equal(oracle.bm3d_plane_int(p.basic, p.final, src, src, mode=mode,
                                                            floor=16, neutral=16, ceil=235), src)
    rgb = np.stack([src, src[::-1], src[:, ::-1]])
    np.testing.assert_array_equal(oracle.bm3d_rgb_int(p.basic, p.final, rgb, mode=1), rgb)


# --- the learned pilot's FP64 oracle (oracle/pilot_oracle.py)
def _pilot_params(rng, scale=1.0, channels=1):
    from oracle import pilot_oracle as po
    out = []
    for cin, cout, relu in po.layers(channels):
        out.append((rng.standard_normal((cout, cin, 9)) * np.sqrt(2.0 / (cin * 9)) * scale,
                    rng.standard_normal(cout) * 0.05))
    return out


def test_pilot_conv_matches_scipy_correlate2d():
    from scipy.signal import correlate2d
    from oracle import pilot_oracle as po
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 7, 9, 3))
    w = rng.standard_normal((4, 3, 9))
    b = rng.standard_normal(4)
    got = po.conv3x3(x, w, b)
    for n in range(2):
        for co in range(4):
            want = b[co] + sum(correlate2d(x[n, :, :, ci], w[co, ci].reshape(3, 3), mode="same", boundary="fill")
                               for ci in range(3))
            np.testing.assert_allclose(got[n, :, :, co], want, atol=1e-12)


def test_pilot_conv_adjoint_identity():
    # <conv(x), y> == <x, dinput(y)>: pins the flipped, transposed kernel of the input gradient
    from oracle import pilot_oracle as po
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 6, 5, 4))
    y = rng.standard_normal((1, 6, 5, 3))
    w = rng.standard_normal((3, 4, 9))
    lhs = np.sum(po.conv3x3(x, w) * y)
    rhs = np.sum(x * po.conv3x3_dinput(y, w))
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(lhs))


@pytest.mark.parametrize("channels", [1, 3])
def test_pilot_gradients_match_finite_differences(channels):
    from oracle import pilot_oracle as po
    rng = np.random.default_rng(2 + channels)
    params = _pilot_params(rng, channels=channels)
    shape = (1, 5, 6) if channels == 1 else (1, channels, 5, 6)
    noisy = rng.uniform(0, 1, shape)
    sigma = np.array([25.0])
    target = rng.uniform(0, 1, shape)
    off = None if channels == 1 else np.array([[0.0, 0.5, 0.5]])
    L0, grads = po.loss_and_grads(params, noisy, sigma, target, offset=off)
    assert abs(L0 - po.loss(params, noisy, sigma, target, offset=off)) < 1e-15
    h = 1e-6
    for l in range(len(params)):
        w, b = params[l]
        for idx in [tuple(rng.integers(0, s) for s in w.shape) for _ in range(3)]:
            wp = [(ww.copy(), bb.copy()) for ww, bb in params]
            wm = [(ww.copy(), bb.copy()) for ww, bb in params]
            wp[l][0][idx] += h
            wm[l][0][idx] -= h
            fd = (po.loss(wp, noisy, sigma, target, offset=off) - po.loss(wm, noisy, sigma, target, offset=off)) / (2 * h)
            assert abs(fd - grads[l][0][idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (l, idx, fd, grads[l][0][idx])
        j = int(rng.integers(0, b.shape[0]))
        bp = [(ww.copy(), bb.copy()) for ww, bb in params]
        bm = [(ww.copy(), bb.copy()) for ww, bb in params]
        bp[l][1][j] += h
        bm[l][1][j] -= h
        fd = (po.loss(bp, noisy, sigma, target, offset=off) - po.loss(bm, noisy, sigma, target, offset=off)) / (2 * h)
        assert abs(fd - grads[l][1][j]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (l, j, fd, grads[l][1][j])


def test_pilot_multichannel_equals_separate_single_channel_parts():
    # a 3-plane network whose weights couple no planes (block-diagonal first
    # and last layers over three disjoint groups of 16 hidden channels, zero
    # cross terms) computes three independent single-plane networks
    from oracle import pilot_oracle as po
    rng = np.random.default_rng(9)
    C, H, W = 3, 6, 7
    noisy = rng.uniform(0, 1, (2, C, H, W))
    sigma = np.array([20.0, 35.0])
    grey = _pilot_params(rng, channels=1)
    rgb = []
    for l, (cin, cout, relu) in enumerate(po.layers(C)):
        w = np.zeros((cout, cin, 9))
        b = np.zeros(cout)
        for c in range(C):
            o = slice(16 * c, 16 * c + 16) if cout == 48 else slice(c, c + 1)
            i = slice(16 * c, 16 * c + 16) if cin == 48 else [c, C]
            gw, gb = grey[l]
            go = slice(0, 16) if cout == 48 else slice(0, 1)
            gi = slice(0, 16) if cin == 48 else [0, 1]
            w[o, i] = gw[go][:, gi]
            b[o] = gb[go]
        rgb.append((w, b))
    out3, _, _ = po.forward(rgb, noisy, sigma)
    for c in range(C):
        g16 = [(w[:16] if w.shape[0] == 48 else w[:1], b[:16] if b.shape[0] == 48 else b[:1]) for w, b in grey]
        g16 = [(w[:, :16] if w.shape[1] == 48 else w, b) for w, b in g16]
        out1, _, _ = po.forward(g16, noisy[:, c], sigma)
        np.testing.assert_allclose(out3[:, c], out1, atol=1e-12)


def test_bf16_round_known_values():
    from oracle import pilot_oracle as po
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1e-3, 0.0])
    got = po.bf16_round(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 1.5 * 2^-8 rounds up
    np.testing.assert_array_equal(got[:4], [1.0, 1.0, 1.0 + 2 ** -7, -2.5])
    assert abs(got[4] - 1e-3) <= 1e-3 * 2 ** -8 and got[5] == 0.0


# --- the configs' variant modes (Haar along the group axis, Kaiser aggregation)
@pytest.fixture
def variant():
    def set_(haar=False, beta=0.0):
        oracle.set_variant(haar, beta)
    yield set_
    oracle.set_variant(False, 0.0)


VARIANTS = [(True, 0.0), (False, 2.0), (True, 2.0)]


@pytest.mark.parametrize("wiener", [False, True])
@pytest.mark.parametrize("haar,beta", VARIANTS)
def test_variant_kernel_vs_numpy(variant, haar, beta, wiener):
    # independent: numpy's Kaiser window, the Haar basis functions written out,
    # scipy's orthonormal DCT; groups of up to 16 (truncated to 8, 4, 2, 1)
    clean, n = smooth_pair(34, 38, 3) if haar else synth.grey_pair(30, 33, 25.0, 4)
    p = para(step=3, R=6, G1=16, G2=16)
    st = p.final if wiener else p.basic
    ref = (clean + 0.01).astype(np.float32) if wiener else n
    variant(haar, beta)
    got = oracle.kernel(st, n, ref, wiener=wiener)
    want = npref.kernel(st, [n], [ref], wiener, haar=haar, kaiser_beta=beta)[0]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)
    counts, _, _ = oracle.match_all(st, ref)
    assert counts.max() > 8  # truncation from above 8 members is exercised


def test_variant_haar_of_two_equals_the_dct(variant):
    # the orthonormal Haar matrix of length 2 is the length-2 DCT-II: with
    # groups of at most 2 the Haar mode is the reference's filter
    clean, n = synth.grey_pair(40, 36, 25.0, 6)
    p = para(step=2, R=5, G1=2, G2=2)
    want = oracle.bm3d_plane(p.basic, p.final, n)
    variant(True, 0.0)
    np.testing.assert_allclose(oracle.bm3d_plane(p.basic, p.final, n), want, rtol=0, atol=1e-12)


def test_variant_constant_image_and_window(variant):
    # a constant image: every group is the block alone, every estimate the
    # constant, so any aggregation window returns it
    src = np.full((36, 40), 0.3)
    p = para()
    variant(True, 2.0)
    np.testing.assert_allclose(oracle.kernel(p.basic, src, src, wiener=False), 0.3, atol=1e-14)
    # numpy's Kaiser window (the variant's weights) is symmetric, peaks at the
    # centre and equals 1/I0(beta) at the ends
    k = np.kaiser(8, 2.0)
    np.testing.assert_allclose(k, k[::-1])
    assert abs(k[0] - 1.0 / np.i0(2.0)) < 1e-15


def test_variant_off_is_the_reference(variant):
    clean, n = synth.grey_pair(32, 32, 25.0, 7)
    p = para(step=3, R=5, G1=8, G2=8)
    want = oracle.bm3d_plane(p.basic, p.final, n)
    variant(True, 2.0)
    other = oracle.bm3d_plane(p.basic, p.final, n)
    variant(False, 0.0)
    np.testing.assert_array_equal(oracle.bm3d_plane(p.basic, p.final, n), want)
    assert np.abs(other - want).max() > 1e-4   # the variants do change the output
