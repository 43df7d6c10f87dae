"""Pins of the NL-means oracle (oracle/bm3d_oracle.c orc_nlm_*, restating
source/NLMeans.cpp) against an independent numpy formulation on brute-force
groups (tests/npref.py nlm) and against closed forms.  CPU only."""
import types

import numpy as np
import pytest

import oracle
from tests import npref


def nparams(**kw):
    p = oracle.nlm_params(**kw)
    return p, types.SimpleNamespace(correction=p.correction, sigma=p.sigma, strength=p.strength,
                                    BlockStep=p.block_step, GroupSize=p.group_size, BMrange=p.bm_range,
                                    BMstep=p.bm_step, thMSE=p.thMSE)


def smooth_noisy(h, w, sigma, seed):
    """A smooth gradient (blocks close to each other) plus Gaussian noise, [0, 1]."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    clean = 0.3 + 0.2 * np.sin(xx / 9.0) * np.cos(yy / 11.0) + 0.1 * xx / w
    return clean, clean + rng.normal(0, sigma / 255.0, (h, w))


def multi_member_fraction(np_, ref):
    g = npref.brute_groups(np_, ref, norm=1.0)
    return np.mean([len(v) >= 2 for v in g.values()])


@pytest.mark.parametrize("corr,bstep,mstep,G,R", [(True, 5, 3, 16, 24), (False, 5, 3, 16, 24), (True, 3, 1, 8, 6),
                                                 (False, 4, 2, 4, 9), (True, 8, 2, 32, 12), (True, 1, 3, 16, 9)])
def test_plane_against_numpy_restatement(corr, bstep, mstep, G, R):
    # thMSE above the noise's 2 sigma^2 = 800 (8-bit MSE units) so most groups have members
    p, q = nparams(correction=corr, sigma=20.0, BlockStep=bstep, BMstep=mstep, GroupSize=G, BMrange=R, thMSE=1500.0)
    _, n = smooth_noisy(37, 45, 20.0, seed=bstep * 7 + mstep)
    assert multi_member_fraction(q, n) > 0.5       # the averaging is exercised, not only copies
    got = oracle.nlm_plane(p, n)
    nums, den = npref.nlm(q, [n], n)
    np.testing.assert_allclose(got, nums[0] / den, rtol=0, atol=1e-12)


def test_plane_separate_guide():
    p, q = nparams(sigma=15.0)
    clean, n = smooth_noisy(40, 48, 15.0, seed=3)
    got = oracle.nlm_plane(p, n, ref=clean)
    nums, den = npref.nlm(q, [n], clean)
    np.testing.assert_allclose(got, nums[0] / den, rtol=0, atol=1e-12)
    assert np.abs(got - oracle.nlm_plane(p, n)).max() > 1e-3


@pytest.mark.parametrize("corr", [True, False])
def test_int_frame_against_numpy_restatement(corr):
    p, q = nparams(correction=corr, sigma=12.0)
    rng = np.random.default_rng(5)
    _, base = smooth_noisy(33, 41, 0.0, seed=1)
    frame = np.stack([base, base * 0.8 + 0.1, 1 - base]) * 1000 + rng.normal(0, 12 * 1000 / 255, (3, 33, 41))
    fi = np.clip(np.rint(frame), 0, 1000).astype(np.int32)
    got = oracle.nlm_rgb_int(p, fi, floor=0, ceil=1000)
    # ConvertToY OPP with clip (include/Conversion.hpp:418-432) into [0, 1]
    y = np.clip(fi.sum(0) * (1.0 / (1000 * 3.0)), 0, 1)
    assert multi_member_fraction(q, y) > 0.5
    nums, den = npref.nlm(q, list(fi), y, value_range=1000.0)
    want = np.floor(np.stack(nums) / den + 0.5)
    d = np.abs(got - want)
    assert d.max() <= 1 and np.mean(d > 0) < 1e-3


def test_int_plane_against_numpy_restatement():
    p, q = nparams(sigma=20.0)
    _, n = smooth_noisy(36, 40, 20.0, seed=9)
    ni = np.rint(n * 255).astype(np.int32)
    ni[0, :4] = [-7, 300, 0, 255]                   # samples outside [0, 255] exercise the input clip
    got = oracle.nlm_plane_int(p, ni, floor=0, neutral=0, ceil=255)
    f = np.clip(ni / 255.0, 0, 1)
    nums, den = npref.nlm(q, [f], f)
    want = np.floor(nums[0] / den * 255 + 0.5)
    d = np.abs(got - want)
    assert d.max() <= 1 and np.mean(d > 0) < 1e-3


def test_constant_image_is_a_fixed_point():
    # every candidate has SSE 0, which matching excludes: groups of one, copies
    # (num / den of k equal copies, exact up to the rounding of the sum)
    p, _ = nparams(sigma=10.0)
    c = np.full((30, 34), 0.37)
    np.testing.assert_allclose(oracle.nlm_plane(p, c), c, rtol=0, atol=1e-15)


def test_offset_invariance():
    # keys are distances and weights depend on them only: a constant offset
    # of the image shifts the output by the same constant
    for corr in (True, False):
        p, _ = nparams(correction=corr, sigma=20.0)
        _, n = smooth_noisy(32, 40, 20.0, seed=11)
        a = oracle.nlm_plane(p, n)
        b = oracle.nlm_plane(p, n + 0.25)
        np.testing.assert_allclose(b - 0.25, a, rtol=0, atol=1e-12)


def test_correction_limits():
    _, n = smooth_noisy(32, 40, 20.0, seed=13)
    # VarN = 0: every pixel with VarX > 0 keeps its own sample X
    p0, _ = nparams(correction=True, sigma=0.0, strength=100.0, thMSE=1000.0)
    np.testing.assert_allclose(oracle.nlm_plane(p0, n), n, rtol=0, atol=1e-14)
    # VarN above every VarX: the corrected form is the weighted mean itself
    pb, _ = nparams(correction=True, sigma=1e6, strength=100.0, thMSE=1000.0)
    pm, _ = nparams(correction=False, sigma=1e6, strength=100.0, thMSE=1000.0)
    np.testing.assert_allclose(oracle.nlm_plane(pb, n), oracle.nlm_plane(pm, n), rtol=0, atol=1e-15)


def test_infinite_strength_is_the_group_mean():
    # strength -> infinity: every weight 1, the estimate is the plain mean of
    # the group's blocks
    p, q = nparams(correction=False, sigma=20.0, strength=1e150)
    _, n = smooth_noisy(29, 37, 20.0, seed=17)
    got = oracle.nlm_plane(p, n)
    num = np.zeros_like(n)
    den = np.zeros_like(n)
    for (j, i), g in npref.brute_groups(q, n, norm=1.0).items():
        num[j:j + 8, i:i + 8] += np.mean([n[y:y + 8, x:x + 8] for _, y, x in g], axis=0)
        den[j:j + 8, i:i + 8] += 1
    np.testing.assert_allclose(got, num / den, rtol=0, atol=1e-12)


@pytest.mark.parametrize("kw", [dict(GroupSize=1), dict(strength=0.0), dict(strength=-1.0), dict(thMSE=0.0),
                                dict(BMrange=2, BMstep=3), dict(BMrange=0)])
def test_degenerate_parameters_copy_the_source(kw):
    # NLMeans::process_* early return (source/NLMeans.cpp:152-158, 234-240, 263-269)
    p, _ = nparams(sigma=10.0, **kw)
    _, n = smooth_noisy(20, 24, 10.0, seed=2)
    np.testing.assert_array_equal(oracle.nlm_plane(p, n), n)
    fi = np.rint(np.stack([n, n, n]) * 255).astype(np.int32)
    np.testing.assert_array_equal(oracle.nlm_rgb_int(p, fi), fi)


def test_uncovered_pixels_with_a_block_step_above_the_block_size():
    # BlockStep 10 > BlockSize 8 leaves columns / rows 8, 9, 18, 19, ... without
    # a block: 0 / 0 in the reference's num / den
    p, _ = nparams(sigma=20.0, BlockStep=10)
    _, n = smooth_noisy(30, 30, 20.0, seed=4)
    got = oracle.nlm_plane(p, n)
    cov = np.zeros(30, bool)
    for j in npref.grid(30, 10):
        cov[j:j + 8] = True
    np.testing.assert_array_equal(np.isnan(got), ~(cov[:, None] & cov[None, :]))
