"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Bars (plain FP32 path): max |GPU - oracle| <= 2e-3 on [0,1] data,
|dPSNR| <= 0.01 dB, >= 99.9% identical ordered groups.  Matching on a shared
guide is expected to be exact: distances are summed in FP64 on the GPU.
"""
import os

import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Basic, BM3D_Final, synth
from isp_mw_amd import _lib as L
from tests.gpu_common import MIN_GROUPS, assert_close, group_identity, para

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)


# ---------------------------------------------------------------- matching
MATCH_CASES = [
    ("np s3 R12", (96, 80), dict(), False),
    ("np s3 R12 final", (96, 80), dict(), True),
    ("fast s8 R9", (120, 100), dict(profile="fast", step=None, R=None), False),
    ("np s4 R16", (90, 110), dict(step=4, R=16), True),
    ("BMstep 2", (70, 66), dict(step=3, R=9, bmstep=2), False),
    ("BMstep 3 R10", (64, 61), dict(step=2, R=10, bmstep=3), True),
    ("G=1 bypass", (40, 40), dict(G1=1), False),
    ("G=2", (48, 40), dict(G1=2), False),
    ("tiny 8x8", (8, 8), dict(), False),
    ("tiny 9x13", (9, 13), dict(), True),
    ("sigma 50", (64, 64), dict(sigma=50.0), False),
]


@pytest.mark.parametrize("name,shape,kw,final", MATCH_CASES, ids=[c[0] for c in MATCH_CASES])
def test_match_groups_identical(gpu, name, shape, kw, final):
    h, w = shape
    say(f"match case {name} {w}x{h} final={final}")
    _, n = synth.grey_pair(h, w, kw.get("sigma", 25.0), h * 7 + w)
    p = para(**kw)
    st = p.final if final else p.basic
    f = BM3D(p)
    c, pos, keys = f.match(n, final_step=final)
    oc, op, ok = oracle.match_all(st, n)
    frac = group_identity(c, pos, oc, op)
    say(f"  identical groups {frac:.5f}, mean N {c.mean():.2f}")
    assert frac == 1.0
    G = st.GroupSize
    m = np.arange(G)[None, None, :] < oc[..., None]
    np.testing.assert_allclose(np.where(m, keys, 0), np.where(m, ok, 0), rtol=1e-12, atol=1e-12)


def test_match_full_window_selection_large_counts(gpu):
    # flat noisy image: hundreds of candidates pass, the radix select decides
    rng = np.random.default_rng(3)
    n = (0.5 + rng.standard_normal((64, 96)) * 0.03).astype(np.float32)
    p = para(G1=16, G2=32)
    f = BM3D(p)
    for final in (False, True):
        c, pos, _ = f.match(n, final_step=final)
        oc, op, _ = oracle.match_all(p.final if final else p.basic, n)
        assert group_identity(c, pos, oc, op) == 1.0
        assert c.min() == (p.final if final else p.basic).GroupSize


def test_match_ties_follow_scan_order(gpu):
    # a periodic image has exactly equal distances; equal keys keep scan order
    y, x = np.mgrid[0:48, 0:48]
    n = (0.5 + 0.2 * np.sin(2 * np.pi * x / 4) + 0.1 * np.cos(2 * np.pi * y / 6)).astype(np.float32)
    p = para(G1=8)
    c, pos, _ = BM3D(p).match(n)
    oc, op, _ = oracle.match_all(p.basic, n)
    assert group_identity(c, pos, oc, op) == 1.0


# ---------------------------------------------------------------- filtering
@pytest.mark.parametrize("mode", ["basic", "final", "both"])
@pytest.mark.parametrize("shape", [(64, 64), (80, 96), (37, 53)])
def test_plane_vs_oracle(gpu, mode, shape):
    h, w = shape
    say(f"plane {mode} {w}x{h}")
    clean, n = synth.grey_pair(h, w, 25.0, h + w)
    p = para()
    if mode == "basic":
        got = BM3D_Basic(p.basic)(n)
        ref = oracle.kernel(p.basic, n, n, wiener=False)
    elif mode == "final":
        guide = oracle.kernel(p.basic, n, n, wiener=False).astype(np.float32)
        got = BM3D_Final(p.final).process(np.empty_like(n), n, guide)
        ref = oracle.kernel(p.final, n, guide, wiener=True)
    else:
        got = BM3D(p)(n)
        ref = oracle.bm3d_plane(p.basic, p.final, n)
    assert_close(got, ref, clean, label=f"plane {mode} {w}x{h}")


@pytest.mark.parametrize("profile", ["fast", "lc", "np", "high", "vn"])
def test_profiles_vs_oracle(gpu, profile):
    say(f"profile {profile}")
    clean, n = synth.grey_pair(72, 88, 20.0, 31)
    p = para(profile, sigma=20.0, step=None, R=None)
    assert_close(BM3D(p)(n), oracle.bm3d_plane(p.basic, p.final, n), clean, label=f"profile {profile}")


def test_separate_ref_plane(gpu):
    clean, n = synth.grey_pair(64, 72, 25.0, 12)
    ref = clean.astype(np.float32)
    p = para()
    got = BM3D(p).process(np.empty_like(n), n, ref)
    want = oracle.bm3d_plane(p.basic, p.final, n, ref)
    assert_close(got, want, clean, label="plane both ref")
    got = BM3D_Basic(p.basic).process(np.empty_like(n), n, ref)
    assert_close(got, oracle.kernel(p.basic, n, ref, wiener=False), label="plane basic ref")


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_rgb_vs_oracle(gpu, mode):
    say(f"rgb mode {mode}")
    clean, n = synth.rgb_pair(56, 72, 25.0, 40 + mode)
    p = para()
    cls = {1: BM3D_Basic, 2: BM3D_Final, 3: BM3D}[mode]
    f = cls(p.basic if mode == 1 else p.final if mode == 2 else p)
    if mode == 2:
        guide = clean.astype(np.float32)
        got = f.process(np.empty_like(n), n, guide)
        ref = oracle.bm3d_rgb(p.basic, p.final, n, guide, mode=2)
    else:
        got = f(n)
        ref = oracle.bm3d_rgb(p.basic, p.final, n, mode=mode)
    assert_close(got, ref, clean, label=f"rgb mode {mode}")


def test_rgb_separate_ref(gpu):
    clean, n = synth.rgb_pair(48, 64, 25.0, 77)
    ref = clean.astype(np.float32)
    p = para()
    got = BM3D(p).process(np.empty_like(n), n, ref)
    assert_close(got, oracle.bm3d_rgb(p.basic, p.final, n, ref, mode=3), clean, label="rgb both ref")


def test_rgb_one_plane_unfiltered(gpu):
    clean, n = synth.rgb_pair(40, 48, 25.0, 5)
    p = para()
    for st in (p.basic, p.final):
        st.sigma = [25.0, 0.0, 25.0]
    assert_close(BM3D(p)(n), oracle.bm3d_rgb(p.basic, p.final, n), clean, label="rgb sigma_U=0")


@pytest.mark.parametrize("name", ["grey64_np_s3r12", "grey96x80_np", "grey72_fast", "rgb48x40_np_s3r12",
                                  "rgb56x64_lc"])
def test_golden_fixtures(gpu, name):
    say(f"golden {name}")
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    from tests.golden.make_golden import para_for
    prof = str(z["profile"])
    ov = tuple(int(v) for v in z["overrides"])
    p = para_for(prof, float(z["sigma"]), ov if ov != (0, 0) else None)
    f = BM3D(p)
    out = f(z["noisy"])
    assert_close(out, z["out"], label=f"golden {name}")
    # step-1 groups of the run itself equal the fixture's
    c1, p1 = f.last_groups(False, z["counts1"].shape, with_pos=True)
    assert group_identity(c1, p1, z["counts1"], z["pos1"]) == 1.0
    c2, p2 = f.last_groups(True, z["counts2"].shape, with_pos=True)
    frac2 = group_identity(c2, p2, z["counts2"], z["pos2"])
    say(f"  step-2 groups (GPU basic vs oracle basic as guide) identical {frac2:.4f}")
    assert frac2 >= 0.98


def test_config0_512_grey(gpu):
    # the reference's CPU-runnable case: 512x512 grey, sigma 25, s=3, R=12, G 16/32
    say("512x512 grey two-step")
    clean, n = synth.grey_pair(512, 512, 25.0, 0)
    p = para()
    f = BM3D(p)
    got = f(n)
    oracle.set_threads(max(1, min(16, os.cpu_count() or 1)), fast=False)
    ref = oracle.bm3d_plane(p.basic, p.final, n)
    oracle.set_threads(1)
    assert_close(got, ref, clean, label="512 grey")
    c1, p1 = f.last_groups(False, (169, 169), with_pos=True)
    oc, op, _ = oracle.match_all(p.basic, n)
    assert group_identity(c1, p1, oc, op) >= MIN_GROUPS


# ---------------------------------------------------------------- edges
WIENER_CAP_CASES = {
    "np s3 R12 G32": dict(),
    "s1 R19 G32": dict(step=1, R=19, G2=32),   # the most contributions per pixel: weight cap 2048
    "high": dict(profile="high", step=None, R=None),
}


@pytest.mark.parametrize("case", list(WIENER_CAP_CASES))
@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_wiener_guide_without_energy(gpu, prec, case):
    # a Wiener guide of 1e-6 * clean: every gain is ~1e-10 and the group
    # weight 1 / sum(W^2) ~ 1e17, beyond what the 64-bit fixed-point sums
    # hold; every weight is capped (isp_bm3d_aggregation_limits), so every
    # pixel's sums reach their largest value without overflowing, and the
    # output stays the near-zero estimate the FP64 oracle gives
    print(f"tiny Wiener guide {prec} {case}", flush=True)
    clean, n = synth.grey_pair(48, 64, 25.0, 21)
    p = para(**WIENER_CAP_CASES[case])
    guide = (clean * 1e-6).astype(np.float32)
    got = BM3D_Final(p.final, precision=prec).process(np.empty_like(n), n, guide)
    want = oracle.kernel(p.final, n, guide, wiener=True)
    print(f"{prec}: max |want| = {np.abs(want).max():.3g}", flush=True)
    assert np.all(np.isfinite(got))
    assert_close(got, want, label=f"tiny Wiener guide {prec} {case}")


def test_degenerate_copy_through(gpu):
    _, n = synth.grey_pair(32, 40, 25.0, 1)
    p = para(sigma=0.0)
    np.testing.assert_array_equal(BM3D(p)(n), n)
    rgb = synth.rgb_pair(24, 24, 25.0, 1)[1]
    p = para()
    p.basic.GroupSize = 1  # process_Frame's wrapper check (source/BM3D.cpp:284-290)
    np.testing.assert_array_equal(BM3D_Basic(p.basic)(rgb), rgb)


def test_constant_plane(gpu):
    p = para()
    src = np.full((40, 48), 0.4, np.float32)
    out = BM3D_Basic(p.basic)(src)
    np.testing.assert_allclose(out, 0.4, atol=1e-6)


def test_errors(gpu):
    p = para()
    with pytest.raises(L.BM3DError) as e:
        BM3D(p)(np.zeros((6, 20), np.float32))
    assert e.value.code == L.E_SIZE
    # positions pack the row in 15 bits: a 16 x 40000 plane is refused, not misread
    with pytest.raises(L.BM3DError) as e:
        BM3D(para())(np.zeros((40000, 16), np.float32))
    assert e.value.code == L.E_UNSUPPORTED
    p = para("vn", step=None, R=None)
    p.final.BlockSize = 17  # above the generic kernels' 16
    with pytest.raises(L.BM3DError) as e:
        BM3D(p)(np.zeros((32, 32), np.float32))
    assert e.value.code == L.E_UNSUPPORTED


def test_device_and_host_agree(gpu):
    import torch
    clean, n = synth.rgb_pair(64, 80, 25.0, 9)
    p = para()
    f = BM3D(p)
    host = f(n)
    x = torch.from_numpy(n).to(gpu)
    dev = f(x)
    torch.cuda.synchronize()
    np.testing.assert_allclose(dev.cpu().numpy(), host, atol=1e-5)


def test_large_host_call_after_device_calls(gpu):
    # host buffers (staged through the context's workspace) after device calls
    # on the same context give the device result
    import torch
    clean, n = synth.rgb_pair(540, 960, 25.0, 19)
    f = BM3D(para(), precision="fp16")
    x = torch.from_numpy(n).to(gpu)
    dev = f(x).cpu().numpy()
    dev2 = f(x).cpu().numpy()
    np.testing.assert_array_equal(dev, dev2)
    for k in range(2):
        host = f(n)
        say(f"  host call {k}: max |host - device| = {np.abs(host - dev).max():.3g}")
        np.testing.assert_array_equal(host, dev)


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_bit_repeatable(gpu, prec):
    # aggregation sums are 64-bit fixed point: the order of the atomic adds
    # cannot change them, so repeated runs are bit-identical
    clean, n = synth.rgb_pair(96, 128, 25.0, 2)
    f = BM3D(para(), precision=prec)
    a, b = f(n), f(n)
    np.testing.assert_array_equal(a, b)
