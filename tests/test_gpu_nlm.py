"""NL-means on the GPU (include/isp_nlm.h) against the FP64 oracle
(orc_nlm_*, source/NLMeans.cpp).  The GPU forms the reference's groups
exactly and averages in FP64 in the reference's order; outputs differ from
the oracle only by the last bits of exp() and of the 64-bit fixed-point
aggregation: float planes within 1e-6, integer outputs equal except one LSB
where the value lies at a rounding tie."""
import numpy as np
import pytest

import oracle
from isp_mw_amd import NLMeans, NLMeans_Para, synth
from isp_mw_amd import _lib as L

pytestmark = pytest.mark.gpu

MAX_ABS_F32 = 1e-6


def smooth_noisy(h, w, sigma, seed, lo=0.3):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    clean = lo + 0.2 * np.sin(xx / 9.0) * np.cos(yy / 11.0) + 0.1 * xx / w
    return clean, clean + rng.normal(0, sigma / 255.0, (h, w))


def opar(para: NLMeans_Para):
    return oracle.nlm_params(para.correction, para.sigma, para.strength, para.BlockSize, para.BlockStep,
                             para.GroupSize, para.BMrange, para.BMstep, para.thMSE)


def check_int(got, want, label, bound=1e-4):
    d = np.abs(got.astype(np.int64) - want.astype(np.int64))
    frac = float((d > 0).mean())
    print(f"{label}: max |gpu-oracle| = {d.max()} LSB, differing {frac:.2e}", flush=True)
    assert d.max() <= 1 and frac <= bound, (label, d.max(), frac)


CASES = [dict(), dict(correction=False, thMSE=1500.0), dict(BlockStep=3, BMstep=1, BMrange=8, GroupSize=8),
         dict(BlockStep=1, BMstep=2, BMrange=6, GroupSize=2), dict(BlockStep=8, BMstep=2, BMrange=12, GroupSize=32),
         dict(BlockStep=4, BMstep=3, BMrange=24, GroupSize=16, correction=False, strength=1e6, thMSE=2000.0)]


@pytest.mark.parametrize("kw", CASES, ids=[str(i) for i in range(len(CASES))])
def test_plane_f32(gpu, kw):
    para = NLMeans_Para(sigma=20.0, **kw)
    print("case", para, flush=True)
    clean, n = smooth_noisy(61, 77, 20.0, seed=len(kw))
    src = n.astype(np.float32)
    got = NLMeans(para)(src)
    want = oracle.nlm_plane(opar(para), src.astype(np.float64))
    err = float(np.abs(got - want).max())
    print(f"plane f32: max|gpu-oracle| = {err:.3g}, PSNR {synth.psnr(src, clean):.2f} -> "
          f"{synth.psnr(got, clean):.2f} dB", flush=True)
    assert err <= MAX_ABS_F32


def test_plane_f32_device_and_guide(gpu):
    import torch
    para = NLMeans_Para(sigma=15.0)
    clean, n = smooth_noisy(48, 64, 15.0, seed=3)
    src = n.astype(np.float32)
    g = clean.astype(np.float32)
    f = NLMeans(para)
    host = f(src, g)
    dev = f(torch.from_numpy(src).to(gpu), torch.from_numpy(g).to(gpu))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), host)
    want = oracle.nlm_plane(opar(para), src.astype(np.float64), g.astype(np.float64))
    assert float(np.abs(host - want).max()) <= MAX_ABS_F32
    assert float(np.abs(host - f(src)).max()) > 1e-3       # the guide changes the groups


@pytest.mark.parametrize("q", [(0, 0, 255), (16, 16, 235), (0, 0, 65535), (16, 128, 240)],
                         ids=["pc8", "tv8", "pc16", "tv8chroma"])
@pytest.mark.parametrize("corr", [True, False])
def test_plane_i32(gpu, q, corr):
    floor, neutral, ceil = q
    para = NLMeans_Para(correction=corr, sigma=20.0, thMSE=None if corr else 1500.0)
    _, n = smooth_noisy(50, 66, 20.0, seed=ceil + corr)
    src = np.rint(n * (ceil - floor) + floor).astype(np.int32)
    src[0, :3] = [floor - 5, ceil + 9, neutral]                 # clipped on input
    got = NLMeans(para)(src, floor=floor, neutral=neutral, ceil=ceil)
    want = oracle.nlm_plane_int(opar(para), src, floor=floor, neutral=neutral, ceil=ceil)
    check_int(got, want, f"plane i32 {q} correction {corr}")


@pytest.mark.parametrize("q", [(0, 255), (16, 235), (0, 65535)], ids=["pc8", "tv8", "pc16"])
@pytest.mark.parametrize("corr", [True, False])
def test_frame_i32(gpu, q, corr):
    floor, ceil = q
    para = NLMeans_Para(correction=corr, sigma=12.0)
    clean, n = synth.rgb_pair(56, 72, 12.0, 5 + corr)
    src = np.clip(np.rint(n * (ceil - floor) + floor), floor, ceil).astype(np.int32)
    got = NLMeans(para)(src, floor=floor, ceil=ceil)
    want = oracle.nlm_rgb_int(opar(para), src, floor=floor, ceil=ceil)
    check_int(got, want, f"frame i32 {q} correction {corr}")


def test_frame_i32_guide_and_device(gpu):
    import torch
    para = NLMeans_Para(sigma=10.0)
    clean, n = synth.rgb_pair(40, 48, 10.0, 9)
    src = np.clip(np.rint(n * 255), 0, 255).astype(np.int32)
    ref = np.clip(np.rint(clean * 255), 0, 255).astype(np.int32)
    f = NLMeans(para)
    host = f(src, ref, floor=0, ceil=255)
    dev = f(torch.from_numpy(src).to(gpu), torch.from_numpy(ref).to(gpu), floor=0, ceil=255)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), host)
    check_int(host, oracle.nlm_rgb_int(opar(para), src, ref, floor=0, ceil=255), "frame i32 guide")


def test_degenerate_parameters_copy(gpu):
    _, n = smooth_noisy(24, 30, 10.0, seed=1)
    src = n.astype(np.float32)
    for kw in (dict(GroupSize=1), dict(strength=0.0), dict(thMSE=-1.0), dict(BMrange=2, BMstep=3)):
        np.testing.assert_array_equal(NLMeans(NLMeans_Para(**kw))(src), src)
    fi = np.clip(np.rint(np.stack([n, n, n]) * 255), 0, 255).astype(np.int32)
    np.testing.assert_array_equal(NLMeans(NLMeans_Para(GroupSize=1))(fi, floor=0, ceil=255), fi)


def test_uncovered_pixels(gpu):
    # BlockStep 10 > BlockSize 8: pixels no block covers are 0/0 as in the reference
    para = NLMeans_Para(sigma=20.0, BlockStep=10)
    _, n = smooth_noisy(30, 30, 20.0, seed=4)
    src = n.astype(np.float32)
    got = NLMeans(para)(src)
    want = oracle.nlm_plane(opar(para), src.astype(np.float64))
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert float(np.abs(got[ok] - want[ok]).max()) <= MAX_ABS_F32


def test_unsupported_parameters_fail_loudly(gpu):
    src = np.zeros((32, 32), np.float32)
    for kw, code in ((dict(BlockSize=4), L.E_UNSUPPORTED), (dict(GroupSize=64), L.E_UNSUPPORTED),
                     (dict(BlockStep=0), L.E_ARG)):
        with pytest.raises(L.BM3DError) as e:
            NLMeans(NLMeans_Para(**kw))(src)
        assert e.value.code == code, kw
    with pytest.raises(L.BM3DError):
        NLMeans()(np.zeros((6, 40), np.float32))


def test_fast_matching_precision(gpu):
    # FP32 distances (fp16 / bf16 contexts): a different tie order, same quality
    para = NLMeans_Para(sigma=20.0)
    clean, n = synth.rgb_pair(96, 128, 20.0, 21)
    src = np.clip(np.rint(n * 255), 0, 255).astype(np.int32)
    exact = NLMeans(para)(src, floor=0, ceil=255)
    fast = NLMeans(para, precision="fp16")(src, floor=0, ceil=255)
    c8 = clean * 255
    p_exact = 10 * np.log10(255.0 ** 2 / np.mean((exact - c8) ** 2))
    p_fast = 10 * np.log10(255.0 ** 2 / np.mean((fast - c8) ** 2))
    print(f"fp32-distance matching: {p_fast:.4f} dB vs exact {p_exact:.4f} dB", flush=True)
    assert abs(p_fast - p_exact) <= 0.05


def test_full_hd_frame(gpu):
    # 1920x1080 RGB, 16-bit samples (the CLI's frame), default parameters
    para = NLMeans_Para(sigma=8.0)
    _, n = synth.rgb_pair(1080, 1920, 8.0, 3)
    src = np.clip(np.rint(n * 255), 0, 255).astype(np.int32) * 257
    got = NLMeans(para)(src, floor=0, ceil=65535)
    want = oracle.nlm_rgb_int(opar(para), src, floor=0, ceil=65535)
    check_int(got, want, "frame 1080p")
