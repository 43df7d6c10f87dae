"""Multi-frame grouping on the GPU (isp_bm3d_stack_f32) against the oracle's
orc_bm3d_stack and against the single-frame path: one frame or radius 0 give
the single-frame result bit for bit; radius >= 1 stays within the plain FP32
bar (max-abs 2e-3, |dPSNR| 0.01 dB) of the FP64 oracle."""
import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Basic, BM3D_Final, synth
from isp_mw_amd import _lib as L
from tests.gpu_common import MAX_ABS, MAX_DPSNR, para

pytestmark = pytest.mark.gpu


def grey_stack(T, h, w, sigma, seed, shift=(0, 1)):
    big = synth.clean_grey(h + 8 * T, w + 8 * T, seed)
    rng = np.random.default_rng(seed)
    cl = np.stack([big[4 + t * shift[0]:4 + t * shift[0] + h, 4 + t * shift[1]:4 + t * shift[1] + w]
                   for t in range(T)])
    return cl, (cl + rng.normal(0, sigma / 255.0, cl.shape)).astype(np.float32)


def rgb_stack(T, h, w, sigma, seed):
    c = synth.clean_rgb(h, w, seed)
    rng = np.random.default_rng(seed)
    cl = np.stack([np.roll(c, t, axis=2) for t in range(T)])
    return cl, (cl + rng.normal(0, sigma / 255.0, cl.shape)).astype(np.float32)


def check(got, want, clean, label):
    err = float(np.abs(got - want).max())
    d = abs(synth.psnr(got, clean) - synth.psnr(want, clean))
    print(f"{label}: max|gpu-oracle| {err:.3g}, dPSNR {d:.5f} dB, PSNR {synth.psnr(got, clean):.3f}", flush=True)
    assert err <= MAX_ABS and d <= MAX_DPSNR, (label, err, d)


@pytest.mark.parametrize("prec", ["fp32", "fp16", "fp32fast", "fp32mfma"])
def test_one_frame_and_radius_zero_equal_single_frame_calls(gpu, prec):
    p = para(R=8)
    _, g = grey_stack(3, 48, 64, 25.0, 3)
    f = BM3D(p, precision=prec)
    one = f.process_stack(np.empty_like(g[:1]), g[:1], radius=2)
    np.testing.assert_array_equal(one[0], f(g[0]))
    r0 = f.process_stack(np.empty_like(g), g, radius=0)
    for t in range(3):
        np.testing.assert_array_equal(r0[t], f(g[t]))
    _, c = rgb_stack(2, 40, 56, 25.0, 4)
    r0 = f.process_stack(np.empty_like(c), c, radius=0)
    for t in range(2):
        np.testing.assert_array_equal(r0[t], f(c[t]))


@pytest.mark.parametrize("radius", [1, 2])
def test_grey_stack_against_oracle(gpu, radius):
    p = para(R=8)
    cl, g = grey_stack(4, 48, 64, 25.0, 5)
    got = BM3D(p).process_stack(np.empty_like(g), g, radius=radius)
    want = oracle.bm3d_stack(p.basic, p.final, g.astype(np.float64), radius)
    check(got, want, cl, f"grey stack radius {radius}")
    r0 = oracle.bm3d_stack(p.basic, p.final, g.astype(np.float64), 0)
    print(f"PSNR radius 0 {synth.psnr(r0, cl):.3f} dB", flush=True)


def test_largest_radius_against_oracle(gpu):
    # radius 4: nine lists per reference block, up to 9 x 31 = 279 candidates
    # in the merge (more than one per lane of a wave), GroupSize 32 in both steps
    p = para(R=8, G1=32, G2=32)
    cl, g = grey_stack(9, 32, 40, 25.0, 10)
    for prec in ("fp32", "fp16"):
        got = BM3D(p, precision=prec).process_stack(np.empty_like(g), g, radius=4)
        if prec == "fp32":
            want = oracle.bm3d_stack(p.basic, p.final, g.astype(np.float64), 4)
            check(got, want, cl, "grey stack radius 4")
            ref32 = got
        else:
            d = synth.psnr(ref32, cl) - synth.psnr(got, cl)
            print(f"fp16 stack radius 4: {d:.4f} dB below fp32", flush=True)
            assert d <= 0.05


def test_rgb_stack_against_oracle(gpu):
    p = para(R=8)
    cl, c = rgb_stack(3, 40, 56, 25.0, 6)
    got = BM3D(p).process_stack(np.empty_like(c), c, radius=1)
    want = oracle.bm3d_stack(p.basic, p.final, c.astype(np.float64), 1)
    check(got, want, cl, "rgb stack radius 1")


@pytest.mark.parametrize("mode", [1, 2])
def test_single_steps_with_guides(gpu, mode):
    p = para(R=8)
    cl, g = grey_stack(3, 40, 48, 25.0, 7)
    ref = (cl + 0.01).astype(np.float32)
    f = BM3D_Basic(p.basic) if mode == 1 else BM3D_Final(p.final)
    got = f.process_stack(np.empty_like(g), g, radius=1, ref=ref)
    want = oracle.bm3d_stack(p.basic, p.final, g.astype(np.float64), 1, ref=ref.astype(np.float64), mode=mode)
    check(got, want, cl, f"grey stack mode {mode} with guide")


def test_device_equals_host_and_fp16_quality(gpu):
    import torch
    p = para(R=8)
    cl, c = rgb_stack(3, 40, 56, 25.0, 8)
    f = BM3D(p)
    host = f.process_stack(np.empty_like(c), c, radius=1)
    x = torch.from_numpy(c).to(gpu)
    dev = f.process_stack(torch.empty_like(x), x, radius=1)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), host)
    h16 = BM3D(p, precision="fp16").process_stack(np.empty_like(c), c, radius=1)
    d = synth.psnr(host, cl) - synth.psnr(h16, cl)
    print(f"fp16 stack: {d:.4f} dB below fp32", flush=True)
    assert d <= 0.05


def test_radius_limits(gpu):
    p = para(R=8)
    _, g = grey_stack(2, 32, 32, 25.0, 9)
    with pytest.raises(L.BM3DError) as e:
        BM3D(p).process_stack(np.empty_like(g), g, radius=5)
    assert e.value.code == L.E_UNSUPPORTED
