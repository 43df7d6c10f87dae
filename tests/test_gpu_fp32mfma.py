"""The FP32_MFMA form (ISP_BM3D_PREC_FP32_MFMA): the FP32-throughout form's
arithmetic with every transform product on the FP16 matrix cores as three
MFMAs of a high / low operand split (22 significant bits, FP32 accumulation).

Its products differ from FP32 VALU sums by rounding only (about 1e-6 of the
coefficients), so its outputs must track the fp32fast form's: the same
matching, decisions that differ only where a coefficient lies within rounding
of the hard threshold.  Against the FP64 oracle it is held to the fp32fast
bars (tests/test_gpu_headline.py covers 1920x1080 RGB).  The plain form's
transforms run on the same split products with its FP64 re-decision; its
vector-ALU kernel stays selectable and is checked beside it.
"""
import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Basic, BM3DBatch, synth
from tests.gpu_common import para

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def run(p, x, prec, cls=BM3D):
    f = cls(p if cls is BM3D else p.basic, precision=prec)
    return f(x).cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("mode", ["basic", "both"])
@pytest.mark.parametrize("rgb", [False, True], ids=["grey", "rgb"])
def test_tracks_the_fp32fast_form(gpu, mode, rgb):
    import torch
    clean, n = (synth.rgb_pair if rgb else synth.grey_pair)(256, 320, 25.0, 21)
    p = para()
    x = torch.from_numpy(n).to(gpu)
    cls = BM3D_Basic if mode == "basic" else BM3D
    fast, mx = run(p, x, "fp32fast", cls), run(p, x, "fp32mfma", cls)
    d = np.abs(mx - fast)
    dp = synth.psnr(mx, clean) - synth.psnr(fast, clean)
    say(f"{'rgb' if rgb else 'grey'} {mode}: |fp32mfma - fp32fast| max {d.max():.3g} mean {d.mean():.3g}; "
        f"{int((d > 1e-4).sum())} of {d.size} beyond 1e-4; dPSNR {dp:+.5f} dB")
    assert np.all(np.isfinite(mx))
    assert abs(dp) <= 0.002
    assert d.mean() <= 2e-6
    assert np.mean(d > 1e-4) <= 1e-3


def test_plane_against_the_oracle(gpu):
    import torch
    clean, n = synth.grey_pair(96, 128, 25.0, 4)
    p = para()
    got = run(p, torch.from_numpy(n).to(gpu), "fp32mfma")
    want = oracle.bm3d_plane(p.basic, p.final, n)
    err = float(np.abs(got - want).max())
    dp = abs(synth.psnr(got, clean) - synth.psnr(want, clean))
    say(f"96x128 grey: max |fp32mfma - oracle| {err:.3g}, dPSNR {dp:.5f} dB")
    assert err <= 0.007 and dp <= 0.01


def test_large_window_shrinks_the_tile(gpu):
    """BMrange 19 with GroupSize 32: the default tile's LDS with both operand
    halves staged would exceed a CU's, so the tile shrinks; the output still
    tracks the fp32fast form."""
    import torch
    clean, n = synth.rgb_pair(200, 256, 25.0, 8)
    p = para(R=19, G1=32, G2=32)
    x = torch.from_numpy(n).to(gpu)
    fast, mx = run(p, x, "fp32fast"), run(p, x, "fp32mfma")
    d = np.abs(mx - fast)
    dp = synth.psnr(mx, clean) - synth.psnr(fast, clean)
    say(f"BMrange 19: |fp32mfma - fp32fast| max {d.max():.3g} mean {d.mean():.3g}; dPSNR {dp:+.5f} dB")
    assert abs(dp) <= 0.002 and d.mean() <= 2e-6


def test_batch_of_rgb_frames_equals_single_calls(gpu):
    import torch
    shapes, sig = [(48, 64), (72, 56), (40, 40)], [25.0, 12.0, 35.0]
    imgs = [synth.rgb_pair(h, w, s, 90 + k)[1] for k, ((h, w), s) in enumerate(zip(shapes, sig))]
    f = BM3D(para(), precision="fp32mfma")
    srcs = [torch.from_numpy(a).to(gpu) for a in imgs]
    dsts = [torch.empty_like(t) for t in srcs]
    b = BM3DBatch(f, dsts, srcs, sigmas=sig)
    b.run()
    torch.cuda.synchronize()
    for k, a in enumerate(imgs):
        want = run(para(sigma=sig[k]), torch.from_numpy(a).to(gpu), "fp32mfma")
        np.testing.assert_array_equal(dsts[k].cpu().numpy().astype(np.float64), want)


def test_plain_form_on_either_engine(gpu, monkeypatch):
    """The plain form's transforms on the split matrix-core products (the
    default) and on the FP32 vector ALUs (ISP_BM3D_FP32_VALU=1, read when a
    context is made): both within the plain bar of the oracle, and within
    rounding of each other (the FP64 re-decision makes the same hard-threshold
    calls on both)."""
    import torch
    from tests.gpu_common import assert_close
    clean, n = synth.rgb_pair(96, 128, 25.0, 6)
    p = para()
    x = torch.from_numpy(n).to(gpu)
    mx = run(p, x, "fp32")
    monkeypatch.setenv("ISP_BM3D_FP32_VALU", "1")
    va = run(p, x, "fp32")
    want = oracle.bm3d_rgb(p.basic, p.final, n)
    assert_close(mx, want, clean, label="plain form, split products")
    assert_close(va, want, clean, label="plain form, FP32 vector ALUs")
    d = float(np.abs(mx - va).max())
    say(f"plain form: max |split - vector| {d:.3g}")
    assert d <= 1e-5
