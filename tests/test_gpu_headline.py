"""Parity of the FP32-distance forms against the FP64 oracle: the headline
form (FP16 / BF16 MFMA transforms), the FP32-throughout form (fp32fast:
FP32 distances, FP32 transforms without FP64 re-decisions) and the same
arithmetic with FP32-accurate products on the matrix cores (fp32mfma).

Matching: the FP32-distance kernels (column walks for BlockStep 3..7, the list
kernel otherwise) must reproduce the oracle's ordered groups
(BlockMatchingMulti, include/Block.h:578-711) for at least 99.9% of the grid
points on the same guide.  End to end at 1920x1080 RGB: per-channel error
against the oracle, and PSNR within 0.05 dB of the plain FP32 GPU form, at
sigma 25 (the benchmark) and sigma 15.
"""
import os

import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Final, synth
from tests.gpu_common import MIN_GROUPS, group_identity, para

pytestmark = pytest.mark.gpu
MAX_DPSNR_MIXED = 0.05
# |mixed - oracle| per channel at 1080p RGB, about twice the measured values
# (sigma 25: fp16 max 0.0097, mean 2.5e-4; bf16 max 0.0128, mean 7.2e-4;
# bf16 keeps 8 mantissa bits against fp16's 11)
MAX_ABS_MIXED = {"fp16": 0.02, "bf16": 0.03}
MEAN_ABS_MIXED = {"fp16": 5e-4, "bf16": 1.5e-3}
# fp32fast against the oracle at 1080p RGB, about twice the measured values
# (sigma 25: max 3.3e-3, 21 of 6.2M samples beyond 2e-3, mean 1.2e-6; sigma
# 15: max 1.6e-3, mean 3.3e-7): FP32 group selection (5 + 2 of 229,401 groups
# differ) and hard-threshold decisions without the FP64 re-decision
MAX_ABS_FAST = 0.007
MEAN_ABS_FAST = 3e-6


def threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)


def key_check(c, keys, oc, ok, G):
    """Keys of the identical groups agree with the FP64 keys to FP32 accuracy."""
    m = (np.arange(G)[None, None, :] < oc[..., None]) & (c == oc)[..., None]
    rel = np.abs(np.where(m, keys - ok, 0.0)) / np.maximum(np.abs(np.where(m, ok, 1.0)), 1e-3)
    return float(rel.max())


@pytest.mark.parametrize("walk", [True, False], ids=["walk", "list"])
@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp32fast"])
def test_fp32_distance_groups_512_grey(gpu, prec, walk, monkeypatch):
    """The 512x512 grey case: BlockStep 3, BMrange 12, G 16 / 32."""
    monkeypatch.setenv("ISP_BM3D_MATCH_WALK", "1" if walk else "0")
    clean, n = synth.grey_pair(512, 512, 25.0, 5)
    p = para()
    f = BM3D(p, precision=prec)
    # step 2 guide: a denoised plane (the oracle's basic estimate)
    oracle.set_threads(threads())
    basic = oracle.kernel(p.basic, n, n, wiener=False).astype(np.float32)
    oracle.set_threads(1)
    for final, guide in ((False, n), (True, basic)):
        st = p.final if final else p.basic
        c, pos, keys = f.match(guide, final_step=final)
        oc, op, ok = oracle.match_all(st, guide)
        frac = group_identity(c, pos, oc, op)
        rel = key_check(c, keys, oc, ok, st.GroupSize)
        say(f"{prec} {'walk' if walk else 'list'} 512^2 step {2 if final else 1}: identical groups "
            f"{frac:.6f} ({int(round((1 - frac) * c.size))} of {c.size} differ), key rel err {rel:.2e}")
        assert frac >= MIN_GROUPS
        assert rel < 1e-4


@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp32fast", "fp32mfma"])
def test_fp32_distance_groups_1080p_rgb(gpu, prec):
    """Groups of the benchmark frame as the fp16 / bf16 form forms them: step 1
    on the FP64 OPP luma of the noisy frame, step 2 (BM3D_Final) on a given
    guide frame's luma."""
    import torch
    clean, noisy = synth.rgb_pair(1080, 1920, 25.0, 1)
    p = para()
    f = BM3D(p, precision=prec)
    x = torch.from_numpy(noisy).to(gpu)
    f(x)
    torch.cuda.synchronize()
    c1, p1 = f.last_groups(False, (359, 639), with_pos=True)
    yuv = oracle.rgb2opp(noisy)
    oc, op, _ = oracle.match_all(p.basic, yuv[0])
    frac1 = group_identity(c1, p1, oc, op)
    say(f"{prec} 1080p step 1: identical groups {frac1:.6f} "
        f"({int(round((1 - frac1) * c1.size))} of {c1.size} differ)")
    assert frac1 >= MIN_GROUPS
    # step 2 on a fixed guide: the clean frame plus a little noise
    guide = (clean + np.random.default_rng(2).standard_normal(clean.shape) * (5 / 255)).astype(np.float32)
    ff = BM3D_Final(p.final, precision=prec)
    g = torch.from_numpy(guide).to(gpu)
    ff.process(torch.empty_like(x), x, g)
    torch.cuda.synchronize()
    c2, p2 = ff.last_groups(True, (359, 639), with_pos=True)
    gy = oracle.rgb2opp(guide)[0]
    oc2, op2, _ = oracle.match_all(p.final, gy)
    frac2 = group_identity(c2, p2, oc2, op2)
    say(f"{prec} 1080p step 2: identical groups {frac2:.6f} "
        f"({int(round((1 - frac2) * c2.size))} of {c2.size} differ)")
    assert frac2 >= MIN_GROUPS


@pytest.mark.parametrize("sigma", [25.0, 15.0])
def test_fp16_1080p_rgb_vs_oracle(gpu, sigma):
    """The mixed-precision 1080p RGB case at sigma 15 and the benchmark (sigma 25): the mixed forms at
    1920x1080 RGB against the FP64 oracle, per channel, and against the plain
    FP32 GPU form in PSNR."""
    import torch
    clean, noisy = synth.rgb_pair(1080, 1920, sigma, 1)
    p = para(sigma=sigma)
    x = torch.from_numpy(noisy).to(gpu)
    outs = {}
    for prec in ("fp32", "fp32fast", "fp32mfma", "fp16", "bf16"):
        outs[prec] = BM3D(p, precision=prec)(x).cpu().numpy().astype(np.float64)
    say(f"1080p sigma {sigma}: oracle")
    oracle.set_threads(threads())
    fast = oracle.host_supports_fast()
    oracle.set_threads(threads(), fast=fast)
    ref = oracle.bm3d_rgb(p.basic, p.final, noisy, fast=fast)
    oracle.set_threads(1)
    oracle.set_threads(1, fast=fast)
    p_ref = synth.psnr(ref, clean)
    p32 = synth.psnr(outs["fp32"], clean)
    say(f"1080p sigma {sigma}: PSNR noisy {synth.psnr(noisy, clean):.3f} oracle {p_ref:.4f} fp32 {p32:.4f}")
    assert abs(p32 - p_ref) <= 0.01
    for prec in ("fp32fast", "fp32mfma"):
        out = outs[prec]
        pf = synth.psnr(out, clean)
        mx = [float(np.abs(out[c] - ref[c]).max()) for c in range(3)]
        mean = [float(np.abs(out[c] - ref[c]).mean()) for c in range(3)]
        big = int(sum((np.abs(out[c] - ref[c]) > 2e-3).sum() for c in range(3)))
        say(f"  {prec}: PSNR {pf:.4f} ({pf - p_ref:+.5f} dB vs oracle); |out - oracle| per channel "
            f"R,G,B max {mx[0]:.3g} {mx[1]:.3g} {mx[2]:.3g} mean {mean[0]:.3g} {mean[1]:.3g} {mean[2]:.3g}; "
            f"{big} samples beyond 2e-3")
        assert abs(pf - p_ref) <= 0.01
        assert max(mx) <= MAX_ABS_FAST
        assert max(mean) <= MEAN_ABS_FAST
    for prec in ("fp16", "bf16"):
        out = outs[prec]
        pm = synth.psnr(out, clean)
        mx = [float(np.abs(out[c] - ref[c]).max()) for c in range(3)]
        mean = [float(np.abs(out[c] - ref[c]).mean()) for c in range(3)]
        say(f"  {prec}: PSNR {pm:.4f} (loss vs fp32 {p32 - pm:+.4f} dB); |out - oracle| per channel "
            f"R,G,B max {mx[0]:.3g} {mx[1]:.3g} {mx[2]:.3g} mean {mean[0]:.3g} {mean[1]:.3g} {mean[2]:.3g}")
        assert p32 - pm <= MAX_DPSNR_MIXED
        assert max(mx) <= MAX_ABS_MIXED[prec]
        assert max(mean) <= MEAN_ABS_MIXED[prec]
