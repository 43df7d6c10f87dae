"""Mixed-precision forms (FP16 / BF16 MFMA transforms, FP32 accumulation):
PSNR within 0.05 dB of the plain FP32 GPU path, error against the FP64
oracle reported per channel."""
import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Basic, BM3D_Final, synth
from tests.gpu_common import para

pytestmark = pytest.mark.gpu
MAX_DPSNR_MIXED = 0.05


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("kind", ["grey", "rgb"])
def test_mixed_vs_fp32(gpu, prec, kind):
    print(f"mixed {prec} {kind}", flush=True)
    if kind == "grey":
        clean, n = synth.grey_pair(256, 320, 25.0, 21)
    else:
        clean, n = synth.rgb_pair(192, 256, 25.0, 22)
    p = para()
    ref32 = BM3D(p)(n)
    mix = BM3D(p, precision=prec)(n)
    d = synth.psnr(ref32, clean) - synth.psnr(mix, clean)
    err = np.abs(mix - ref32)
    print(f"  psnr fp32 {synth.psnr(ref32, clean):.3f} {prec} {synth.psnr(mix, clean):.3f} "
          f"loss {d:.4f} dB, max|diff| {err.max():.3g}, mean {err.mean():.3g}", flush=True)
    assert d <= MAX_DPSNR_MIXED
    assert err.mean() < (2e-3 if prec == "fp16" else 8e-3)


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_mixed_single_steps(gpu, prec):
    clean, n = synth.grey_pair(128, 160, 25.0, 23)
    p = para()
    b32 = BM3D_Basic(p.basic)(n)
    bm = BM3D_Basic(p.basic, precision=prec)(n)
    assert synth.psnr(b32, clean) - synth.psnr(bm, clean) <= MAX_DPSNR_MIXED
    guide = b32
    f32 = BM3D_Final(p.final).process(np.empty_like(n), n, guide)
    fm = BM3D_Final(p.final, precision=prec).process(np.empty_like(n), n, guide)
    assert synth.psnr(f32, clean) - synth.psnr(fm, clean) <= MAX_DPSNR_MIXED


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_mixed_small_groups_and_edges(gpu, prec):
    # odd sizes, G=1 bypass and short groups exercise the zero padding
    clean, n = synth.grey_pair(37, 53, 25.0, 24)
    for kw in (dict(), dict(G1=1, G2=1), dict(G1=3, G2=5)):
        p = para(**kw)
        a = BM3D(p)(n)
        b = BM3D(p, precision=prec)(n)
        assert np.all(np.isfinite(b))
        assert synth.psnr(a, clean) - synth.psnr(b, clean) <= 0.1


def test_mixed_error_vs_fp64_oracle_per_channel(gpu):
    clean, n = synth.rgb_pair(96, 128, 25.0, 25)
    p = para()
    ref = oracle.bm3d_rgb(p.basic, p.final, n)
    for prec in ("fp32", "fp16", "bf16"):
        out = BM3D(p, precision=prec)(n)
        errs = [float(np.abs(out[c] - ref[c]).max()) for c in range(3)]
        print(f"{prec}: max|out-oracle| per channel R,G,B = {errs}", flush=True)
        assert max(errs) < {"fp32": 2e-3, "fp16": 0.03, "bf16": 0.05}[prec]


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("tile", [(16, 8), (8, 16), (16, 16), (4, 4)])
def test_filter_tile_shapes_do_not_change_results(gpu, prec, tile):
    """Filter tiles of 128 reference blocks or more run 16-wave workgroups,
    smaller ones 8-wave workgroups; the fixed-point sums are integers, so the
    output equals the default 8x8 tile's bit for bit."""
    clean, n = synth.rgb_pair(150, 200, 25.0, 31)
    p = para()
    want = BM3D(p, precision=prec)(n)
    f = BM3D(p, precision=prec)
    f.set_tuning(f.size_class(200, 150), {"match_tile": (4, 16), "filter_tile": tile, "final_group_size": 0})
    print(f"{prec} filter tile {tile}", flush=True)
    np.testing.assert_array_equal(f(n), want)


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_wide_workgroups_at_bmrange_19(gpu, prec):
    """At BMrange 19 the default 8x8 tile's footprint leaves room for one
    workgroup per CU, which then takes 16 waves; a 4x2 tile fits two 8-wave
    workgroups.  Both give the same bits, single calls and batches alike."""
    import torch
    from isp_mw_amd import BM3DBatch
    clean, n = synth.rgb_pair(120, 160, 25.0, 33)
    p = para(R=19, G1=32, G2=32)
    wide = BM3D(p, precision=prec)(n)
    f = BM3D(p, precision=prec)
    f.set_tuning(f.size_class(160, 120), {"match_tile": (4, 16), "filter_tile": (4, 2), "final_group_size": 0})
    np.testing.assert_array_equal(f(n), wide)
    g = BM3D(p, precision=prec)
    x = torch.from_numpy(n).cuda()
    y = torch.empty_like(x)
    BM3DBatch(g, [y], [x]).run()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), wide)
