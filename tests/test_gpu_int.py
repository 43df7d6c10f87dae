"""Integer Plane / Frame paths (RangeConvert in and out, OPP from and to int)
against the oracle: outputs equal up to +-1 LSB where the float value lies at
a rounding tie."""
import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Basic, BM3D_Final, synth
from tests.gpu_common import para

pytestmark = pytest.mark.gpu


def quantize(x, floor, ceil):
    return np.clip(np.round(x * (ceil - floor) + floor), floor, ceil).astype(np.int32)


def check_int(got, ref, label, levels):
    # FP32 filtering leaves ~3e-7 of the [0,1] range: outputs one LSB apart
    # where the exact value sits within that of a rounding boundary
    d = np.abs(got.astype(np.int64) - ref.astype(np.int64))
    frac = float((d > 0).mean())
    bound = max(1e-3, 2 * 3e-7 * levels)
    print(f"{label}: max |gpu-oracle| = {d.max()} LSB, differing {frac:.2e} (bound {bound:.1e})", flush=True)
    assert d.max() <= 1 and frac <= bound, (label, d.max(), frac)


@pytest.mark.parametrize("q", [(0, 0, 255), (16, 16, 235), (0, 0, 1023), (0, 128, 255), (16, 128, 240)],
                         ids=["pc8", "tv8", "pc10", "pc8chroma", "tv8chroma"])
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_plane_int(gpu, q, mode):
    floor, neutral, ceil = q
    clean, n = synth.grey_pair(64, 80, 25.0, 61 + mode)
    chroma = floor < neutral
    src = quantize(n - (0.5 if chroma else 0.0) + (0.5 if chroma else 0.0), floor, ceil)
    p = para()
    cls = {1: BM3D_Basic, 2: BM3D_Final, 3: BM3D}[mode]
    f = cls(p.basic if mode == 1 else p.final if mode == 2 else p)
    ref = quantize(clean, floor, ceil) if mode == 2 else None
    got = f.process(np.empty_like(src), src, ref, floor=floor, neutral=neutral, ceil=ceil)
    want = oracle.bm3d_plane_int(p.basic, p.final, src, ref, mode=mode, floor=floor, neutral=neutral, ceil=ceil)
    check_int(got, want, f"plane int {q} mode {mode}", ceil - floor)


@pytest.mark.parametrize("q", [(0, 255), (16, 235), (0, 65535)], ids=["pc8", "tv8", "pc16"])
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_frame_int(gpu, q, mode):
    floor, ceil = q
    clean, n = synth.rgb_pair(56, 64, 25.0, 71 + mode)
    src = quantize(n, floor, ceil)
    p = para()
    cls = {1: BM3D_Basic, 2: BM3D_Final, 3: BM3D}[mode]
    f = cls(p.basic if mode == 1 else p.final if mode == 2 else p)
    ref = quantize(clean, floor, ceil) if mode == 2 else None
    got = f.process(np.empty_like(src), src, ref, floor=floor, ceil=ceil)
    want = oracle.bm3d_rgb_int(p.basic, p.final, src, ref, mode=mode, floor=floor, ceil=ceil)
    check_int(got, want, f"frame int {q} mode {mode}", ceil - floor)


def test_frame_int_device(gpu):
    import torch
    clean, n = synth.rgb_pair(48, 64, 25.0, 9)
    src = quantize(n, 0, 255)
    f = BM3D(para())
    host = f(src, floor=0, ceil=255)
    dev = f(torch.from_numpy(src).to(gpu), floor=0, ceil=255).cpu().numpy()
    np.testing.assert_array_equal(host, dev)


def test_frame_int_degenerate_copies(gpu):
    src = quantize(synth.rgb_pair(32, 32, 25.0, 3)[1], 0, 255)
    p = para()
    p.basic.GroupSize = 1  # BM3D_Base::process_Frame early return
    np.testing.assert_array_equal(BM3D_Basic(p.basic)(src, floor=0, ceil=255), src)
