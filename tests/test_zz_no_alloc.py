"""isp_bm3d_plane_f32_noalloc: the allocation-free entry point, called through
the C ABI on device buffers, against isp_bm3d_plane_f32 (its allocating
counterpart) on the same inputs, bit for bit.  Two calls in a row on one
context, the second on a smaller plane; every mode; both engines; a row
stride wider than the plane; a plane too large for the scratch array; and the
device's free memory unchanged by the calls."""
import ctypes

import numpy as np
import pytest

from isp_mw_amd import BM3D, BM3D_Basic, BM3D_Final, synth
from isp_mw_amd import _lib as L
from tests.gpu_common import para

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def noalloc(f, dst, src, ref, w, h, stride, mode):
    import torch
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream(src.device).cuda_stream)
    return L.lib().isp_bm3d_plane_f32_noalloc(f._ctx, ptr(dst), ptr(src), ptr(ref), w, h, stride, mode, stream)


def allocating(f, dst, src, ref, w, h, stride, mode):
    import torch
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream(src.device).cuda_stream)
    return L.lib().isp_bm3d_plane_f32(f._ctx, ptr(dst), ptr(src), ptr(ref), w, h, stride, mode, L.MEM_DEVICE,
                                      stream)


def planes(gpu, h, w, stride, seed):
    """noisy plane and a guide (a lightly noisy clean plane), [h, stride] with w valid columns"""
    import torch
    clean, n = synth.grey_pair(h, w, 25.0, seed)
    g = (clean + np.random.default_rng(seed).standard_normal(clean.shape) * (5 / 255)).astype(np.float32)
    pad = lambda a: np.pad(a, ((0, 0), (0, stride - w)), constant_values=0.25).astype(np.float32)  # noqa: E731
    return torch.from_numpy(pad(n)).to(gpu), torch.from_numpy(pad(g)).to(gpu)


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
@pytest.mark.parametrize("mode", [L.BOTH, L.BASIC, L.FINAL], ids=["both", "basic", "final"])
def test_noalloc_equals_allocating_entry_point(gpu, prec, mode):
    import torch
    p = para()
    cls = {L.BOTH: BM3D, L.BASIC: BM3D_Basic, L.FINAL: BM3D_Final}[mode]
    f = cls(p if mode == L.BOTH else (p.basic if mode == L.BASIC else p.final), precision=prec)
    # a large plane, then a smaller one with a wider stride, on the same context
    for (h, w, stride, seed) in ((720, 1280, 1280, 3), (136, 200, 256, 4)):
        say(f"noalloc {prec} mode {mode}: {w}x{h} stride {stride}")
        src, guide = planes(gpu, h, w, stride, seed)
        ref = guide if mode == L.FINAL else None
        want = torch.full_like(src, -1.0)
        got = torch.full_like(src, -2.0)
        assert allocating(f, want, src, ref, w, h, stride, mode) == L.OK
        assert noalloc(f, got, src, ref, w, h, stride, mode) == L.OK, L.lib().isp_bm3d_last_error(f._ctx)
        torch.cuda.synchronize()
        a, b = want[:, :w].cpu().numpy(), got[:, :w].cpu().numpy()
        assert np.all(np.isfinite(b))
        say(f"  max |noalloc - allocating| = {float(np.abs(a - b).max()):.3g}")
        np.testing.assert_array_equal(b, a)
        # the padding columns are not written
        assert torch.all(got[:, w:] == -2.0)


def test_noalloc_separate_guide_and_repeat(gpu):
    """BOTH with a separate step-1 guide; the second call of the same plane
    after a different one gives the same bits (nothing left in the scratch
    array changes a result)."""
    import torch
    f = BM3D(para(), precision="fp16")
    src, guide = planes(gpu, 300, 400, 400, 7)
    want = torch.empty_like(src)
    assert allocating(f, want, src, guide, 400, 300, 400, L.BOTH) == L.OK
    outs = []
    for (h, w, seed) in ((300, 400, 7), (96, 120, 8), (300, 400, 7)):
        s2, g2 = planes(gpu, h, w, w, seed)
        o = torch.empty_like(s2)
        assert noalloc(f, o, s2, g2, w, h, w, L.BOTH) == L.OK
        outs.append(o)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(outs[0].cpu().numpy(), want.cpu().numpy())
    np.testing.assert_array_equal(outs[2].cpu().numpy(), want.cpu().numpy())


def test_noalloc_rejects_planes_beyond_the_scratch(gpu):
    import torch
    f = BM3D(para())
    x = torch.zeros((2048, 4096), dtype=torch.float32, device=gpu)
    rc = noalloc(f, x, x, None, 4096, 2048, 4096, L.BOTH)
    msg = L.lib().isp_bm3d_last_error(f._ctx).decode()
    say(f"4096x2048: rc {rc}: {msg}")
    assert rc == L.E_UNSUPPORTED and "scratch" in msg
    assert noalloc(f, x, x, None, 4, 4, 4, L.BOTH) == L.E_SIZE
    assert noalloc(f, None, x, None, 64, 64, 64, L.BOTH) == L.E_ARG


def test_noalloc_leaves_free_memory_unchanged(gpu):
    import torch
    src, _ = planes(gpu, 512, 640, 640, 9)
    out = torch.empty_like(src)
    # load the kernels and the runtime's per-kernel resources first, on another context
    warm = BM3D(para(), precision="fp16")
    assert allocating(warm, out, src, None, 640, 512, 640, L.BOTH) == L.OK
    f = BM3D(para(), precision="fp16")
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info(gpu)
    for _ in range(2):
        assert noalloc(f, out, src, None, 640, 512, 640, L.BOTH) == L.OK
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info(gpu)
    say(f"free device memory before {free0} after {free1}")
    assert free1 == free0
