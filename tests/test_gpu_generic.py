"""The generic kernels: BlockSize other than 8 (profile vn's final step uses
11, include/BM3D.h:127-130; the CLI's -BS1 / -BS2) and GroupSize above 32
(-GS1 / -GS2), against the FP64 oracle.  They compute distances, transforms
and shrinkage in FP64, so groups are identical and outputs agree to the
fixed-point resolution of the aggregation."""
import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Basic, BM3D_Final, BM3D_Para, synth
from isp_mw_amd import _lib as L
from tests.gpu_common import assert_close, group_identity

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def gpara(B1=8, B2=8, G1=16, G2=32, step=3, R=8, sigma=25.0, bmstep=1, profile="np"):
    p = BM3D_Para(profile)
    for st in (p.basic, p.final):
        st.BlockStep, st.BMrange, st.BMstep = step, R, bmstep
    p.basic.BlockSize, p.final.BlockSize = B1, B2
    p.basic.GroupSize, p.final.GroupSize = G1, G2
    p.set_sigma(sigma)
    return p


def smooth_pair(h, w, seed, sigma=10.0):
    """smooth content with weak noise: full groups"""
    y, x = np.mgrid[0:h, 0:w]
    clean = 0.4 + 0.1 * np.sin(x / 7.0) * np.cos(y / 9.0) + 0.002 * x
    n = clean + np.random.default_rng(seed).standard_normal((h, w)) * (sigma / 255)
    return clean, n.astype(np.float32)


CASES = {
    "B11 G32": dict(B1=11, B2=11, G1=32, G2=32, step=6, R=10),
    "B5 G48": dict(B1=5, B2=5, G1=48, G2=48, step=2, R=6),
    "B16 G24": dict(B1=16, B2=16, G1=24, G2=24, step=5, R=8),
    "B8 G64": dict(B1=8, B2=8, G1=64, G2=64, step=3, R=7),
    "B3 G40 BMstep2": dict(B1=3, B2=3, G1=40, G2=40, step=2, R=8, bmstep=2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_generic_groups_identical(gpu, name):
    p = gpara(**CASES[name])
    for kind in ("smooth", "synth"):
        _, n = smooth_pair(70, 66, 3) if kind == "smooth" else synth.grey_pair(70, 66, 25.0, 3)
        f = BM3D(p)
        for final in (False, True):
            st = p.final if final else p.basic
            c, pos, keys = f.match(n, final_step=final)
            oc, op, ok = oracle.match_all(st, n)
            frac = group_identity(c, pos, oc, op)
            say(f"{name} {kind} step {2 if final else 1}: identical {frac:.5f}, mean N {c.mean():.2f}, "
                f"max N {c.max()}")
            assert frac == 1.0
            m = np.arange(st.GroupSize)[None, None, :] < oc[..., None]
            np.testing.assert_allclose(np.where(m, keys, 0), np.where(m, ok, 0), rtol=1e-12, atol=1e-12)
        if kind == "smooth":
            assert c.max() == p.final.GroupSize  # full groups are exercised


@pytest.mark.parametrize("mode", ["basic", "final", "both"])
@pytest.mark.parametrize("name", list(CASES))
def test_generic_plane_vs_oracle(gpu, name, mode):
    p = gpara(**CASES[name])
    clean, n = synth.grey_pair(72, 80, 25.0, 11)
    say(f"generic plane {name} {mode}")
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
    assert_close(got, ref, clean, max_abs=1e-6, label=f"{name} {mode}")


@pytest.mark.parametrize("sigma", [10.0, 30.0])
def test_profile_vn_plane_and_rgb(gpu, sigma):
    """Profile vn: basic BlockSize 8 / GroupSize 32 (the tuned kernels), final
    BlockSize 11 (the generic kernels)."""
    p = BM3D_Para("vn").set_sigma(sigma)
    assert p.final.BlockSize == 11
    clean, n = synth.grey_pair(96, 104, sigma, 17)
    assert_close(BM3D(p)(n), oracle.bm3d_plane(p.basic, p.final, n), clean, label=f"vn plane sigma {sigma}")
    clean3, n3 = synth.rgb_pair(64, 72, sigma, 18)
    assert_close(BM3D(p)(n3), oracle.bm3d_rgb(p.basic, p.final, n3), clean3, label=f"vn rgb sigma {sigma}")


def test_profile_vn_integer_frame(gpu):
    p = BM3D_Para("vn").set_sigma(20.0)
    _, n = synth.rgb_pair(48, 56, 20.0, 19)
    frame = np.clip(np.rint(n * 255), 0, 255).astype(np.int32)
    got = BM3D(p)(frame, floor=0, ceil=255)
    want = oracle.bm3d_rgb_int(p.basic, p.final, frame, mode=3, floor=0, ceil=255)
    d = np.abs(got.astype(np.int64) - want)
    say(f"vn int frame: max {d.max()} levels, {np.mean(d > 0):.2e} differ")
    assert d.max() <= 1 and np.mean(d > 0) < 1e-3


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_generic_steps_in_mixed_contexts(gpu, prec):
    """Steps the tuned kernels do not take run the generic FP64 kernels in every
    precision: both steps generic gives the FP32 context's bits."""
    p = gpara(**CASES["B11 G32"])
    _, n = synth.grey_pair(64, 70, 25.0, 23)
    np.testing.assert_array_equal(BM3D(p, precision=prec)(n), BM3D(p)(n))


def test_generic_device_buffers_and_noalloc(gpu):
    import ctypes
    import torch
    p = gpara(**CASES["B5 G48"])
    _, n = synth.grey_pair(90, 100, 25.0, 29)
    f = BM3D(p)
    host = f(n)
    x = torch.from_numpy(n).to(gpu)
    dev = f(x).cpu().numpy()
    np.testing.assert_array_equal(dev, host)
    y = torch.empty_like(x)
    s = ctypes.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    rc = L.lib().isp_bm3d_plane_f32_noalloc(f._ctx, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                            None, 100, 90, 100, L.BOTH, s)
    assert rc == L.OK
    np.testing.assert_array_equal(y.cpu().numpy(), host)


def test_generic_limits(gpu):
    _, n = synth.grey_pair(40, 40, 25.0, 31)
    for kw, code in ((dict(B1=17, B2=17), L.E_UNSUPPORTED),       # BlockSize above 16
                     (dict(G1=65, G2=65), L.E_UNSUPPORTED),       # GroupSize above 64
                     (dict(B1=16, B2=16, G1=64, G2=64), L.E_UNSUPPORTED),  # 64 x 16^2 samples per group
                     (dict(B1=12, B2=12), None)):
        p = gpara(**kw)
        if code is None:
            assert np.all(np.isfinite(BM3D(p)(n)))
            continue
        with pytest.raises(L.BM3DError) as e:
            BM3D(p)(n)
        assert e.value.code == code
    # an image smaller than the block
    with pytest.raises(L.BM3DError) as e:
        BM3D(gpara(B1=11, B2=11))(np.zeros((10, 40), np.float32))
    assert e.value.code == L.E_SIZE
    # multi-frame groups and batches keep to the tuned kernels' parameters
    import torch
    from isp_mw_amd import BM3DBatch
    vn = BM3D_Para("vn").set_sigma(20.0)
    stack = np.stack([n, n])
    with pytest.raises(L.BM3DError) as e:
        BM3D(vn).process_stack(np.empty_like(stack), stack, radius=1)
    assert e.value.code == L.E_UNSUPPORTED
    x = torch.from_numpy(n).to(gpu)
    with pytest.raises(L.BM3DError) as e:
        BM3DBatch(BM3D(vn), [torch.empty_like(x)], [x]).run()
    assert e.value.code == L.E_UNSUPPORTED
