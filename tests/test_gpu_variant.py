"""The configs' variant modes (isp_bm3d_set_variant): the Haar transform along
the group axis with groups truncated to a power of two, and Kaiser-windowed
aggregation (beta 2), against the oracle in the same modes (oracle.set_variant,
pinned by an independent numpy formulation in tests/test_oracle.py)."""
import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Basic, BM3D_Final, BM3DBatch, synth
from isp_mw_amd import _lib as L
from tests.gpu_common import assert_close, para

pytestmark = pytest.mark.gpu
VARIANTS = {"haar": (True, 0.0), "kaiser": (False, 2.0), "haar+kaiser": (True, 2.0)}


@pytest.fixture
def ovariant():
    def set_(haar, beta):
        oracle.set_variant(haar, beta)
    yield set_
    oracle.set_variant(False, 0.0)


@pytest.mark.parametrize("mode", ["basic", "final", "both"])
@pytest.mark.parametrize("name", list(VARIANTS))
def test_variant_plane_vs_oracle(gpu, ovariant, name, mode):
    haar, beta = VARIANTS[name]
    print(f"variant {name} {mode}", flush=True)
    clean, n = synth.grey_pair(64, 72, 25.0, 41)
    p = para(G1=16, G2=32)
    ovariant(haar, beta)
    if mode == "basic":
        f = BM3D_Basic(p.basic)
        f.set_variant(haar, beta)
        got = f(n)
        ref = oracle.kernel(p.basic, n, n, wiener=False)
    elif mode == "final":
        guide = oracle.kernel(p.basic, n, n, wiener=False).astype(np.float32)
        f = BM3D_Final(p.final)
        f.set_variant(haar, beta)
        got = f.process(np.empty_like(n), n, guide)
        ref = oracle.kernel(p.final, n, guide, wiener=True)
    else:
        f = BM3D(p)
        f.set_variant(haar, beta)
        got = f(n)
        ref = oracle.bm3d_plane(p.basic, p.final, n)
    assert_close(got, ref, clean, max_abs=1e-6, label=f"variant {name} {mode}")


def test_variant_rgb_and_precisions(gpu, ovariant):
    clean, n = synth.rgb_pair(48, 64, 25.0, 43)
    p = para(G1=16, G2=32)
    ovariant(True, 2.0)
    ref = oracle.bm3d_rgb(p.basic, p.final, n)
    outs = []
    for prec in ("fp32", "fp16", "bf16"):
        f = BM3D(p, precision=prec)
        f.set_variant(True, 2.0)
        outs.append(f(n))
        assert_close(outs[-1], ref, clean, max_abs=1e-6, label=f"variant rgb {prec}")
    np.testing.assert_array_equal(outs[1], outs[0])   # every precision runs the same FP64 kernels


def test_variant_off_restores_the_reference(gpu):
    _, n = synth.grey_pair(56, 60, 25.0, 45)
    p = para()
    want = BM3D(p)(n)
    f = BM3D(p)
    f.set_variant(True, 2.0)
    other = f(n)
    f.set_variant()
    np.testing.assert_array_equal(f(n), want)
    assert float(np.abs(other - want).max()) > 1e-4
    with pytest.raises(L.BM3DError) as e:
        L.check(L.lib().isp_bm3d_set_variant(f._ctx, 8, 0.0), f._ctx)
    assert e.value.code == L.E_ARG


def test_variant_refused_by_batches_and_stacks(gpu):
    import torch
    _, n = synth.grey_pair(40, 40, 25.0, 47)
    f = BM3D(para())
    f.set_variant(True, 0.0)
    x = torch.from_numpy(n).cuda()
    with pytest.raises(L.BM3DError) as e:
        BM3DBatch(f, [torch.empty_like(x)], [x]).run()
    assert e.value.code == L.E_UNSUPPORTED
    stack = np.stack([n, n])
    with pytest.raises(L.BM3DError) as e:
        f.process_stack(np.empty_like(stack), stack, radius=1)
    assert e.value.code == L.E_UNSUPPORTED


def test_variant_quality_512(gpu):
    """The 512x512 grey case with the variant transform and window (Haar, Kaiser 2),
    beside the reference's semantics: both are reported, the variant must
    denoise about as well."""
    clean, n = synth.grey_pair(512, 512, 25.0, 5)
    # on one plane the reference scales the thresholds by normY = 1/sqrt(3)
    # (include/BM3D.h:246-258): sigma * sqrt(3) matches noise of sigma 25
    p = para(sigma=25.0 * np.sqrt(3.0))
    ref_mode = BM3D(p)(n)
    f = BM3D(p)
    f.set_variant(True, 2.0)
    var = f(n)
    pr, pv = synth.psnr(ref_mode, clean), synth.psnr(var, clean)
    print(f"512^2 grey: PSNR noisy {synth.psnr(n, clean):.3f}, reference semantics {pr:.3f}, "
          f"Haar + Kaiser 2 {pv:.3f} ({pv - pr:+.3f} dB)", flush=True)
    assert pv > synth.psnr(n, clean) + 6.0
    assert abs(pv - pr) < 1.0
