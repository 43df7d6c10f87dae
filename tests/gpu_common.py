"""Shared helpers of the GPU parity tests."""
import numpy as np

from isp_mw_amd import BM3D_Para, synth

MAX_ABS = 2e-3      # max |GPU - oracle| on [0, 1] data, plain FP32 path
MAX_DPSNR = 0.01    # |PSNR(GPU) - PSNR(oracle)| in dB, plain FP32 path
MIN_GROUPS = 0.999  # fraction of grid points with identical ordered groups


def para(profile="np", sigma=25.0, step=3, R=12, G1=None, G2=None, bmstep=1):
    p = BM3D_Para(profile)
    for st in (p.basic, p.final):
        if step is not None:
            st.BlockStep = step
        if R is not None:
            st.BMrange = R
        st.BMstep = bmstep
    if G1:
        p.basic.GroupSize = G1
    if G2:
        p.final.GroupSize = G2
    p.set_sigma(sigma)
    return p


def assert_close(gpu, ref, clean=None, max_abs=MAX_ABS, max_dpsnr=MAX_DPSNR, label=""):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    assert np.all(np.isfinite(gpu)), label
    err = float(np.abs(gpu - ref).max())
    print(f"{label}: max|gpu-oracle| = {err:.3g}", flush=True)
    assert err <= max_abs, (label, err)
    if clean is not None:
        d = abs(synth.psnr(gpu, clean) - synth.psnr(ref, clean))
        print(f"{label}: dPSNR = {d:.5f} dB", flush=True)
        assert d <= max_dpsnr, (label, d)


def group_identity(c_gpu, p_gpu, c_ref, p_ref):
    """Fraction of grid points whose ordered group (count and positions) agree."""
    same = c_gpu == c_ref
    G = min(p_gpu.shape[-1], p_ref.shape[-1])
    idx = np.arange(G)[None, None, :]
    mask = idx < c_ref[..., None]
    eqpos = np.all((p_gpu[..., :G] == p_ref[..., :G]) | ~mask, axis=-1)
    return float(np.mean(same & eqpos))
