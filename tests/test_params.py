"""Parameter presets (include/BM3D.h:15-168) through the Python mirror and the
C ABI.  The table below is copied as numbers from the reference's lines."""
import ctypes
import re

import pytest

from isp_mw_amd import BM3D_Basic_Para, BM3D_Final_Para, BM3D_Para, c_params_default
from isp_mw_amd import _lib as L

# profile: (basic BlockSize, BlockStep, GroupSize, BMrange, lambda), (final BlockSize, BlockStep, GroupSize, BMrange)
TABLE = {
    "fast": ((8, 8, 16, 9, 2.7), (8, 7, 16, 9)),
    "lc": ((8, 6, 16, 9, 2.7), (8, 5, 16, 9)),
    "np": ((8, 4, 16, 16, 2.7), (8, 3, 32, 16)),
    "high": ((8, 3, 16, 16, 2.7), (8, 2, 32, 16)),
    "vn": ((8, 4, 32, 16, 2.8), (11, 6, 32, 16)),
}


@pytest.mark.parametrize("profile", list(TABLE))
def test_profile_table(profile):
    (bb, bs, bg, br, lam), (fb, fs, fg, fr) = TABLE[profile]
    p = BM3D_Para(profile)
    assert (p.basic.BlockSize, p.basic.BlockStep, p.basic.GroupSize, p.basic.BMrange, p.basic.lambda_) == \
        (bb, bs, bg, br, lam)
    assert (p.final.BlockSize, p.final.BlockStep, p.final.GroupSize, p.final.BMrange) == (fb, fs, fg, fr)
    assert p.basic.BMstep == p.final.BMstep == 1
    assert p.basic.sigma == [10.0, 10.0, 10.0]
    # thMSE_Default with sigma 10 (include/BM3D.h:89-97, 142-150)
    assert p.basic.thMSE == (1000 + 1500 if profile == "vn" else 400 + 800)
    assert p.final.thMSE == (400 + 400 if profile == "vn" else 200 + 100)


@pytest.mark.parametrize("profile", list(TABLE))
def test_c_abi_defaults_match_mirror(profile):
    c = c_params_default(profile)
    p = BM3D_Para(profile)
    for cs, ps in ((c.basic, p.basic), (c.final_, p.final)):
        assert cs.profile.decode() == profile
        assert (cs.block_size, cs.block_step, cs.group_size, cs.bm_range, cs.bm_step) == \
            (ps.BlockSize, ps.BlockStep, ps.GroupSize, ps.BMrange, ps.BMstep)
        assert cs.thMSE == ps.thMSE and list(cs.sigma) == ps.sigma
    assert c.basic.lambda_ == p.basic.lambda_


def test_default_profile_is_fast():
    p = L.Params()
    assert L.lib().isp_bm3d_params_default(ctypes.byref(p), None) == 0
    assert p.basic.profile == b"fast" and p.basic.block_step == 8


def test_unknown_profile():
    p = L.Params()
    assert L.lib().isp_bm3d_params_default(ctypes.byref(p), b"bogus") == L.E_PROFILE
    assert b"bogus" in L.lib().isp_bm3d_last_error(None)
    with pytest.raises(ValueError):
        BM3D_Basic_Para("bogus")
    with pytest.raises(ValueError):
        BM3D_Final_Para("bogus")


def test_set_sigma_replicates_and_recomputes():
    p = BM3D_Para("np").set_sigma(25.0)
    assert p.basic.sigma == [25.0] * 3 and p.final.sigma == [25.0] * 3
    assert p.basic.thMSE == 400 + 25 * 80 and p.final.thMSE == 200 + 25 * 10
    p.set_sigma([5.0, 7.0])
    assert p.basic.sigma == [5.0, 7.0, 7.0]


def test_thmse_default_c():
    s = L.StepParams()
    assert L.lib().isp_bm3d_step_params_default(ctypes.byref(s), b"vn", 0) == 0
    s.sigma[0] = 20.0
    assert L.lib().isp_bm3d_thmse_default(ctypes.byref(s), 0) == 0 and s.thMSE == 1000 + 20 * 150
    assert L.lib().isp_bm3d_thmse_default(ctypes.byref(s), 1) == 0 and s.thMSE == 400 + 20 * 40
