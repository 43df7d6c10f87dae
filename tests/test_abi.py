"""The C-ABI library loads and exports every function include/isp_bm3d.h
declares (no GPU needed)."""
import ctypes
import os
import re

from isp_mw_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header="isp_bm3d.h", prefix="isp_bm3d_"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared() == sorted(L.EXPORTS)
    assert declared("isp_nlm.h", "isp_nlm_") == sorted(L.NLM_EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared() + declared("isp_nlm.h", "isp_nlm_") + declared("isp_pilot.h", "isp_pilot_")
    assert len(names) >= 40
    for name in names:
        assert hasattr(lib, name), name


def test_nlm_params_and_argument_checks():
    lib = L.lib()
    # 4 + pad + 2 doubles + 5 ints + pad + double
    assert ctypes.sizeof(L.NlmParams) == 8 + 16 + 20 + 4 + 8
    p = L.NlmParams()
    assert lib.isp_nlm_params_default(ctypes.byref(p), 1, 8.0) == L.OK
    # NLMeans_Para (include/NLMeans.h:11-22)
    assert (p.correction, p.sigma, p.strength, p.thMSE) == (1, 8.0, 40.0, 400.0)
    assert (p.block_size, p.block_step, p.group_size, p.bm_range, p.bm_step) == (8, 5, 16, 24, 3)
    assert lib.isp_nlm_params_default(ctypes.byref(p), 0, 10.0) == L.OK
    assert (p.correction, p.strength, p.thMSE) == (0, 15.0, 250.0)
    assert lib.isp_nlm_params_default(None, 1, 8.0) == L.E_ARG
    assert lib.isp_nlm_plane_f32(None, ctypes.byref(p), None, None, None, 8, 8, 8, L.MEM_HOST, None) == L.E_ARG


def test_null_and_bad_arguments():
    lib = L.lib()
    assert lib.isp_bm3d_plane_f32(None, None, None, None, 8, 8, 8, L.BOTH, L.MEM_HOST, None) == L.E_ARG
    assert lib.isp_bm3d_create(None, None, 0) == L.E_ARG
    assert lib.isp_bm3d_destroy(None) == L.OK
    assert lib.isp_bm3d_last_error(None)


def test_sizes_of_structs():
    # 8 (profile) + 24 (sigma) + 5 ints (20) + pad 4 + thMSE 8 + lambda 8
    assert ctypes.sizeof(L.StepParams) == 72 and ctypes.sizeof(L.Params) == 144


def test_batch_struct_sizes():
    # 9 plane pointers, width, height, stride, sigma
    assert ctypes.sizeof(L.Image) == 9 * 8 + 4 + 4 + 8 + 8
    # 5 ints, pad, 2 doubles
    assert ctypes.sizeof(L.Tuning) == 5 * 4 + 4 + 16


def test_batch_and_tuning_argument_checks():
    lib = L.lib()
    plan = ctypes.c_void_p()
    assert lib.isp_bm3d_batch_create(None, ctypes.byref(plan), None, 1, 1, L.BOTH, 0) == L.E_ARG
    assert lib.isp_bm3d_batch_run(None, None, 1) == L.E_ARG
    assert lib.isp_bm3d_batch_destroy(None) == L.OK
    assert lib.isp_bm3d_autotune(None, 64, 64, 25.0, 0, None) == L.E_ARG
    assert lib.isp_bm3d_set_tuning(None, 16, None) == L.E_ARG
    # size class = ceil(log2(width * height))
    assert lib.isp_bm3d_size_class(256, 256) == 16
    assert lib.isp_bm3d_size_class(257, 256) == 17
    assert lib.isp_bm3d_size_class(1920, 1080) == 21
    assert lib.isp_bm3d_size_class(0, 5) == L.E_ARG


def test_plan_on_the_host():
    """isp_bm3d_plan: grid sizes, generic steps and the no-alloc scratch without a GPU."""
    lib = L.lib()
    p = L.Params()
    info = L.PlanInfo()
    assert lib.isp_bm3d_params_default(ctypes.byref(p), b"np") == L.OK
    for st in (p.basic, p.final_):
        st.block_step, st.bm_range = 3, 12
    p.basic.group_size, p.final_.group_size = 16, 32
    assert lib.isp_bm3d_plan(ctypes.byref(p), 1920, 1080, 3, ctypes.byref(info)) == L.OK
    assert (info.ngx[0], info.ngy[0], info.ngx[1], info.ngy[1]) == (639, 359, 639, 359)
    assert list(info.generic) == [0, 0] and list(info.degenerate) == [0, 0]
    assert info.weight_cap[1] == 32768.0
    assert info.noalloc_bytes <= info.noalloc_capacity == 96 << 20  # a 1080p plane fits the scratch
    assert lib.isp_bm3d_plan(ctypes.byref(p), 3840, 2160, 1, ctypes.byref(info)) == L.OK
    assert info.noalloc_bytes > info.noalloc_capacity
    assert lib.isp_bm3d_params_default(ctypes.byref(p), b"vn") == L.OK
    assert lib.isp_bm3d_plan(ctypes.byref(p), 640, 480, 1, ctypes.byref(info)) == L.OK
    assert list(info.generic) == [0, 1]          # final BlockSize 11: the generic kernels
    assert info.ngx[1] == (640 - 11) // 6 + 1 + (1 if (640 - 11) % 6 else 0)
    p.final_.block_size = 17
    assert lib.isp_bm3d_plan(ctypes.byref(p), 640, 480, 1, ctypes.byref(info)) == L.E_UNSUPPORTED
    assert b"BlockSize 17" in lib.isp_bm3d_last_error(None)
    p.final_.block_size = 11
    assert lib.isp_bm3d_plan(ctypes.byref(p), 16, 40000, 1, ctypes.byref(info)) == L.E_UNSUPPORTED
    assert lib.isp_bm3d_plan(ctypes.byref(p), 10, 10, 1, ctypes.byref(info)) == L.E_SIZE
    p.basic.sigma[0] = p.final_.sigma[0] = 0.0
    assert lib.isp_bm3d_plan(ctypes.byref(p), 10, 10, 1, ctypes.byref(info)) == L.OK
    assert list(info.degenerate) == [1, 1]       # sigma 0: both steps copy through
