"""ctypes binding of libisp_bm3d.so (include/isp_bm3d.h, include/isp_nlm.h).

The library is required: importing the binding fails loudly when it has not
been built (``python -m isp_mw_amd.build``); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libisp_bm3d.so")
# ISP_BM3D_LIB names another build of the same library (the host-sanitized
# libisp_bm3d_san.so of the sanitizer tests); it must exist like the default
LIB_PATH = os.environ.get("ISP_BM3D_LIB") or LIB_PATH

OK = 0
E_ARG, E_UNSUPPORTED, E_SIZE, E_HIP, E_PROFILE = -1, -2, -3, -4, -5
BASIC, FINAL, BOTH = 1, 2, 3
MEM_HOST, MEM_DEVICE = 0, 1
PREC_FP32, PREC_FP16, PREC_BF16, PREC_FP32_FAST, PREC_FP32_MFMA = 0, 1, 2, 3, 4
VARIANT_HAAR, VARIANT_KAISER = 1, 2

# Every function include/isp_bm3d.h declares.
EXPORTS = (
    "isp_bm3d_params_default",
    "isp_bm3d_step_params_default",
    "isp_bm3d_thmse_default",
    "isp_bm3d_create",
    "isp_bm3d_destroy",
    "isp_bm3d_last_error",
    "isp_bm3d_set_precision",
    "isp_bm3d_set_variant",
    "isp_bm3d_plane_f32",
    "isp_bm3d_plane_f32_noalloc",
    "isp_bm3d_rgb_f32",
    "isp_bm3d_plane_i32",
    "isp_bm3d_rgb_i32",
    "isp_bm3d_stack_f32",
    "isp_bm3d_match",
    "isp_bm3d_grid_size",
    "isp_bm3d_aggregation_limits",
    "isp_bm3d_plan",
    "isp_bm3d_set_timing",
    "isp_bm3d_get_timing",
    "isp_bm3d_reset_timing",
    "isp_bm3d_last_groups",
    "isp_bm3d_rgb_to_opp",
    "isp_bm3d_opp_to_rgb",
    "isp_bm3d_batch_create",
    "isp_bm3d_batch_run",
    "isp_bm3d_batch_update",
    "isp_bm3d_batch_info",
    "isp_bm3d_batch_destroy",
    "isp_bm3d_size_class",
    "isp_bm3d_autotune",
    "isp_bm3d_set_tuning",
    "isp_bm3d_get_tuning",
)
# Every function include/isp_nlm.h declares.
NLM_EXPORTS = (
    "isp_nlm_params_default",
    "isp_nlm_plane_f32",
    "isp_nlm_plane_i32",
    "isp_nlm_rgb_i32",
)
BATCH_AUTO_GROUP = 1
TUNE_RGB = 2
NUM_STAGES = 10


class StepParams(ctypes.Structure):
    """isp_bm3d_step_params (mirror of BM3D_Para_Base, include/BM3D.h:15-45)."""

    _fields_ = [
        ("profile", ctypes.c_char * 8),
        ("sigma", ctypes.c_double * 3),
        ("block_size", ctypes.c_int),
        ("block_step", ctypes.c_int),
        ("group_size", ctypes.c_int),
        ("bm_range", ctypes.c_int),
        ("bm_step", ctypes.c_int),
        ("thMSE", ctypes.c_double),
        ("lambda_", ctypes.c_double),
    ]


class PlanInfo(ctypes.Structure):
    """isp_bm3d_plan_info."""

    _fields_ = [
        ("degenerate", ctypes.c_int * 2),
        ("generic", ctypes.c_int * 2),
        ("ngx", ctypes.c_int * 2),
        ("ngy", ctypes.c_int * 2),
        ("weight_cap", ctypes.c_double * 2),
        ("workspace_bytes", ctypes.c_size_t),
        ("noalloc_bytes", ctypes.c_size_t),
        ("noalloc_capacity", ctypes.c_size_t),
    ]


class NlmParams(ctypes.Structure):
    """isp_nlm_params (mirror of NLMeans_Para, include/NLMeans.h:11-22)."""

    _fields_ = [
        ("correction", ctypes.c_int),
        ("sigma", ctypes.c_double),
        ("strength", ctypes.c_double),
        ("block_size", ctypes.c_int),
        ("block_step", ctypes.c_int),
        ("group_size", ctypes.c_int),
        ("bm_range", ctypes.c_int),
        ("bm_step", ctypes.c_int),
        ("thMSE", ctypes.c_double),
    ]


class Params(ctypes.Structure):
    """isp_bm3d_params (mirror of BM3D_Para, include/BM3D.h:154-168)."""

    _fields_ = [("basic", StepParams), ("final_", StepParams)]


class Image(ctypes.Structure):
    """isp_bm3d_image: one image of a batch (device planes)."""

    _fields_ = [
        ("dst", ctypes.c_void_p * 3),
        ("src", ctypes.c_void_p * 3),
        ("ref", ctypes.c_void_p * 3),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("stride", ctypes.c_size_t),
        ("sigma", ctypes.c_double),
    ]


class Tuning(ctypes.Structure):
    """isp_bm3d_tuning: tile shapes (and final GroupSize) of one size class."""

    _fields_ = [
        ("match_tile", ctypes.c_int * 2),
        ("filter_tile", ctypes.c_int * 2),
        ("final_group_size", ctypes.c_int),
        ("ms_per_image", ctypes.c_double),
        ("psnr_db", ctypes.c_double),
    ]


class BM3DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"isp_bm3d error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m isp_mw_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    # One HIP runtime per process: when torch is present its libamdhip64.so.7
    # is loaded first, and the library's NEEDED entry (same soname) binds to
    # it, so torch tensors and streams can be passed straight through.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    sp = ctypes.POINTER(StepParams)
    pp = ctypes.POINTER(Params)
    L.isp_bm3d_params_default.argtypes = [pp, ctypes.c_char_p]
    L.isp_bm3d_step_params_default.argtypes = [sp, ctypes.c_char_p, i]
    L.isp_bm3d_thmse_default.argtypes = [sp, i]
    L.isp_bm3d_create.argtypes = [ctypes.POINTER(vp), pp, i]
    L.isp_bm3d_destroy.argtypes = [vp]
    L.isp_bm3d_last_error.argtypes = [vp]
    L.isp_bm3d_last_error.restype = ctypes.c_char_p
    L.isp_bm3d_set_precision.argtypes = [vp, i]
    L.isp_bm3d_set_variant.argtypes = [vp, i, ctypes.c_double]
    L.isp_bm3d_plane_f32.argtypes = [vp, vp, vp, vp, i, i, sz, i, i, vp]
    L.isp_bm3d_plane_f32_noalloc.argtypes = [vp, vp, vp, vp, i, i, sz, i, vp]
    L.isp_bm3d_rgb_f32.argtypes = [vp] + [vp] * 9 + [i, i, sz, i, i, vp]
    L.isp_bm3d_plane_i32.argtypes = [vp, vp, vp, vp, i, i, sz, i, i, i, i, i, vp]
    L.isp_bm3d_rgb_i32.argtypes = [vp] + [vp] * 9 + [i, i, sz, i, i, i, i, vp]
    L.isp_bm3d_stack_f32.argtypes = [vp, vp, vp, vp, i, i, i, i, sz, i, i, i, vp]
    L.isp_bm3d_match.argtypes = [vp, vp, i, i, sz, i, i, vp, vp, vp]
    L.isp_bm3d_grid_size.argtypes = [vp, i, i, i, ctypes.POINTER(i), ctypes.POINTER(i)]
    L.isp_bm3d_aggregation_limits.argtypes = [sp, i, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double)]
    L.isp_bm3d_plan.argtypes = [pp, i, i, i, ctypes.POINTER(PlanInfo)]
    L.isp_bm3d_set_timing.argtypes = [vp, i]
    L.isp_bm3d_get_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_long), i]
    L.isp_bm3d_reset_timing.argtypes = [vp]
    L.isp_bm3d_last_groups.argtypes = [vp, i, vp, vp, sz]
    L.isp_bm3d_rgb_to_opp.argtypes = [vp] * 6 + [i, i, sz, vp]
    L.isp_bm3d_opp_to_rgb.argtypes = [vp] * 6 + [i, i, sz, i, vp]
    L.isp_bm3d_batch_create.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(Image), i, i, i, i]
    L.isp_bm3d_batch_run.argtypes = [vp, vp, i]
    L.isp_bm3d_batch_update.argtypes = [vp, ctypes.POINTER(Image)]
    L.isp_bm3d_batch_info.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(sz), ctypes.POINTER(i)]
    L.isp_bm3d_batch_destroy.argtypes = [vp]
    L.isp_bm3d_size_class.argtypes = [i, i]
    L.isp_bm3d_autotune.argtypes = [vp, i, i, ctypes.c_double, i, ctypes.POINTER(Tuning)]
    L.isp_bm3d_set_tuning.argtypes = [vp, i, ctypes.POINTER(Tuning)]
    L.isp_bm3d_get_tuning.argtypes = [vp, i, ctypes.POINTER(Tuning)]
    npp = ctypes.POINTER(NlmParams)
    L.isp_nlm_params_default.argtypes = [npp, i, ctypes.c_double]
    L.isp_nlm_plane_f32.argtypes = [vp, npp, vp, vp, vp, i, i, sz, i, vp]
    L.isp_nlm_plane_i32.argtypes = [vp, npp, vp, vp, vp, i, i, sz, i, i, i, i, vp]
    L.isp_nlm_rgb_i32.argtypes = [vp, npp] + [vp] * 9 + [i, i, sz, i, i, i, vp]
    for name in EXPORTS + NLM_EXPORTS:
        getattr(L, name).restype = getattr(L, name).restype or ctypes.c_int
    L.isp_bm3d_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc: int, ctx=None) -> None:
    if rc != OK:
        msg = lib().isp_bm3d_last_error(ctx)
        raise BM3DError(rc, msg.decode() if msg else "")
