"""Host-side harness of libisp_bm3d.so's C ABI, for the sanitizer test
(tests/test_sanitizers.py runs it under AddressSanitizer + UBSan against the
library built with instrumented host code).  It needs no GPU: parameter
presets, thMSE defaults, the aggregation bound, the host-only plan (parameter
derivation, grid, workspace layout) over many parameter sets and sizes,
argument checks of every entry point, and context creation's error path.
Exits non-zero on a wrong result; the sanitizers abort on a memory or
undefined-behaviour error."""
import ctypes
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from isp_mw_amd import _lib as L  # noqa: E402


def main() -> int:
    lib = L.lib()
    assert L.LIB_PATH.endswith(os.environ.get("HARNESS_EXPECT_LIB", ".so")), L.LIB_PATH
    p = L.Params()
    for prof in (b"fast", b"lc", b"np", b"high", b"vn", None):
        assert lib.isp_bm3d_params_default(ctypes.byref(p), prof) == L.OK
    assert lib.isp_bm3d_params_default(ctypes.byref(p), b"bogus") == L.E_PROFILE
    assert lib.isp_bm3d_params_default(ctypes.byref(p), b"x" * 4096) == L.E_PROFILE
    assert lib.isp_bm3d_params_default(None, b"np") == L.E_ARG
    sp = L.StepParams()
    for fin in (0, 1):
        assert lib.isp_bm3d_step_params_default(ctypes.byref(sp), b"vn", fin) == L.OK
        assert lib.isp_bm3d_thmse_default(ctypes.byref(sp), fin) == L.OK
    assert lib.isp_bm3d_size_class(1920, 1080) == 21
    b, c = ctypes.c_double(), ctypes.c_double()
    for s in (1, 2, 3, 5, 8):
        for R in (0, 4, 12, 19, 64, 300):
            for G in (1, 16, 32, 64):
                sp.block_size, sp.block_step, sp.group_size, sp.bm_range, sp.bm_step = 8, s, G, R, 1
                lib.isp_bm3d_aggregation_limits(ctypes.byref(sp), 1, ctypes.byref(b), ctypes.byref(c))
    # the plan over parameter sets and sizes, valid and not
    info = L.PlanInfo()
    cases = 0
    for prof in (b"fast", b"np", b"high", b"vn"):
        for (w, h) in ((8, 8), (9, 13), (64, 48), (1920, 1080), (4096, 2160), (65535, 16), (16, 32767),
                       (16, 40000), (70000, 16), (7, 100), (0, 5)):
            for nplanes in (1, 3):
                lib.isp_bm3d_params_default(ctypes.byref(p), prof)
                rc = lib.isp_bm3d_plan(ctypes.byref(p), w, h, nplanes, ctypes.byref(info))
                cases += 1
                if w < 8 or h < 8:
                    assert rc in (L.E_SIZE, L.OK), (prof, w, h, rc)  # vn final BlockSize 11 > 8 too
                elif h > 32767 or w > 65535:
                    assert rc == L.E_UNSUPPORTED, (prof, w, h, rc)
                elif rc == L.OK:
                    assert info.noalloc_capacity == 96 << 20
                    assert info.workspace_bytes > 0 and info.noalloc_bytes > 0
                    assert info.generic[1] == (1 if prof == b"vn" else 0), (prof, info.generic[1])
    for B, G in ((1, 1), (3, 40), (11, 32), (16, 32), (16, 64), (17, 8), (8, 65), (0, 8), (8, 0)):
        lib.isp_bm3d_params_default(ctypes.byref(p), b"np")
        p.final_.block_size, p.final_.group_size = B, G
        rc = lib.isp_bm3d_plan(ctypes.byref(p), 300, 200, 3, ctypes.byref(info))
        cases += 1
        assert rc in (L.OK, L.E_UNSUPPORTED), (B, G, rc)
        assert (rc == L.OK) == (B <= 16 and G <= 64 and B * B * G <= 8192 or B <= 0 or G == 0), (B, G, rc)
    assert lib.isp_bm3d_plan(None, 8, 8, 1, ctypes.byref(info)) == L.E_ARG
    assert lib.isp_bm3d_plan(ctypes.byref(p), 8, 8, 2, ctypes.byref(info)) == L.E_ARG
    # every entry point rejects a null context or null buffers without touching memory
    vp = None
    assert lib.isp_bm3d_plane_f32(vp, vp, vp, vp, 64, 64, 64, L.BOTH, L.MEM_HOST, vp) == L.E_ARG
    assert lib.isp_bm3d_plane_f32_noalloc(vp, vp, vp, vp, 64, 64, 64, L.BOTH, vp) == L.E_ARG
    assert lib.isp_bm3d_plane_i32(vp, vp, vp, vp, 64, 64, 64, 0, 0, 255, L.BOTH, L.MEM_HOST, vp) == L.E_ARG
    assert lib.isp_bm3d_rgb_f32(vp, *([vp] * 9), 64, 64, 64, L.BOTH, L.MEM_HOST, vp) == L.E_ARG
    assert lib.isp_bm3d_rgb_i32(vp, *([vp] * 9), 64, 64, 64, 0, 255, L.BOTH, L.MEM_HOST, vp) == L.E_ARG
    assert lib.isp_bm3d_stack_f32(vp, vp, vp, vp, 2, 1, 64, 64, 64, 1, L.BOTH, L.MEM_HOST, vp) == L.E_ARG
    assert lib.isp_bm3d_batch_create(vp, vp, vp, 1, 1, L.BOTH, 0) == L.E_ARG
    assert lib.isp_bm3d_set_precision(vp, 1) == L.E_ARG
    assert lib.isp_bm3d_last_error(vp)
    # context creation without a device: an error status and message, nothing leaked or overrun
    ctx = ctypes.c_void_p()
    lib.isp_bm3d_params_default(ctypes.byref(p), b"vn")
    rc = lib.isp_bm3d_create(ctypes.byref(ctx), ctypes.byref(p), 0)
    assert rc in (L.OK, L.E_HIP), rc
    if rc == L.OK:
        lib.isp_bm3d_destroy(ctx)
    np_ = L.NlmParams()
    assert lib.isp_nlm_params_default(ctypes.byref(np_), 1, 8.0) == L.OK
    print(f"host harness: {cases} plan cases, library {os.path.basename(L.LIB_PATH)}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
