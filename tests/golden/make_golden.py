"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference ships no tests, golden data or images, and it cannot be built
here (FFTW3 and MSVC-only code), so these vectors come from the oracle, whose
own correctness is pinned independently in tests/test_oracle.py.  They freeze
its outputs so that later changes to it, and the GPU path, are checked
against fixed data.  Inputs are regenerated from isp_mw_amd.synth seeds and
stored too.  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from isp_mw_amd import BM3D_Para, synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [
    # name, kind, shape, sigma, seed, profile, overrides (BlockStep, BMrange)
    ("grey64_np_s3r12", "grey", (64, 64), 25.0, 101, "np", (3, 12)),
    ("grey96x80_np", "grey", (80, 96), 25.0, 102, "np", None),
    ("grey72_fast", "grey", (72, 72), 15.0, 103, "fast", None),
    ("rgb48x40_np_s3r12", "rgb", (40, 48), 25.0, 104, "np", (3, 12)),
    ("rgb56x64_lc", "rgb", (64, 56), 10.0, 105, "lc", None),
]


def para_for(profile, sigma, ov):
    p = BM3D_Para(profile)
    if ov:
        for st in (p.basic, p.final):
            st.BlockStep, st.BMrange = ov
    p.set_sigma(sigma)
    return p


def main():
    oracle.set_threads(1)
    for name, kind, (h, w), sigma, seed, profile, ov in CASES:
        p = para_for(profile, sigma, ov)
        if kind == "grey":
            clean, noisy = synth.grey_pair(h, w, sigma, seed)
            basic = oracle.kernel(p.basic, noisy, noisy, wiener=False)
            out = oracle.bm3d_plane(p.basic, p.final, noisy)
            c1, p1, _ = oracle.match_all(p.basic, noisy)
            c2, p2, _ = oracle.match_all(p.final, basic)
        else:
            clean, noisy = synth.rgb_pair(h, w, sigma, seed)
            out = oracle.bm3d_rgb(p.basic, p.final, noisy)
            yuv = oracle.rgb2opp(noisy)
            basic = np.stack(oracle.kernel(p.basic, list(yuv), list(yuv), wiener=False))
            c1, p1, _ = oracle.match_all(p.basic, yuv[0])
            c2, p2, _ = oracle.match_all(p.final, basic[0])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), noisy=noisy, basic=basic, out=out,
                            counts1=c1, pos1=p1, counts2=c2, pos2=p2,
                            meta=np.array([h, w, seed], np.int64), sigma=np.float64(sigma),
                            profile=np.array(profile), overrides=np.array(ov if ov else (0, 0), np.int64))
        print(name, out.shape, f"psnr noisy {synth.psnr(noisy, clean):.2f} out {synth.psnr(out, clean):.2f}")


if __name__ == "__main__":
    main()
