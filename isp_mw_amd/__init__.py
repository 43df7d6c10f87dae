"""MI355X-native BM3D denoiser with ISP_MW's BM3D interface.

The compute path is libisp_bm3d.so: hand-written HIP kernels for gfx950 behind
a C ABI (include/isp_bm3d.h).  This package mirrors the reference's C++ classes
(include/BM3D.h) in Python, with the NL-means filter on the same matching
kernel (include/NLMeans.h), and generates the synthetic benchmark images.
"""
from .batch import BM3DBatch  # noqa: F401
from .bm3d import (BM3D, BM3D_Basic, BM3D_Basic_Para, BM3D_Final, BM3D_Final_Para,  # noqa: F401
                   BM3D_Para, BM3D_Para_Base, c_params_default)
from .nlm import NLMeans, NLMeans_Para  # noqa: F401

__all__ = ["BM3DBatch", "BM3D", "BM3D_Basic", "BM3D_Final", "BM3D_Para", "BM3D_Basic_Para", "BM3D_Final_Para",
           "BM3D_Para_Base", "c_params_default", "NLMeans", "NLMeans_Para"]
