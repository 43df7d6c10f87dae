/*
 * isp_nlm.h - C ABI of the MI355X NL-means denoiser (in libisp_bm3d.so).
 *
 * Drop-in replacement for the NLMeans filter of ISP_MW (include/NLMeans.h,
 * source/NLMeans.cpp): block matching as BM3D's, then a weighted average of
 * each group (weights exp(-MSE / strength^2)), optionally corrected by a
 * stationarity test, aggregated uniformly.  The filter runs on an
 * isp_bm3d_ctx (include/isp_bm3d.h), which supplies the device, workspace,
 * stream timing and the matching precision: ISP_BM3D_PREC_FP32 (the default)
 * forms the reference's groups exactly (FP64 distances); FP16 / BF16
 * contexts match with FP32 distances.  Averaging is FP64 in every case.
 *
 * Units follow the reference: sigma, strength and thMSE on the 8-bit scale;
 * float planes with luma in [0,1] (|samples| < 2^20).  Status codes are
 * those of isp_bm3d.h.
 */
#ifndef ISP_NLM_H_
#define ISP_NLM_H_

#include "isp_bm3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of NLMeans_Para (include/NLMeans.h:11-22). */
typedef struct isp_nlm_params {
    int correction;   /* 1: WeightedAverage_Correction, 0: WeightedAverage       */
    double sigma;     /* noise sigma, 8-bit scale                                */
    double strength;  /* weights exp(-key / strength^2)                          */
    int block_size;   /* BlockSize (the GPU kernels implement 8)                 */
    int block_step;   /* BlockStep                                               */
    int group_size;   /* GroupSize (2..32 on the GPU; 1 copies src through)      */
    int bm_range;     /* BMrange                                                 */
    int bm_step;      /* BMstep                                                  */
    double thMSE;     /* matching threshold (not scaled by the OPP norm)         */
} isp_nlm_params;

/* NLMeans_Para with the given correction and sigma and the defaults of
 * include/NLMeans.h:11-22 (also NLMeans_IO, :148-149): strength = 5 sigma
 * (1.5 sigma without correction), thMSE = 50 sigma (25 sigma), BlockSize 8,
 * BlockStep 5, GroupSize 16, BMrange 24, BMstep 3. */
int isp_nlm_params_default(isp_nlm_params *p, int correction, double sigma);

/* Plane_FL: NLMeans::process_Plane_FL (source/NLMeans.cpp:150-231).  ref may
 * be NULL (= src); it is the matching guide.  stride in elements for dst, src
 * and ref; mem and stream as in isp_bm3d_plane_f32. */
int isp_nlm_plane_f32(isp_bm3d_ctx *ctx, const isp_nlm_params *p, float *dst, const float *src, const float *ref,
                      int width, int height, size_t stride, int mem, void *stream);

/* Plane (int32 with Floor / Neutral / Ceil): NLMeans::process_Plane
 * (source/NLMeans.cpp:232-257): RangeConvert to [0,1] with clip, the float
 * filter, RangeConvert back with +0.5 rounding. */
int isp_nlm_plane_i32(isp_bm3d_ctx *ctx, const isp_nlm_params *p, int *dst, const int *src, const int *ref,
                      int width, int height, size_t stride, int floor_, int neutral, int ceil_, int mem,
                      void *stream);

/* Frame (int32 R, G, B with Floor / Ceil): NLMeans::process_Frame
 * (source/NLMeans.cpp:260-375): groups matched on the OPP luma of ref
 * (ConvertToY with clip, include/Conversion.hpp:398-432) and applied to each
 * of R, G, B; output (int)(num / den + 0.5).  ref_* may all be NULL (= src).
 * Samples must lie within 32x the larger of |floor_| and |ceil_|. */
int isp_nlm_rgb_i32(isp_bm3d_ctx *ctx, const isp_nlm_params *p, int *dst_r, int *dst_g, int *dst_b,
                    const int *src_r, const int *src_g, const int *src_b, const int *ref_r, const int *ref_g,
                    const int *ref_b, int width, int height, size_t stride, int floor_, int ceil_, int mem,
                    void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ISP_NLM_H_ */
