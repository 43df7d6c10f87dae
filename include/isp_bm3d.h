/*
 * isp_bm3d.h - C ABI of the MI355X BM3D denoiser (libisp_bm3d.so).
 *
 * Drop-in replacement for the BM3D filter of ISP_MW: the Basic / Final /
 * two-step filters of include/BM3D.h on Plane_FL-like float planes and on
 * float RGB frames.  No C++ or device-library types cross this boundary:
 * plain pointers, sizes and int status codes.  include/isp_bm3d.hpp layers
 * the reference's C++ class names (BM3D_Para, BM3D_Basic, BM3D_Final, BM3D,
 * Plane_FL, Frame) on top of these functions.
 *
 * Units follow the reference: sigma in 0..255, thMSE in MSE units of the
 * 8-bit scale, float samples with luma range [0,1] and chroma [-0.5,0.5].
 */
#ifndef ISP_BM3D_H_
#define ISP_BM3D_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Functions return ISP_BM3D_OK (0) or a negative code and
 * record a message readable with isp_bm3d_last_error(). */
#define ISP_BM3D_OK 0
#define ISP_BM3D_E_ARG -1         /* invalid argument (NULL pointer, bad mode, ...) */
#define ISP_BM3D_E_UNSUPPORTED -2 /* parameters outside what the GPU kernels implement */
#define ISP_BM3D_E_SIZE -3        /* image smaller than one block */
#define ISP_BM3D_E_HIP -4         /* HIP runtime error */
#define ISP_BM3D_E_PROFILE -5     /* unknown profile name */

/* Steps (the `mode` argument). */
#define ISP_BM3D_BASIC 1 /* BM3D_Basic: hard-threshold step only        */
#define ISP_BM3D_FINAL 2 /* BM3D_Final: Wiener step only, ref = guide   */
#define ISP_BM3D_BOTH 3  /* BM3D: basic then final                      */

/* Where the caller's buffers live (the `mem` argument). */
#define ISP_BM3D_MEM_HOST 0
#define ISP_BM3D_MEM_DEVICE 1

/* Arithmetic of the transforms and shrinkage, and of the matching distances.
 * PREC_FP32: FP32-accurate transforms and FP32 shrinkage (hard-threshold
 * decisions within a narrow band of the threshold re-made in FP64), FP64
 * distances summed in the reference's order (its tie order exactly).
 * PREC_FP32_FAST: FP32 transforms on the vector ALUs and FP32 shrinkage
 * without the FP64 re-decision, and FP32 distances (fused multiply-add) as
 * the mixed forms match.  PREC_FP16 / PREC_BF16: 16-bit MFMA operands with
 * FP32 accumulation, FP32 shrinkage, FP32 distances.  PREC_FP32_MFMA: the
 * FP32_FAST form's arithmetic with the transform products on the FP16 matrix
 * cores, each operand split into fp16(a) + fp16(a - fp16(a)) (22 significant
 * bits) and each product taken as three MFMAs with FP32 accumulation.  The
 * plain form's transforms are the same split products (the environment
 * variable ISP_BM3D_FP32_VALU=1 puts them back on the FP32 vector ALUs).
 * Every form aggregates in 64-bit fixed point (exact, order-independent sums;
 * see isp_bm3d_aggregation_limits). */
#define ISP_BM3D_PREC_FP32 0      /* plain FP32 path, the reference's groups exactly */
#define ISP_BM3D_PREC_FP16 1      /* FP16 MFMA transforms, FP32 accumulation        */
#define ISP_BM3D_PREC_BF16 2      /* BF16 MFMA transforms, FP32 accumulation        */
#define ISP_BM3D_PREC_FP32_FAST 3 /* FP32 throughout: FP32 distances, FP32 filter   */
#define ISP_BM3D_PREC_FP32_MFMA 4 /* FP32-accurate products on the matrix cores     */

/* Mirror of BM3D_Para_Base (include/BM3D.h:15-45). */
typedef struct isp_bm3d_step_params {
    char profile[8];  /* "fast", "lc", "np", "high" or "vn"            */
    double sigma[3];  /* per plane, 0..255 units                        */
    int block_size;   /* BlockSize                                      */
    int block_step;   /* BlockStep                                      */
    int group_size;   /* GroupSize                                      */
    int bm_range;     /* BMrange                                        */
    int bm_step;      /* BMstep                                         */
    double thMSE;     /* before the OPP-norm scaling of the constructor */
    double lambda;    /* hard-threshold multiplier (basic step)         */
} isp_bm3d_step_params;

/* Mirror of BM3D_Para (include/BM3D.h:154-168). */
typedef struct isp_bm3d_params {
    isp_bm3d_step_params basic;
    isp_bm3d_step_params final_;
} isp_bm3d_params;

typedef struct isp_bm3d_ctx isp_bm3d_ctx;

/* BM3D_Para(profile) including thMSE_Default() (include/BM3D.h:27-168).
 * profile NULL means "fast", the reference default. */
int isp_bm3d_params_default(isp_bm3d_params *params, const char *profile);

/* BM3D_Basic_Para(profile) / BM3D_Final_Para(profile) (include/BM3D.h:48-151).
 * final_step = 0 for the basic step, 1 for the final step. */
int isp_bm3d_step_params_default(isp_bm3d_step_params *p, const char *profile, int final_step);

/* thMSE_Default() (include/BM3D.h:89-97 basic, 142-150 final): recompute
 * thMSE from sigma[0] and the profile. */
int isp_bm3d_thmse_default(isp_bm3d_step_params *p, int final_step);

/* Creates a filter context on HIP device `device` (BM3D(const BM3D_Para&),
 * include/BM3D.h:349-351; tables of source/BM3D.cpp:9-73).  The context owns
 * its device tables and workspace; it is immutable after creation except for
 * the precision (isp_bm3d_set_precision), the variant modes
 * (isp_bm3d_set_variant), tile tuning and timing. */
int isp_bm3d_create(isp_bm3d_ctx **ctx, const isp_bm3d_params *params, int device);
int isp_bm3d_destroy(isp_bm3d_ctx *ctx);

/* Message of the last failed call on ctx (or of the last failed create when
 * ctx is NULL).  Never NULL. */
const char *isp_bm3d_last_error(const isp_bm3d_ctx *ctx);

/* Transform arithmetic; ISP_BM3D_PREC_FP32 by default. */
int isp_bm3d_set_precision(isp_bm3d_ctx *ctx, int precision);

/* Variant modes of the filter, off by default (flags 0: the reference's
 * semantics).  No reference counterpart: the reference transforms the group
 * axis with a DCT of the group's length and aggregates each block with one
 * weight (source/BM3D.cpp:14-24, 378-384, 450-456).
 *   ISP_BM3D_VARIANT_HAAR: the orthonormal Haar transform along the group
 *     axis, each group truncated to the largest power of two of its members;
 *   ISP_BM3D_VARIANT_KAISER: each block aggregated with a separable Kaiser
 *     window of shape kaiser_beta (2.0 is the usual choice).
 * Both keep the orthonormal thresholds (lambda sigma) and Wiener noise
 * (sigma^2 / 2^zeta).  Variant calls run the generic FP64 kernels in every
 * precision (BlockSize <= 16, GroupSize <= 64); batches, multi-frame stacks
 * and autotuning refuse them. */
#define ISP_BM3D_VARIANT_HAAR 1
#define ISP_BM3D_VARIANT_KAISER 2
int isp_bm3d_set_variant(isp_bm3d_ctx *ctx, int flags, double kaiser_beta);

/* One float plane (Plane_FL): FilterIF2::process(dst, src, ref) for
 * BM3D_Basic (mode BASIC, source/BM3D.cpp:80-136,339-385), BM3D_Final (mode
 * FINAL, :80-136,411-457) or BM3D (mode BOTH, :464-472, with the
 * intermediate estimate sized like src).  ref may be NULL (= src).  stride is
 * in elements and applies to dst, src and ref.  mem selects host or device
 * pointers; stream is a hipStream_t (NULL = default stream).  Host calls
 * return after the result is in dst; device calls are stream-ordered. */
int isp_bm3d_plane_f32(isp_bm3d_ctx *ctx, float *dst, const float *src, const float *ref,
                       int width, int height, size_t stride, int mode, int mem, void *stream);

/* isp_bm3d_plane_f32 on device buffers that allocates nothing.  No
 * reference counterpart: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The call's workspace (the step-1 estimate, one plane of
 * fixed-point sums, the group lists) lives in a device array of 96 MiB per
 * GPU that the library holds from context creation until the process ends;
 * no hipMalloc, hipFree or other allocating routine runs in the call.  A plane
 * whose workspace does not fit (about 2 Mpixel at GroupSize 32 and
 * BlockStep 3, e.g. 1920x1080 at stride 1920) returns ISP_BM3D_E_UNSUPPORTED.
 * dst, src and ref (or NULL) are device pointers with row stride `stride`;
 * the call is stream-ordered on `stream` and stage timing is off during it.
 * Every call on one GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU (other streams, other threads,
 * other contexts) race and must be serialised by the caller.  The groups it
 * forms are not kept for isp_bm3d_last_groups. */
int isp_bm3d_plane_f32_noalloc(isp_bm3d_ctx *ctx, float *dst, const float *src, const float *ref,
                               int width, int height, size_t stride, int mode, void *stream);

/* Float RGB frame (planar R, G, B in [0,1]) with the semantics of
 * process_Frame (source/BM3D.cpp:282-313 for BASIC / FINAL, :501-527 for
 * BOTH): OPP conversion without clip, matching on Y, OPP inverse with clip to
 * [0,1].  ref_* may all be NULL (= src).  For BOTH a separate ref only guides
 * step-1 matching (its Y); for FINAL it is the Wiener guide. */
int isp_bm3d_rgb_f32(isp_bm3d_ctx *ctx, float *dst_r, float *dst_g, float *dst_b,
                     const float *src_r, const float *src_g, const float *src_b,
                     const float *ref_r, const float *ref_g, const float *ref_b,
                     int width, int height, size_t stride, int mode, int mem, void *stream);

/* Integer Plane (int32 samples with the quantisation of a Plane:
 * Floor / Neutral / Ceil, include/Image_Type.h:51-180; chroma iff
 * Floor < Neutral): BM3D_Base::process_Plane / BM3D::process_Plane
 * (source/BM3D.cpp:248-279, 475-498): RangeConvert to float, the filter,
 * RangeConvert back with +0.5 rounding and clip.  Otherwise as
 * isp_bm3d_plane_f32. */
int isp_bm3d_plane_i32(isp_bm3d_ctx *ctx, int *dst, const int *src, const int *ref, int width, int height,
                       size_t stride, int floor_, int neutral, int ceil_, int mode, int mem, void *stream);

/* Integer RGB Frame (int32 R, G, B with Floor / Ceil): process_Frame
 * (source/BM3D.cpp:282-313, 501-527) exactly, including OPP from int
 * (include/Conversion.hpp:547-578) and OPP to int with +0.5 rounding and clip
 * (:660-762).  Otherwise as isp_bm3d_rgb_f32. */
int isp_bm3d_rgb_i32(isp_bm3d_ctx *ctx, int *dst_r, int *dst_g, int *dst_b, const int *src_r, const int *src_g,
                     const int *src_b, const int *ref_r, const int *ref_g, const int *ref_b, int width, int height,
                     size_t stride, int floor_, int ceil_, int mode, int mem, void *stream);

/* Multi-frame grouping: BM3D of a stack of `frames` grey planes (nplanes 1)
 * or float RGB frames (nplanes 3, the semantics of isp_bm3d_rgb_f32) in which
 * the groups of frame t draw blocks from frames |z - t| <= radius (0..4) and
 * each filtered block is aggregated into its own frame.  The reference has
 * the Pos3 group codes for this (BlockGroup, include/Block.h:843-846, 927-936,
 * 1141-1208) but forms no such groups: candidates are the single-frame search
 * window's positions in every frame of the window, frames visited t, t-1,
 * t+1, t-2, ..., with BlockMatchingMulti's test and (key, scan order) sort.
 * With one frame or radius 0 the result equals the single-frame call on each
 * frame.  dst, src and ref (or NULL) hold frames * nplanes plane pointers,
 * frame-major ([t * nplanes + c]); mode, mem and stream as in
 * isp_bm3d_plane_f32. */
int isp_bm3d_stack_f32(isp_bm3d_ctx *ctx, float *const *dst, const float *const *src, const float *const *ref,
                       int frames, int nplanes, int width, int height, size_t stride, int radius, int mode, int mem,
                       void *stream);
/* Block matching of one step only (BM3D_Base::BlockMatching,
 * source/BM3D.cpp:221-236, include/Block.h:658-703) on a float guide plane,
 * for group-index parity checks.  For grid point g = gy*ngx + gx:
 * counts[g] = N, pos[g*G + z] = (y << 16) | x, keys[g*G + z] = distance key
 * (MSE units; may be NULL).  Outputs are host arrays; the guide follows mem.
 * final_step selects which step's parameters are used. */
int isp_bm3d_match(isp_bm3d_ctx *ctx, const float *guide, int width, int height, size_t stride,
                   int final_step, int mem, int *counts, int *pos, double *keys);

/* Per-stage device timing (no reference counterpart; the reference has no
 * timers).  When enabled, every kernel launch of the library is bracketed by
 * hipEvents on the caller's stream.  isp_bm3d_get_timing() waits for the
 * recorded events, adds their elapsed times into per-stage totals and copies
 * up to `max` stages: names[i] (static strings), total_ms[i], launches[i].
 * It returns the number of stages (ISP_BM3D_NUM_STAGES). */
#define ISP_BM3D_NUM_STAGES 10
int isp_bm3d_set_timing(isp_bm3d_ctx *ctx, int enable);
int isp_bm3d_get_timing(isp_bm3d_ctx *ctx, const char **names, double *total_ms, long *launches, int max);
int isp_bm3d_reset_timing(isp_bm3d_ctx *ctx);

/* Groups formed by the most recent plane/RGB call for one step, copied to
 * host arrays: counts[g] (capacity >= grid points) and, when pos is not NULL,
 * pos[g*G + z] (capacity*G entries).  Waits for the device.  Returns the
 * number of grid points.  No reference counterpart (the reference keeps its
 * match codes local to Kernel, source/BM3D.cpp:122,190); for parity checks. */
int isp_bm3d_last_groups(isp_bm3d_ctx *ctx, int final_step, int *counts, int *pos, size_t capacity);

/* Grid sizes of a step for an image (source/BM3D.cpp:98-127). */
int isp_bm3d_grid_size(const isp_bm3d_ctx *ctx, int width, int height, int final_step,
                       int *ngx, int *ngy);

/* What a call of this size would do, computed on the host alone (no device,
 * no context; no reference counterpart): per step [basic, final], whether it
 * copies src through (degenerate parameters), whether it runs the generic
 * kernels (BlockSize != 8 or GroupSize > 32), its reference grid and its
 * Wiener weight cap; the device workspace a plane / frame call of this size
 * allocates, and the scratch isp_bm3d_plane_f32_noalloc needs (stride =
 * width) against the scratch it has.  Returns ISP_BM3D_OK, or the status the
 * filter calls would return for these parameters and sizes. */
typedef struct isp_bm3d_plan_info {
    int degenerate[2];
    int generic[2];
    int ngx[2], ngy[2];
    double weight_cap[2];
    size_t workspace_bytes;
    size_t noalloc_bytes;
    size_t noalloc_capacity;
} isp_bm3d_plan_info;
int isp_bm3d_plan(const isp_bm3d_params *params, int width, int height, int nplanes, isp_bm3d_plan_info *info);

/* Headroom of the aggregation sums (no reference counterpart: the reference
 * sums in FP64).  num and den are 64-bit fixed point with 32 fraction bits.
 * For a step's parameters and a temporal radius (0 for single frames):
 * *max_contributions = an upper bound on the (group, member) contributions
 * to one pixel; *weight_cap = the cap applied to Wiener group weights
 * 1 / sum(W^2), the largest power of two <= 2^16 with
 * max_contributions * weight_cap * 4 < 2^30, so sums of samples and
 * reconstructions of magnitude <= 4 cannot overflow.  Returns
 * ISP_BM3D_E_UNSUPPORTED (weight_cap 0) when even weights <= 1 could. */
int isp_bm3d_aggregation_limits(const isp_bm3d_step_params *p, int radius, double *max_contributions,
                                double *weight_cap);

/* OPP colour conversions of float device planes (MatrixConvert_RGB2YUV /
 * MatrixConvert_YUV2RGB with ColorMatrix::OPP, include/Conversion.hpp:547-578,
 * 660-762): Y=(R+G+B)/3, U=(R-B)/2, V=(R-2G+B)/4 without clip, and the inverse
 * with clip to [0,1] when clip != 0.  Enqueued on `stream`, nothing synchronises. */
int isp_bm3d_rgb_to_opp(float *y, float *u, float *v, const float *r, const float *g, const float *b, int width,
                        int height, size_t stride, void *stream);
int isp_bm3d_opp_to_rgb(float *r, float *g, float *b, const float *y, const float *u, const float *v, int width,
                        int height, size_t stride, int clip, void *stream);

/* ---------------------------------------------------------------------------
 * Batches: many images of any sizes filtered together, one launch per stage
 * for the whole batch, replayed as one hipGraph.  No reference counterpart
 * (ISP_MW filters one image per call, include/Filter.h:173-216); each image's
 * result equals a single call on it with the same parameters and tiles, bit
 * for bit.
 * ------------------------------------------------------------------------- */
typedef struct isp_bm3d_image {
    float *dst[3];          /* device planes: [0] for grey (nplanes 1), [0..2] R, G, B */
    const float *src[3];
    const float *ref[3];    /* or NULL (= src): BASIC and BOTH, the step-1 matching guide (its luma
                               for RGB); FINAL, the Wiener guide of BM3D_Final::process(dst, src, ref) */
    int width, height;
    size_t stride;          /* row pitch of dst, src and ref in elements */
    double sigma;           /* > 0: this image's sigma (0..255 units) for every plane of both
                               steps, thMSE from thMSE_Default; <= 0: the context's parameters */
} isp_bm3d_image;

typedef struct isp_bm3d_batch isp_bm3d_batch;

/* flags */
#define ISP_BM3D_BATCH_AUTO_GROUP 1  /* the final step's GroupSize of each image from the
                                        tuning table of its size class, when one is set */
#define ISP_BM3D_TUNE_RGB 2          /* isp_bm3d_autotune: measure RGB frames (three planes) */

/* mode: ISP_BM3D_BASIC, ISP_BM3D_FINAL (ref: the guide frames) or ISP_BM3D_BOTH.
 * Buffers are device memory; the plan
 * keeps the pointers (change them with isp_bm3d_batch_update). */
int isp_bm3d_batch_create(isp_bm3d_ctx *ctx, isp_bm3d_batch **batch, const isp_bm3d_image *images, int count,
                          int nplanes, int mode, int flags);
/* Runs the whole batch on `stream` (hipStream_t): use_graph 1 replays the
 * captured hipGraph, 0 enqueues the same launches one by one. */
int isp_bm3d_batch_run(isp_bm3d_batch *batch, void *stream, int use_graph);
/* New buffers and sigmas for images of the same sizes (waits for the device). */
int isp_bm3d_batch_update(isp_bm3d_batch *batch, const isp_bm3d_image *images);
/* kernel launches per run, device bytes held, workgroups of the largest launch */
int isp_bm3d_batch_info(const isp_bm3d_batch *batch, int *launches, size_t *device_bytes, int *max_workgroups);
int isp_bm3d_batch_destroy(isp_bm3d_batch *batch);

/* ---------------------------------------------------------------------------
 * Autotuning: tile shapes (reference blocks per workgroup) per size class,
 * measured on a batch of seeded synthetic images; used by every later call
 * and batch on the context.  Tile shapes never change results.  With
 * ISP_BM3D_BATCH_AUTO_GROUP the tuner also tries a final-step GroupSize of 16
 * (against the context's) and keeps it when it is faster and its PSNR on the
 * synthetic images is within 0.05 dB.  No reference counterpart.
 * ------------------------------------------------------------------------- */
typedef struct isp_bm3d_tuning {
    int match_tile[2];      /* x, y */
    int filter_tile[2];     /* of the engine the context's precision selects */
    int final_group_size;   /* 0: the context's */
    double ms_per_image;    /* measured with the chosen shapes */
    double psnr_db;         /* of the synthetic images with the chosen shapes */
} isp_bm3d_tuning;

int isp_bm3d_size_class(int width, int height);  /* ceil(log2(width * height)) */
int isp_bm3d_autotune(isp_bm3d_ctx *ctx, int width, int height, double sigma, int flags, isp_bm3d_tuning *out);
/* t NULL clears the entry */
int isp_bm3d_set_tuning(isp_bm3d_ctx *ctx, int size_class, const isp_bm3d_tuning *t);
/* returns 1 and fills t when the class has an entry, 0 otherwise */
int isp_bm3d_get_tuning(const isp_bm3d_ctx *ctx, int size_class, isp_bm3d_tuning *t);

#ifdef __cplusplus
}
#endif

#endif /* ISP_BM3D_H_ */
