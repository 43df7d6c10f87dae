/*
 * FP64 CPU restatement of ISP_MW's BM3D filter.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP path and the timed CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product library
 * (libisp_bm3d.so) never links or calls it.
 *
 * Everything is double precision, as in the reference CPU build where
 * FLType is double (include/Type.h:26-30).
 */
#ifndef ISP_BM3D_ORACLE_H_
#define ISP_BM3D_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of BM3D_Para_Base (include/BM3D.h:15-45): values as the user gives
 * them; sigma in 0..255 units, thMSE in MSE units of the 8-bit scale. */
typedef struct orc_step_params {
    int block_size;
    int block_step;
    int group_size;
    int bm_range;
    int bm_step;
    double thMSE;
    double lambda;
    double sigma[3];
} orc_step_params;

/* Variant modes (no reference counterpart; the configs' "1D Haar" and
 * "Kaiser window beta=2.0"), for every later filter call of the process:
 * haar != 0 transforms the group axis with the orthonormal Haar matrix and
 * truncates each group to the largest power of two of its members;
 * kaiser_beta != 0 aggregates each block with a Kaiser window of that shape.
 * (0, 0) restores the reference's semantics. */
int orc_set_variant(int haar, double kaiser_beta);

/* Reference grid of one axis (source/BM3D.cpp:98-127).  Writes at most
 * `cap` positions, returns the count. */
int orc_grid(int n, int block_size, int block_step, int *out, int cap);

/* Block::_SearchBoundary (include/Block.h:716-745). */
int orc_search_boundary(int pos, int plane_boundary, int search_range, int search_step);

/* Block matching for every reference position of the grid
 * (BM3D_Base::BlockMatching, source/BM3D.cpp:221-236, and
 * Block::BlockMatchingMulti, include/Block.h:578-711) on one plane `ref`
 * (range 1, row stride `stride`).  For grid point g = gy*ngx+gx:
 * counts[g] = N (1..group_size), pos[g*group_size+z] = (y<<16)|x, and
 * keys[g*group_size+z] = distance key (MSE units).  keys may be NULL. */
int orc_match_all(const orc_step_params *p, const double *ref, int w, int h, int stride,
                  int *counts, int32_t *pos, double *keys);

/* BM3D_Base::Kernel (source/BM3D.cpp:80-218).  nplanes = 1 or 3; matching
 * runs on ref[0]; plane c is filtered when sigma[c] > 0 (nplanes=3) or
 * sigma[0] > 0 (nplanes=1).  wiener = 0 -> BM3D_Basic, 1 -> BM3D_Final.
 * dst planes are written in full (planes that are not filtered are copied
 * from src, the behaviour of the two-step Frame path).
 * Returns 0 on success, 1 if the parameters made it copy src through,
 * negative on error. */
int orc_kernel(const orc_step_params *p, int wiener, int nplanes,
               double *const *dst, const double *const *src, const double *const *ref,
               int w, int h);

/* Two-step BM3D on one plane with the intermediate estimate sized like src
 * (BM3D::process_Plane_FL, source/BM3D.cpp:464-472, with tmp sized). */
int orc_bm3d_plane(const orc_step_params *basic, const orc_step_params *final_,
                   double *dst, const double *src, const double *ref, int w, int h);

/* Two-step BM3D on a float RGB frame in [0,1] with the semantics of
 * BM3D::process_Frame (source/BM3D.cpp:501-527): OPP conversion without clip,
 * step 1 matched on Y of ref (or of src when ref is NULL), step 2 matched on
 * the basic Y, OPP inverse with clip to [0,1].  mode: 1 basic only
 * (BM3D_Basic::process_Frame), 2 final only (ref is then the guide frame,
 * BM3D_Final::process_Frame), 3 both. */
int orc_bm3d_rgb(const orc_step_params *basic, const orc_step_params *final_, int mode,
                 double *const *dst, const double *const *src, const double *const *ref,
                 int w, int h);

/* OPP conversions on float planes (include/Conversion.hpp:547-578, 739-762)
 * for a float RGB frame with Floor 0, Ceil 1. */
void orc_rgb2opp(double *y, double *u, double *v, const double *r, const double *g,
                 const double *b, long n);
void orc_opp2rgb(double *r, double *g, double *b, const double *y, const double *u,
                 const double *v, long n, int clip);

/* Integer Plane / Frame paths (BM3D_Base::process_Plane, BM3D::process_Plane,
 * source/BM3D.cpp:248-279, 475-498; process_Frame on an int Frame,
 * :282-313, 501-527).  Quantisation Floor / Neutral / Ceil as in Plane
 * (include/Image_Type.hpp:85-117; chroma iff Floor < Neutral); mode as orc_bm3d_rgb. */
void orc_range_to_float(double *dst, const int32_t *src, long n, int floor_, int neutral, int ceil_);
void orc_range_to_int(int32_t *dst, const double *src, long n, int floor_, int neutral, int ceil_);
int orc_bm3d_plane_int(const orc_step_params *basic, const orc_step_params *final_, int mode,
                       int32_t *dst, const int32_t *src, const int32_t *ref, int w, int h,
                       int floor_, int neutral, int ceil_);
int orc_bm3d_rgb_int(const orc_step_params *basic, const orc_step_params *final_, int mode,
                     int32_t *const *dst, const int32_t *const *src, const int32_t *const *ref,
                     int w, int h, int floor_, int ceil_);

/* NLMeans (include/NLMeans.h:11-22 NLMeans_Para): the same block matching
 * with thMSE not scaled by normY, weights exp(-key / strength^2), and the
 * stationarity correction when `correction` is set. */
typedef struct orc_nlm_params {
    int correction;
    double sigma, strength;
    int block_size, block_step, group_size, bm_range, bm_step;
    double thMSE;
} orc_nlm_params;

/* NLMeans::process_Plane_FL (source/NLMeans.cpp:150-228); ref may be NULL (src). */
int orc_nlm_plane(const orc_nlm_params *p, double *dst, const double *src, const double *ref, int w, int h);
/* NLMeans::process_Frame on an int RGB frame (source/NLMeans.cpp:279-375):
 * matching on the OPP luma of ref (or src), averaging of the R, G, B samples,
 * output (int)(num/den + 0.5). */
int orc_nlm_rgb_int(const orc_nlm_params *p, int32_t *const *dst, const int32_t *const *src,
                    const int32_t *const *ref, int w, int h, int floor_, int ceil_);
/* NLMeans::process_Plane on an int plane (source/NLMeans.cpp:232-257). */
int orc_nlm_plane_int(const orc_nlm_params *p, int32_t *dst, const int32_t *src, const int32_t *ref, int w, int h,
                      int floor_, int neutral, int ceil_);

/* Unnormalised 3D DCT on an n0 x n1 x n2 array (n2 fastest): FFTW REDFT10
 * forward (inverse = 0) or REDFT01 backward (inverse = 1) on every axis. */
void orc_dct3d(double *data, int n0, int n1, int n2, int inverse);

/* Row-band threads used by orc_kernel / orc_bm3d_* (1 = the reference's
 * single-threaded loop order).  Returns the previous value. */
int orc_set_threads(int n);

/* Multi-frame grouping over a stack of frames (Pos3 groups, include/Block.h:843-846,
 * 927-936, 1141-1208): candidates from frames |z - t| <= radius, visited t, t-1, t+1,
 * ..., members aggregated into their own frames.  Planes are [frame * nplanes + c]. */
int orc_match_stack(const orc_step_params *p, const double *const *guide, int frames, int radius, int t, int w,
                    int h, int *counts, int32_t *pos, int *fz);
int orc_kernel_stack(const orc_step_params *p, int wiener, int nplanes, int frames, int radius,
                     double *const *dst, const double *const *src, const double *const *ref, int w, int h);
int orc_bm3d_stack(const orc_step_params *basic, const orc_step_params *final_, int mode, int nplanes, int frames,
                   int radius, double *const *dst, const double *const *src, const double *const *ref, int w, int h);

#ifdef __cplusplus
}
#endif

#endif
