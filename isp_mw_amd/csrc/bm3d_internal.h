// Internal interface between the C ABI (bm3d_capi.cpp) and the HIP kernels
// (bm3d_kernels.hip).  Not installed; no reference counterpart.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ispbm3d {

constexpr int kBlock = 8;        // BlockSize implemented by the GPU kernels
constexpr int kMaxGroup = 32;    // largest GroupSize implemented
// Default tile shapes (reference blocks per workgroup, x by y); the tile
// shape is a launch parameter (MatchArgs / FilterArgs trx, try_) that the
// autotuner may change.
constexpr int kWalkTileX = 4;    // matching by column walks: 4 columns of 16 blocks
constexpr int kWalkTileY = 16;
constexpr int kTileRX = 8;       // matching and the FP32 filter
constexpr int kTileRY = 4;
constexpr int kStrip = 5;        // candidates per lane strip in matching
constexpr int kStripMax = 6;     // widest strip (column walks of one round); the tile's margin
// Capacity of a wave's accepted-candidate list in matching.  Windows with
// more candidates reduce the list to its G-1 best whenever the next batch of
// 64 strips could overflow it (capacity >= 31 + 64 * kStrip).
constexpr int kMatchListCap = 384;
constexpr int kWaves = 4;        // waves per workgroup
constexpr int kThreads = 64 * kWaves;
// MFMA filter: 8 waves over a tile of 8 x 8 reference blocks by default
constexpr int kMTileRX = 8;
constexpr int kMTileRY = 8;
constexpr int kMWaves = 8;
constexpr int kMThreads = 64 * kMWaves;
constexpr int kMaxTileRefs = 256;  // trx * try_ bound
constexpr int kPairTables = 16 * 17;  // (N1, N2) operand tables of two groups in one MFMA operand

// Batched launches: one launch covers the tiles (or rows) of many images.
// Workgroup b reads unit = map[b]: image (unit >> 20) and, for tiled kernels,
// tile ((unit >> 10) & 1023, unit & 1023) as (y, x), for row kernels row
// (unit & 0xFFFFF); its arguments are args[image].
constexpr int kBatchImageShift = 20;
constexpr int kMaxBatchImages = 4096;
constexpr int kMaxBatchTiles = 1024;  // tiles per image per axis
constexpr int kMaxBatchRows = 1 << 20;

// Fixed-point scale of the aggregation sums: value * 2^32, two's complement
// in 64 bits, so a pixel's sums must stay below 2^63 / 2^32 = 2^31 in value.
// A pixel p is covered by block origins o with o <= p < o + B on each axis;
// o can be a member of the group of every reference block whose window holds
// it, once per group.  Per axis the references whose windows reach
// [p - B + 1, p] lie in a span of 2R + B positions, so at most
// floor((2R + B - 1) / s) + 2 grid points (the +2: both ends, and the grid's
// last position clamped to n - B), and a group holds at most min(G, B^2) of
// the origins covering p (fewer at BMstep > 1).  That gives
// aggregation_bound().  Hard-threshold weights are <= 1; Wiener weights
// 1 / sum(W^2) exceed 1 only where every
// gain of the group is small, and are capped at wiener_weight_cap(): the
// largest power of two (at most 2^16) for which bound * cap * kSampleBound
// stays below 2^30, a factor 2 inside the 2^31 limit.  With samples of
// magnitude <= kSampleBound, num and den then cannot overflow, and each
// addend (|value| <= 4 * 2^16 = 2^18) lies inside the 2^19 for which the
// FP64 rounding of the device helper (fix_bits) is exact.  The cap acts
// only on groups whose guide carries almost no energy (sum(W^2) < 1 / cap),
// whose estimates are near zero.
constexpr float kFixScale = 4294967296.0f;         // 2^32
constexpr double kFixInv = 1.0 / 4294967296.0;     // 2^-32
constexpr double kFixSumLimit = 1073741824.0;      // 2^30: bound * cap * kSampleBound stays below
constexpr double kSampleBound = 4.0;               // |samples| and |reconstructions| the sums allow
constexpr float kMaxWienerWeight = 65536.0f;       // 2^16, the cap's upper end

// Largest number of (group, member) contributions to one pixel's sums for a
// step with BlockSize B, BlockStep s, BMrange R (rounded to BMstep st),
// GroupSize G, and 2 * radius + 1 frames whose groups reach a frame
// (multi-frame groups).  A window's members lie on a lattice of step st, so at
// most ceil(B / st) of them per axis cover a pixel.
inline double aggregation_bound(int B, int s, int R, int st, int G, int radius)
{
    const double axis = (double)((2 * R + B - 1) / s + 2);
    const int per_axis = (B + st - 1) / st;
    const int members = per_axis * per_axis;
    return axis * axis * (double)(G < members ? G : members) * (double)(2 * radius + 1);
}

// Cap of the Wiener group weight (a power of two in [1, 2^16]); 0 when even
// weights <= 1 could overflow the sums.
inline float wiener_weight_cap(int B, int s, int R, int st, int G, int radius)
{
    const double bound = aggregation_bound(B, s, R, st, G, radius);
    if (bound * kSampleBound >= kFixSumLimit) return 0.f;
    float cap = kMaxWienerWeight;
    while (cap > 1.f && (double)cap * bound * kSampleBound >= kFixSumLimit) cap *= 0.5f;
    return cap;
}

// Device scratch of the allocation-free entry point (scratch.hip): a
// file-scope device array of kScratchBytes per GPU.  scratch_address() gives
// the array's address on `device`, which must be the current device; the
// runtime is asked once per device.
constexpr size_t kScratchBytes = size_t(96) << 20;  // 96 MiB
hipError_t scratch_address(int device, void **out);

// Positions are packed (y << 16) | x.
struct MatchArgs {
    const void *guide;       // matching guide plane (float or double by launch)
    int gpitch;              // its row pitch in elements
    int trx, try_;           // tile shape in reference blocks
    int W, H;
    int ngx, ngy;
    int bstep;               // BlockStep: grid position k is min(k*bstep, last)
    int range;               // BMrange rounded down to a multiple of BMstep
    int step;                // BMstep
    int G;                   // GroupSize
    int bypass;              // 1: every group is the reference block alone
    int exact;               // 1: FP64 distances rounded term by term as the reference does;
                             // 0: FP32 distances (fused multiply-add) and 32-bit keys
    double thSSE;            // thMSE * normY * MSE2SSE
    double distMul;          // 1 / MSE2SSE
    int tileW, tileH, tilePitch;  // staged guide tile (doubles)
    int maxCand;             // per-wave candidate list capacity: min(window size, kMatchListCap)
    int *cnt_out;            // [ngy*ngx]
    int *pos_out;            // [ngy*ngx*G]
    double *key_out;         // [ngy*ngx*G] or nullptr
    // Multi-frame grouping: candidates from another frame's plane `search`
    // (same type and pitch as guide) against the reference blocks of guide.
    // Such a list has no self entry: cnt_out = N <= G-1, pos_out[0..N).
    const void *search;      // nullptr: the guide itself (the single-frame search)
    int refW, refH, refPitch;  // staged reference-block tile of a cross-frame search
    int walk;                // 1: match_walk_kernel (see match_walk_eligible)
    int B;                   // BlockSize (match_generic_kernel; the other kernels use kBlock)
};

// The column-walk matching kernel takes FP32 distances of the single-frame
// search with BlockStep 3..7 (neighbouring blocks of a column share rows, and
// a row touches at most 4 of them).  Windows beyond 128 lane strips of kStrip
// offsets are walked in chunks; scan indices must fit 16 bits (at most 255
// offsets per axis).
inline bool match_walk_eligible(const MatchArgs &a)
{
    if (a.exact || a.search || a.bstep < 3 || a.bstep >= 8 || a.step < 1) return false;
    return 2 * (a.range / a.step) + 1 <= 255;
}

// Multi-frame grouping: merge of the per-frame candidate lists of one
// reference frame into its groups.
constexpr int kMaxStackRadius = 4;
constexpr int kMaxStackLists = 2 * kMaxStackRadius + 1;
struct StackMergeArgs {
    int refs;                 // grid points
    int G;                    // GroupSize: row length of every list and of the output
    int nlists;               // list 0: the reference frame's own (self first); then by frame priority
    const int *cnt[kMaxStackLists];
    const int *pos[kMaxStackLists];
    const double *key[kMaxStackLists];
    int slot[kMaxStackLists]; // frame slot (window-relative) of each list's members
    int *cnt_out, *pos_out, *slot_out;
};

struct FilterArgs {
    int trx, try_;            // tile shape in reference blocks
    int W, H, pitch;          // geometry of src/guide/num/den planes
    int ngx, ngy;
    int bstep;
    int range;
    int G;
    const int *cnt;           // groups from matching
    const int *pos;
    const float *src[3];      // noisy planes
    const float *guide[3];    // Wiener guide planes (final step)
    // Aggregation accumulators, zeroed before launch: 64-bit fixed point
    // with kFixScale fractional bits (exact, order-independent sums).
    unsigned long long *num[3];
    unsigned long long *den[3];
    int active[3];            // plane filtered?
    float thr[3];             // lambda * sigma_c (orthonormal domain)
    float sig2[3];            // sigma_c^2 (orthonormal domain)
    float wmax;               // cap of the Wiener group weight (wiener_weight_cap)
    int FW, FH;               // accumulation footprint of a tile
    const float *c8;          // 8x8 orthonormal DCT-II matrix, [k][n]
    const float *cn;          // per-N orthonormal DCT-II matrices, padded to NP x NP
    int cn_off[kMaxGroup + 1];
    // Hard-threshold decisions within `band` of the threshold are re-made
    // from the coefficient recomputed in FP64 (FP64 matrices, unpadded N x N).
    const double *c8d;
    const double *cnd;
    int cnd_off[kMaxGroup + 1];
    double thr64[3];
    float band[3];
    // MFMA operand tables (16-bit elements, f16 or bf16 by precision):
    // kb[4][2][64][8]  K64 as B of T = X K64 (forward, spatial)
    // ka[2][4][64][8]  K64 as A of X^T = K64 V^T (inverse, spatial)
    // cna[N][2][64][8] C_N as A of Y = C_N T (forward, group axis)
    // cnb[N][2][64][8] C_N as B of V^T = Y^T C_N (inverse, group axis)
    // cnap / cnbp[(N1-1)*17 + N2][2][64][8]: the same for two groups of
    // N1 (1..16) and N2 (0..16) blocks in one operand, diag(C_N1, C_N2)
    const uint16_t *kb, *ka, *cna, *cnb, *cnap, *cnbp;
    // The FP32_MFMA form only (else null): the low halves of the FP16 tables
    // above, fp16(v - fp16(v)), in the same layouts.
    const uint16_t *kbl, *kal, *cnal, *cnbl, *cnapl, *cnbpl;
    // Multi-frame groups (stack kernels only).  One launch covers every
    // frame of the stack: the groups of reference frame t are cnt / pos /
    // slot + t * (ngx*ngy) (times G for pos and slot); member z of a group
    // lies in frame slot[...] (absolute), whose src / guide / num / den
    // planes are the pointers plus frame * fstride elements.  A workgroup
    // walks the frames in order for its tile, with `nslots` accumulator
    // slots used as a ring (frame f in slot f % nslots), and flushes frame
    // f once frame f + radius is done.
    const int *slot;
    int nslots;
    int frames, radius;
    size_t fstride;
};

struct FinalizeArgs {
    int W, H, pitch;
    const unsigned long long *num[3];
    const unsigned long long *den[3];
    int den_box;              // 0: den per pixel; kBlock: den per block origin,
                              // a pixel's den the sum over the 8 x 8 origins
                              // whose block covers it
    const float *src[3];      // copied through for inactive planes
    int active[3];
    int nplanes;
    int to_rgb;               // OPP -> RGB (include/Conversion.hpp:660-762)
    int out_int;              // 0: float outputs, 1: int32 outputs (+0.5 rounding)
    double gain_y, gain_c;    // RGB: YUV->RGB gains (dRange / sRange)
    double offset;            // RGB: dFloor (+0.5 for int); plane int: Neutral + 0.5
    double q_gain;            // plane int: range
    int clip;                 // clip outputs to [lo, hi]
    double lo, hi;
    void *out[3];
    size_t out_pitch;
};

struct OppArgs {
    int W, H;
    const void *r, *g, *b;     // float or int32 samples
    int in_int;
    double gainY, offsetY, gainU, gainV, offsetC;
    size_t in_pitch;
    float *y, *u, *v;          // may be null
    int pitch;
    double *y64;               // may be null (matching guide)
    int pitch64;
    int clipY;                 // clip Y to [0, 1] (ConvertToY's default)
};

// int32 plane -> float plane: v * gain + offset (RangeConvert), optional clip
struct ToFloatArgs {
    int W, H;
    const int *src;
    size_t spitch;
    float *dst;
    int dpitch;
    double *dst64;             // may be null: the same values in FP64 (matching guide)
    double gain, offset;
    int clip;
    double lo, hi;
};

// NL-means weighted averaging of the groups of one match launch.
struct NlmArgs {
    int trx, try_;            // tile shape in reference blocks
    int W, H;
    int ngx, ngy, bstep;
    int G;                    // GroupSize: row length of pos / keys
    const int *cnt;
    const int *pos;
    const double *keys;       // distance keys (MSE units), member 0 = the reference block
    int nplanes;
    int src_type;             // 0 float, 1 double, 2 int32 samples
    const void *src[3];
    size_t spitch;            // of src, in elements
    int correction;           // WeightedAverage_Correction
    double exp_mul;           // -1 / strength^2
    double varN;              // (sigma * range / 255)^2
    double fix;               // fixed-point scale of the accumulators (a power of two)
    unsigned long long *num[3];
    int pitch;                // of num
    int FW, FH;               // footprint of a tile: (trx - 1) * bstep + 8 by (try_ - 1) * bstep + 8
};

struct NlmFinalizeArgs {
    int W, H, pitch, bstep, nplanes;
    const unsigned long long *num[3];
    double inv_fix;           // 1 / NlmArgs::fix
    int out_int;              // 0: float outputs; 1: int32 of val * gain + offset (truncated)
    double gain, offset;
    int clip;
    double lo, hi;
    void *out[3];
    size_t out_pitch;
};

// Generic kernels (generic_kernels.hip) for the parameters the tuned kernels
// do not take: BlockSize != 8 or GroupSize > 32.  FP64 distances, transforms
// and shrinkage; one workgroup per group; aggregation by global atomics into
// the same fixed-point planes.
constexpr int kGenMaxBlock = 16;
constexpr int kGenMaxGroup = 64;
constexpr int kGenMaxElems = 8192;  // G * B^2 of one group (two FP64 groups in LDS for Wiener)

struct GenFilterArgs {
    int W, H, pitch;
    int ngx, ngy;
    int B, G;
    const int *cnt, *pos;      // groups from matching, G entries per grid point
    const float *src[3];
    const float *guide[3];     // Wiener guides
    unsigned long long *num[3], *den[3];  // zeroed 64-bit fixed-point sums (kFixScale)
    int active[3];
    double thr[3];             // lambda * sigma_c (orthonormal domain)
    double sig2[3];            // sigma_c^2
    float wmax;                // cap of the Wiener group weight
    const double *cb;          // B x B orthonormal DCT-II [k][n]
    const double *cn;          // C_N [k][n] for N = 1..G, at (N-1) N (2N-1) / 6
    // variant modes (no reference counterpart): haar: C_N of a power-of-two N
    // is the orthonormal Haar matrix and a group keeps the largest power of
    // two of its members; kaiser (B x B, or null): each block's aggregation
    // weight times the window
    int haar;
    const double *kaiser;
};

size_t match_generic_lds_bytes(const MatchArgs &a);
hipError_t launch_match_generic(const MatchArgs &a, bool guide64, hipStream_t s);
size_t filter_generic_lds_bytes(const GenFilterArgs &a, bool wiener);
hipError_t launch_filter_generic(const GenFilterArgs &a, bool wiener, hipStream_t s);

size_t nlm_lds_bytes(const NlmArgs &a);
hipError_t launch_nlm(const NlmArgs &a, hipStream_t s);
hipError_t launch_nlm_finalize(const NlmFinalizeArgs &a, hipStream_t s);

size_t match_lds_bytes(const MatchArgs &a);
size_t filter_lds_bytes(const FilterArgs &a, bool wiener);

// A batched launch: `units` workgroups (or rows), map and args in device memory.
struct BatchLaunch {
    const uint32_t *map;
    const void *args;      // device array of the kernel's argument struct
    unsigned units;
    int maxW;              // row kernels: widest image
};

// guide64: the guide is FP64 (else FP32).  `b` null: a single-image launch.
hipError_t launch_match(const MatchArgs &a, bool guide64, hipStream_t s, const BatchLaunch *b = nullptr);
hipError_t launch_filter(const FilterArgs &a, bool wiener, hipStream_t s, const BatchLaunch *b = nullptr);
// precision: 1 = FP16, 2 = BF16 operands with FP32 accumulation, 3 = FP16
// operands split in high and low halves (FP32-accurate products; needs the
// low-half tables in FilterArgs)
hipError_t launch_filter_mfma(const FilterArgs &a, bool wiener, int precision, hipStream_t s,
                              const BatchLaunch *b = nullptr);
size_t filter_mfma_lds_bytes(const FilterArgs &a, bool wiener);
hipError_t launch_finalize(const FinalizeArgs &a, hipStream_t s, const BatchLaunch *b = nullptr);
// multi-frame groups: merge, and the filters over a window of frame slots
hipError_t launch_stack_merge(const StackMergeArgs &a, hipStream_t s);
size_t filter_stack_lds_bytes(const FilterArgs &a, bool wiener, int precision);
hipError_t launch_filter_stack(const FilterArgs &a, bool wiener, int precision, hipStream_t s);
hipError_t launch_opp(const OppArgs &a, hipStream_t s, const BatchLaunch *b = nullptr);
hipError_t launch_to_float(const ToFloatArgs &a, hipStream_t s);
hipError_t launch_opp_inverse(float *r, float *g, float *b, const float *y, const float *u, const float *v, int W,
                              int H, size_t pitch, int clip, hipStream_t s);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s);
hipError_t launch_copy2d(float *dst, size_t dpitch, const float *src, size_t spitch, int W, int H,
                         hipStream_t s);

}  // namespace ispbm3d
