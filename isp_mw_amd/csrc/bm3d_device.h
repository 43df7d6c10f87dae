// Device helpers shared by the HIP translation units of the BM3D library
// (bm3d_kernels.hip, generic_kernels.hip): wave primitives, the reference
// grid and search window, the 64-bit fixed-point accumulation and the
// wave-level radix selection of matching.  Internal; no reference
// counterpart beyond the functions cited below.
#pragma once

#include "bm3d_internal.h"
#include "tile_order.h"

namespace ispbm3d {
namespace {

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of lanes below this one whose bit is set in mask.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Lane q's value of v, q the same in every lane: v_readlane from registers
// (a shuffle or an LDS array would wait on an LDS round trip per call).
__device__ __forceinline__ uint32_t lane_value(uint32_t v, uint32_t q)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)q);
}
__device__ __forceinline__ uint64_t lane_value(uint64_t v, uint32_t q)
{
    return ((uint64_t)lane_value((uint32_t)(v >> 32), q) << 32) | lane_value((uint32_t)v, q);
}

// The image and tile (or row) a workgroup works on: its own launch's
// arguments and blockIdx, or, in a batched launch, the map entry of this
// workgroup and that image's arguments.
template <bool BATCHED, typename A>
__device__ __forceinline__ const A &tile_unit(const A &a0, const A *__restrict__ bargs,
                                              const uint32_t *__restrict__ bmap, int &tx, int &ty)
{
    if constexpr (BATCHED) {
        const uint32_t u = bmap[blockIdx.x];
        tx = (int)(u & 1023u);
        ty = (int)((u >> 10) & 1023u);
        return bargs[u >> kBatchImageShift];
    } else {
        xcd_tile(tx, ty);
        return a0;
    }
}

template <bool BATCHED, typename A>
__device__ __forceinline__ const A &row_unit(const A &a0, const A *__restrict__ bargs,
                                             const uint32_t *__restrict__ bmap, int &y)
{
    if constexpr (BATCHED) {
        const uint32_t u = bmap[blockIdx.y];
        y = (int)(u & 0xFFFFFu);
        return bargs[u >> kBatchImageShift];
    } else {
        y = (int)blockIdx.y;
        return a0;
    }
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Inclusive prefix sum over the wave's 64 lanes by DPP, without LDS: row
// shifts by 1, 2, 4, 8 (a lane with no source in its row of 16 adds 0) scan
// each row; row_bcast:15 then adds row 0's total to row 1 and row 2's to row
// 3, and row_bcast:31 the total of rows 0-1 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
    return v;
}

// Minimum / maximum over the wave by DPP, in every lane: the row scans of
// wave_incl_scan with min / max (a lane with no source takes the identity),
// then lane 63's value.
__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
    const int id = (int)~0u;
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x111, 0xf, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x112, 0xf, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x114, 0xf, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x118, 0xf, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x142, 0xa, 0xf, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, 0x143, 0xc, 0xf, false));
    return lane_value(v, 63);
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return lane_value(v, 63);
}

// Block::_SearchBoundary (include/Block.h:716-745); range already a multiple of step.
__device__ __forceinline__ int search_boundary(int pos, int boundary, int range, int step)
{
    int sb;
    if (pos == boundary) {
        sb = boundary;
    } else if (pos > boundary) {
        sb = pos - range;
        while (sb < boundary) sb += step;
    } else {
        sb = pos + range;
        while (sb > boundary) sb -= step;
    }
    return sb;
}

// Reference grid (source/BM3D.cpp:98-127): 0, s, 2s, ... with any position
// past n-B replaced by n-B, i.e. position k is min(k*s, n-B).
__device__ __forceinline__ int grid_pos(int k, int s, int last) { return min(k * s, last); }

__host__ __device__ constexpr size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

typedef unsigned long long u64;

// Fixed-point addends of the aggregation sums (LDS integer atomics are full
// rate; two's complement wraps negative values correctly).  The value added
// for x with weight w is x * w * 2^32 rounded to the nearest integer (ties to
// even), formed by one FP64 fused multiply-add: doubles in [2^56, 2^57) are
// spaced 16 apart, so t = fma(x, w * 2^36, 1.5 * 2^56) equals
// 1.5 * 2^56 + 16 * round(x * w * 2^32) while |x * w * 2^32| < 2^51, and the
// bit pattern of t minus that of 1.5 * 2^56 is that integer in two's
// complement.  The constant's low word is zero, so only the high word is
// subtracted: three ALU operations (FP32 -> FP64, FP64 FMA, 32-bit add) where
// an integer and a fraction part take five.
constexpr double kFixMagic = 108086391056891904.0;     // 1.5 * 2^56
constexpr u64 kFixMagicBits = 0x4378000000000000ull;   // its bit pattern

// the FMA operand of weight w: w * 2^32 * 16 (exact: a power-of-two scaling)
__device__ __forceinline__ double fix_weight(float w) { return (double)w * (16.0 * (double)kFixScale); }

__device__ __forceinline__ u64 fix_bits(float x, double wfix)
{
    const double t = __builtin_fma((double)x, wfix, kFixMagic);
    return (u64)__double_as_longlong(t) - kFixMagicBits;
}

// adds round(w * 2^32)
__device__ __forceinline__ void fix_add(u64 *acc, float w) { atomicAdd(acc, fix_bits(1.f, fix_weight(w))); }

// adds round(x * w * 2^32), wfix = fix_weight(w)
__device__ __forceinline__ void fix_add_scaled(u64 *acc, float x, double wfix) { atomicAdd(acc, fix_bits(x, wfix)); }

__host__ __device__ inline size_t match_wave_bytes(int maxCand, int keyBytes)
{
    // keys K[maxCand] + selk K[32] + idxs u32[maxCand] + memb u32[maxCand]
    // + hist u32[256] + seli u32[32]
    return align16((size_t)maxCand * keyBytes + 32 * (size_t)keyBytes + (size_t)maxCand * 8 + 256 * 4 + 32 * 4);
}

// ---------------------------------------------------------------------------
// Block matching.  One workgroup stages the guide around a tile of
// kTileRX x kTileRY reference blocks in LDS as FP64; each wave takes one
// reference block at a time.  Lane strips of kStrip horizontally adjacent
// candidates share loaded rows.  Distances are summed in FP64 in the
// reference's row-major order; accepted candidates (SSE <= thSSE, SSE != 0)
// are compacted into a per-wave LDS list, and the G-1 smallest by
// (key, scan index) are found with an MSB-first radix select on the key bits.
// EXACT: each term rounded as d*d then added, as the reference's FP64 loop
// does (include/Block.h:600-603), so equal-looking distances of quantised
// data tie exactly as they do there; otherwise fused multiply-add.
// DT: the distance type.  double (the plain form: the reference's FP64
// distances, 64-bit keys) or float (the mixed-precision forms: FP32 distances
// and 32-bit keys, half the LDS and half the radix passes).
__device__ __forceinline__ double sq_acc_exact(double d, double acc)
{
#pragma clang fp contract(off)
    const double t = d * d;
    return acc + t;
}

template <bool EXACT>
__device__ __forceinline__ double sq_acc(double d, double acc)
{
    if (EXACT) return sq_acc_exact(d, acc);
    return __builtin_fma(d, d, acc);
}

template <bool EXACT>
__device__ __forceinline__ float sq_acc(float d, float acc)
{
    return fmaf(d, d, acc);
}

// Sort keys of non-negative distances: their IEEE bit patterns.
template <typename DT>
struct KeyOf;
template <>
struct KeyOf<double> {
    typedef uint64_t T;
    static constexpr int BITS = 64;
    static __device__ __forceinline__ T key(double v) { return (uint64_t)__double_as_longlong(v); }
    static __device__ __forceinline__ double value(T k) { return __longlong_as_double((long long)k); }
};
template <>
struct KeyOf<float> {
    typedef uint32_t T;
    static constexpr int BITS = 32;
    static __device__ __forceinline__ T key(float v) { return __float_as_uint(v); }
    static __device__ __forceinline__ double value(T k) { return (double)__uint_as_float(k); }
};

// The K smallest (key, scan index) of the wave's list keys / idxs[0..cnt)
// into selk / seli, unordered; returns how many (min(cnt, K)).  An MSB-first
// radix select on the key bits (256-bin LDS histograms, wave prefix scan);
// ties within the last bucket are ranked exactly by scan index.
template <typename KT, int BITS>
__device__ __forceinline__ uint32_t select_smallest(const KT *keys, const uint32_t *idxs, uint32_t cnt, uint32_t K,
                                                    KT *selk, uint32_t *seli, uint32_t *hist, uint32_t *memb,
                                                    int lane)
{
    uint32_t nsel = 0;
    if (cnt <= K) {
        for (uint32_t e = lane; e < cnt; e += 64) {
            selk[e] = keys[e];
            seli[e] = idxs[e];
        }
        nsel = cnt;
    } else {
        // bits above the highest one in which the list's keys differ are
        // common to all of them: the radix starts below them (the exponent
        // bits of FP32 keys would otherwise put every key in one or two bins,
        // whose atomics serialise)
        KT kmin = ~(KT)0, kmax = 0;
        for (uint32_t e = lane; e < cnt; e += 64) {
            const KT k = keys[e];
            kmin = k < kmin ? k : kmin;
            kmax = k > kmax ? k : kmax;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const KT x = __shfl_xor(kmin, o), y = __shfl_xor(kmax, o);
            kmin = x < kmin ? x : kmin;
            kmax = y > kmax ? y : kmax;
        }
        const KT diff = kmin ^ kmax;
        const int top = diff == 0 ? 0 : (BITS == 64 ? 63 - __clzll((long long)diff) : 31 - __clz((int)diff));
        int shift = top < 8 ? 0 : top - 7;
        KT pmask = shift + 8 >= BITS ? (KT)0 : ~(KT)0 << (shift + 8);
        KT prefix = kmin & pmask;
        uint32_t need = K;
        bool whole = false;
        for (;; shift = shift < 8 ? 0 : shift - 8) {
            for (int q = lane; q < 256; q += 64) hist[q] = 0;
            wave_sync();
            for (uint32_t e = lane; e < cnt; e += 64) {
                const KT k = keys[e];
                if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
            }
            wave_sync();
            const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1];
            const uint32_t h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
            const uint32_t s = h0 + h1 + h2 + h3;
            const uint32_t incl = wave_incl_scan(s, lane);
            const uint32_t excl = incl - s;
            const uint64_t bm = __ballot(excl < need && need <= incl);
            const int L = __ffsll((unsigned long long)bm) - 1;
            uint32_t dig = 0, blw = 0, mm = 0;
            if (lane == L) {
                const uint32_t hh[4] = {h0, h1, h2, h3};
                uint32_t c = excl;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (mm == 0 && c + hh[q] >= need) {
                        dig = 4 * lane + q;
                        blw = c;
                        mm = hh[q];
                    }
                    c += hh[q];
                }
            }
            dig = lane_value(dig, (uint32_t)L);
            blw = lane_value(blw, (uint32_t)L);
            mm = lane_value(mm, (uint32_t)L);
            need -= blw;
            prefix |= (KT)dig << shift;
            pmask |= (KT)0xFF << shift;
            if (mm == need) {
                whole = true;
                break;
            }
            if (mm <= 64 || shift == 0) break;
            wave_sync();
        }
        uint32_t nm = 0;
        for (uint32_t e0 = 0; e0 < cnt; e0 += 64) {
            const uint32_t e = e0 + lane;
            const bool inr = e < cnt;
            const KT k = inr ? keys[e] : ~(KT)0;
            const bool below = inr && (k & pmask) < prefix;
            const bool inb = inr && (k & pmask) == prefix;
            const bool take = below || (whole && inb);
            const uint64_t tm = __ballot(take);
            if (take) {
                const uint32_t sl = nsel + lanes_below(tm);
                selk[sl] = k;
                seli[sl] = idxs[e];
            }
            nsel += (uint32_t)__popcll(tm);
            if (!whole) {
                const uint64_t mb = __ballot(inb);
                if (inb) memb[nm + lanes_below(mb)] = e;
                nm += (uint32_t)__popcll(mb);
            }
        }
        if (!whole) {
            wave_sync();
            for (uint32_t q0 = 0; q0 < nm; q0 += 64) {
                const uint32_t q = q0 + lane;
                const bool v = q < nm;
                KT kq = 0;
                uint32_t iq = 0, rank = 0;
                // (the bucket can hold hundreds of equal keys here, where
                // LDS broadcast reads pipeline better than per-member readlanes)
                if (v) {
                    const uint32_t e = memb[q];
                    kq = keys[e];
                    iq = idxs[e];
                    for (uint32_t r2 = 0; r2 < nm; ++r2) {
                        const uint32_t e2 = memb[r2];
                        const KT k2 = keys[e2];
                        const uint32_t i2 = idxs[e2];
                        rank += (k2 < kq) || (k2 == kq && i2 < iq);
                    }
                }
                const bool take = v && rank < need;
                const uint64_t tm = __ballot(take);
                if (take) {
                    const uint32_t sl = nsel + lanes_below(tm);
                    selk[sl] = kq;
                    seli[sl] = iq;
                }
                nsel += (uint32_t)__popcll(tm);
            }
        }
    }
    wave_sync();
    return nsel;
}

}  // namespace
}  // namespace ispbm3d
