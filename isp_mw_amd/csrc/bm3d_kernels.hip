// HIP kernels of the MI355X BM3D denoiser (gfx950, wave64).
//
// Stages of one BM3D step (source/BM3D.cpp:80-218):
//   match_kernel     - exhaustive block matching in the clipped search window
//                      (include/Block.h:578-751) with FP64 distances, threshold,
//                      and top-(G-1) selection by a wave-level radix select.
//   filter_kernel    - gather, separable orthonormal 3D DCT, hard threshold or
//                      Wiener shrinkage, inverse, weighted aggregation into an
//                      LDS tile of 64-bit fixed-point sums, flushed to HBM with
//                      integer atomics (source/BM3D.cpp:339-385, 411-457;
//                      include/Block.h:1240-1357).  FP32 on the vector ALUs.
//   filter_mfma_kernel - the same stage with the transforms as chained FP16 or
//                      BF16 matrix-core products (FP32 accumulation).
//   to_float_kernel  - RangeConvert of integer samples to float
//                      (include/Conversion.hpp:134-256).
//   finalize_kernel  - num/den (source/BM3D.cpp:132-135, 204-217), optionally
//                      fused with OPP->RGB and clip (include/Conversion.hpp:739-762).
//   opp_kernel       - RGB->OPP without clip (include/Conversion.hpp:547-578).
#include <type_traits>

#include "bm3d_device.h"

namespace ispbm3d {
namespace {


// CROSS: candidates from another frame's plane (a.search) against the
// reference blocks of a.guide, staged in a second LDS tile; the list has no
// self entry.
template <typename GT, typename DT, bool ST1, bool EXACT, bool BATCHED, bool CROSS = false>
__global__ __launch_bounds__(kThreads) void match_kernel(MatchArgs a0, const uint32_t *__restrict__ bmap,
                                                         const MatchArgs *__restrict__ bargs)
{
    typedef typename KeyOf<DT>::T KT;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int tx, ty;
    const MatchArgs &a = tile_unit<BATCHED>(a0, bargs, bmap, tx, ty);
    const GT *__restrict__ guide = static_cast<const GT *>(a.guide);
    constexpr bool cross = CROSS;
    const GT *__restrict__ sguide = cross ? static_cast<const GT *>(a.search) : guide;
    const int gpitch = a.gpitch;
    const int trx = a.trx, tnum = a.trx * a.try_;
    const int tx0 = tx * trx, ty0 = ty * a.try_;
    const int rx0 = grid_pos(tx0, a.bstep, a.W - kBlock), ry0 = grid_pos(ty0, a.bstep, a.H - kBlock);
    const int ox = rx0 - a.range, oy = ry0 - a.range;

    DT *tile = reinterpret_cast<DT *>(smem);
    const size_t tileBytes = align16((size_t)a.tileH * a.tilePitch * sizeof(DT));
    DT *rtile = reinterpret_cast<DT *>(smem + tileBytes);  // cross: the tile's reference blocks
    const size_t rtileBytes = cross ? align16((size_t)a.refH * a.refPitch * sizeof(DT)) : 0;
    unsigned char *wb = smem + tileBytes + rtileBytes + (size_t)wave * match_wave_bytes(a.maxCand, sizeof(KT));
    KT *keys = reinterpret_cast<KT *>(wb);
    KT *selk = keys + a.maxCand;
    uint32_t *idxs = reinterpret_cast<uint32_t *>(selk + 32);
    uint32_t *memb = idxs + a.maxCand;
    uint32_t *hist = memb + a.maxCand;
    uint32_t *seli = hist + 256;

    for (int k = tid; k < a.tileH * a.tileW; k += kThreads) {
        const int ty = k / a.tileW, tx = k - ty * a.tileW;
        const int yy = oy + ty, xx = ox + tx;
        DT v = 0;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) v = (DT)sguide[(size_t)yy * gpitch + xx];
        tile[ty * a.tilePitch + tx] = v;
    }
    if (cross) {
        for (int k = tid; k < a.refH * a.refW; k += kThreads) {
            const int ty = k / a.refW, tx = k - ty * a.refW;
            const int yy = ry0 + ty, xx = rx0 + tx;
            DT v = 0;
            if (yy < a.H && xx < a.W) v = (DT)guide[(size_t)yy * gpitch + xx];
            rtile[ty * a.refPitch + tx] = v;
        }
    }
    __syncthreads();

    for (int r = wave; r < tnum; r += kWaves) {
        const int gxi = tx0 + (r % trx), gyi = ty0 + (r / trx);
        if (gxi >= a.ngx || gyi >= a.ngy) continue;
        const int i = grid_pos(gxi, a.bstep, a.W - kBlock), j = grid_pos(gyi, a.bstep, a.H - kBlock);
        const size_t refIdx = (size_t)gyi * a.ngx + gxi;
        const size_t base = refIdx * a.G;
        if (a.bypass && !cross) {
            if (lane == 0) {
                a.cnt_out[refIdx] = 1;
                a.pos_out[base] = (j << 16) | i;
                if (a.key_out) a.key_out[base] = 0.0;
            }
            continue;
        }
        const int st = a.step;
        const int l = search_boundary(i, 0, a.range, st);
        const int rr = search_boundary(i, a.W - kBlock, a.range, st);
        const int t = search_boundary(j, 0, a.range, st);
        const int b = search_boundary(j, a.H - kBlock, a.range, st);
        const int nx = (rr - l) / st + 1, ny = (b - t) / st + 1;
        const int selfx = (i - l) / st, selfy = (j - t) / st;
        const int nsx = (nx + kStrip - 1) / kStrip, nstrips = nsx * ny;
        const int rp = cross ? a.refPitch : a.tilePitch;
        const DT *tref = cross ? rtile + (j - ry0) * rp + (i - rx0) : tile + (j - oy) * rp + (i - ox);
        const DT thSSE = (DT)a.thSSE, distMul = (DT)a.distMul;

        // The list holds a.maxCand entries (fewer than a large window's
        // candidates): before a batch of strips that could overflow it, it is
        // reduced to its K best, which still contain the window's K best.
        // After such a reduction to K entries, a later candidate whose key is
        // not below the largest of them ranks behind all K (its scan index is
        // larger) and is not listed: kbound.
        const uint32_t K = (uint32_t)(a.G - 1);
        const bool may_fill = (uint32_t)a.maxCand < (uint32_t)(nx * ny);
        uint32_t cnt = 0;
        KT kbound = ~(KT)0;
        for (int s0 = 0; s0 < nstrips; s0 += 64) {
            if (may_fill && cnt + 64 * kStrip > (uint32_t)a.maxCand) {
                wave_sync();
                const uint32_t nk =
                    select_smallest<KT, KeyOf<DT>::BITS>(keys, idxs, cnt, K, selk, seli, hist, memb, lane);
                KT kx = 0;
                for (uint32_t e = lane; e < nk; e += 64) {
                    keys[e] = selk[e];
                    idxs[e] = seli[e];
                    kx = selk[e] > kx ? selk[e] : kx;
                }
                if (nk == K) {
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        const KT y = __shfl_xor(kx, o);
                        kx = y > kx ? y : kx;
                    }
                    kbound = kx;
                }
                cnt = nk;
                wave_sync();
            }
            const int sidx = s0 + lane;
            const bool valid = sidx < nstrips;
            const int ky = valid ? sidx / nsx : 0;
            const int kx0 = valid ? (sidx - ky * nsx) * kStrip : 0;
            DT acc[kStrip];
#pragma unroll
            for (int c = 0; c < kStrip; ++c) acc[c] = 0;
            if (valid) {
                const DT *tc = tile + (t - oy + ky * st) * a.tilePitch + (l - ox + kx0 * st);
#pragma unroll
                for (int u = 0; u < kBlock; ++u) {
                    DT rv[kBlock];
#pragma unroll
                    for (int v = 0; v < kBlock; ++v) rv[v] = tref[u * rp + v];
                    // EXACT: one sum over the 64 terms in row-major order; FP32:
                    // each row's 8 fused terms summed from 0, then the row sums
                    // added in row order (the order of match_walk_kernel)
                    if constexpr (ST1) {
                        DT cv[kBlock + kStrip - 1];
#pragma unroll
                        for (int q = 0; q < kBlock + kStrip - 1; ++q) cv[q] = tc[u * a.tilePitch + q];
#pragma unroll
                        for (int c = 0; c < kStrip; ++c) {
                            DT rs = 0;
#pragma unroll
                            for (int v = 0; v < kBlock; ++v) {
                                const DT d = rv[v] - cv[c + v];
                                if constexpr (EXACT) acc[c] = sq_acc<true>(d, acc[c]);
                                else rs = sq_acc<false>(d, rs);
                            }
                            if constexpr (!EXACT) acc[c] += rs;
                        }
                    } else {
#pragma unroll
                        for (int c = 0; c < kStrip; ++c) {
                            DT rs = 0;
#pragma unroll
                            for (int v = 0; v < kBlock; ++v) {
                                const DT d = rv[v] - tc[u * a.tilePitch + c * st + v];
                                if constexpr (EXACT) acc[c] = sq_acc<true>(d, acc[c]);
                                else rs = sq_acc<false>(d, rs);
                            }
                            if constexpr (!EXACT) acc[c] += rs;
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < kStrip; ++c) {
                const int kx = kx0 + c;
                const KT key = KeyOf<DT>::key(acc[c] * distMul);
                const bool ok = valid && kx < nx && (cross || !(kx == selfx && ky == selfy)) && acc[c] <= thSSE &&
                                acc[c] != 0 && key < kbound;
                const uint64_t m = __ballot(ok);
                if (ok) {
                    const uint32_t slot = cnt + lanes_below(m);
                    keys[slot] = key;
                    idxs[slot] = (uint32_t)(ky * nx + kx);
                }
                cnt += (uint32_t)__popcll(m);
            }
        }
        wave_sync();

        // ---- selection of the K = G-1 smallest (key, scan index) ----
        const uint32_t nsel =
            select_smallest<KT, KeyOf<DT>::BITS>(keys, idxs, cnt, K, selk, seli, hist, memb, lane);

        // ---- order the selection and write the group ----
        // (a cross-frame list: the selection alone, without the self entry)
        const int first = cross ? 0 : 1;
        if (lane == 0) {
            a.cnt_out[refIdx] = (int)(first + nsel);
            if (!cross) {
                a.pos_out[base] = (j << 16) | i;
                if (a.key_out) a.key_out[base] = 0.0;
            }
        }
        // each lane holds one selected entry and counts the entries before
        // it: from the other lanes' registers by lane_value at BMstep 1, from
        // LDS otherwise.  Measured (same build, A/B): at BMstep 1 (the plain
        // form's FP64 matching at 1080p) lane_value took 1.46 + 1.16 ->
        // 1.42 + 1.12 ms; at BMstep 3 (NL-means' exact form) it made
        // nlm_match 1.13 -> 1.49 ms, a difference not traced to the loop
        KT kq = 0;
        uint32_t iq = 0, rank = 0;
        if ((uint32_t)lane < nsel) {
            kq = selk[lane];
            iq = seli[lane];
        }
        if constexpr (ST1) {
            // a uniform trip count and no short-circuit: no divergent loop
            const uint32_t nq = (uint32_t)__builtin_amdgcn_readfirstlane((int)nsel);
            for (uint32_t q = 0; q < nq; ++q) {
                const KT k2 = lane_value(kq, q);
                const uint32_t i2 = lane_value(iq, q);
                rank += (uint32_t)((k2 < kq) | ((k2 == kq) & (i2 < iq)));
            }
        } else if ((uint32_t)lane < nsel) {
            for (uint32_t q = 0; q < nsel; ++q) {
                const KT k2 = selk[q];
                rank += (k2 < kq) || (k2 == kq && seli[q] < iq);
            }
        }
        if ((uint32_t)lane < nsel) {
            const int ky = (int)(iq / (uint32_t)nx), kx = (int)(iq - (uint32_t)ky * nx);
            const int y = t + ky * st, x = l + kx * st;
            a.pos_out[base + first + rank] = (y << 16) | x;
            if (a.key_out) a.key_out[base + first + rank] = KeyOf<DT>::value(kq);
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Block matching by column walks: FP32 distances, the single-frame search,
// BlockStep 3..7, and windows of at most 2 x 64 lane strips.  One wave walks
// down one grid column of the tile.  The windows of one column share their
// lattice of offsets (dy, dx) (the range is a multiple of BMstep; each
// block's own clipping decides which offsets are candidates).  Every lane
// owns RND strips of STRIP offsets for the whole walk (two of 5, or one of
// 6 where 64 such strips cover the window).  Each image row's
// 8-term sum of squared differences is formed once per offset and added to
// every block of the column that covers the row (held in a ring of
// kWalkSlots register slots).  A block's distance
// is then the sum of its eight row sums, added in row order from 0: the same
// wherever the block lies in a walk, and the same as in match_kernel.  A
// block is selected as soon as its last row is in; its slot then takes the
// block kWalkSlots further down.  At BlockStep 3 a column of 16 blocks forms
// 53 row sums for 128 uses.  The keys stay in registers; the selection reads
// them there.
// Register slots of the ring: the blocks a row can touch.  Block k's rows
// end at j_k + 7; block k + P starts at j_k + (P-1) s + 1 or later (the last
// grid position may be clamped), so P slots suffice where (P-1) s >= 7:
// P = 4 at BlockStep 3, P = 3 at BlockStep 4..7.
constexpr uint32_t kNoKey = ~0u;

// per wave: hist u32[256], memk u32[64], memi u32[64], selk u32[32], seli u32[32]
__host__ __device__ constexpr size_t walk_wave_bytes() { return (256 + 64 + 64 + 32 + 32) * 4; }
// per wave, MULTI only: [try_][2][32] kept pairs and [try_] counts
__host__ __device__ inline size_t walk_prior_bytes(int try_) { return align16((size_t)try_ * 65 * 4); }

// The K smallest (key, index) of the wave's register keys into selk / seli,
// unordered; returns how many (min(candidates, K)).  key[r][c] is kNoKey for
// no candidate; its index is ibase[r] + c, and indices grow in (r, lane, c)
// order.  The radix select of select_smallest with the histogram passes over
// registers; the last bucket's members (at most 64) are ranked exactly in
// LDS, or, when more of them share one key, taken in index order.
template <int RND, int STRIP>
__device__ __forceinline__ uint32_t select_walk(const uint32_t (&key)[RND][STRIP],
                                                const int (&ibase)[RND], uint32_t K, uint32_t *selk,
                                                uint32_t *seli, uint32_t *hist, uint32_t *memk, uint32_t *memi,
                                                int lane)
{
    int n = 0;
    uint32_t kmin = kNoKey, kmax = 0;
#pragma unroll
    for (int r = 0; r < RND; ++r)
#pragma unroll
        for (int c = 0; c < STRIP; ++c) {
            const uint32_t k = key[r][c];
            if (k != kNoKey) {
                ++n;
                kmin = k < kmin ? k : kmin;
                kmax = k > kmax ? k : kmax;
            }
        }
    const uint32_t total = lane_value(wave_incl_scan((uint32_t)n, lane), 63);
    uint32_t nsel = 0;
    if (total <= K) {
#pragma unroll
        for (int r = 0; r < RND; ++r)
#pragma unroll
            for (int c = 0; c < STRIP; ++c) {
                const bool take = key[r][c] != kNoKey;
                const uint64_t tm = __ballot(take);
                if (take) {
                    const uint32_t sl = nsel + lanes_below(tm);
                    selk[sl] = key[r][c];
                    seli[sl] = (uint32_t)(ibase[r] + c);
                }
                nsel += (uint32_t)__popcll(tm);
            }
        wave_sync();
        return nsel;
    }
    kmin = wave_min(kmin);
    kmax = wave_max(kmax);
    const uint32_t diff = kmin ^ kmax;
    const int top = diff == 0 ? 0 : 31 - __clz((int)diff);
    int shift = top < 8 ? 0 : top - 7;
    uint32_t pmask = shift + 8 >= 32 ? 0u : ~0u << (shift + 8);
    uint32_t prefix = kmin & pmask;
    uint32_t need = K, mm = 0;
    bool whole = false;
    for (;; shift = shift < 8 ? 0 : shift - 8) {
        for (int q = lane; q < 256; q += 64) hist[q] = 0;
        wave_sync();
#pragma unroll
        for (int r = 0; r < RND; ++r)
#pragma unroll
            for (int c = 0; c < STRIP; ++c) {
                const uint32_t k = key[r][c];
                if (k != kNoKey && (k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
            }
        wave_sync();
        const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1];
        const uint32_t h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
        const uint32_t s = h0 + h1 + h2 + h3;
        const uint32_t incl = wave_incl_scan(s, lane);
        const uint32_t excl = incl - s;
        const uint64_t bm = __ballot(excl < need && need <= incl);
        const int L = __ffsll((unsigned long long)bm) - 1;
        uint32_t dig = 0, blw = 0;
        mm = 0;
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
        prefix |= dig << shift;
        pmask |= 0xFFu << shift;
        if (mm == need) {
            whole = true;
            break;
        }
        if (mm <= 64 || shift == 0) break;
        wave_sync();
    }
    // below the bucket (and the bucket itself when all of it is taken)
    const bool rank = !whole && mm <= 64;
    uint32_t nm = 0;
#pragma unroll
    for (int r = 0; r < RND; ++r)
#pragma unroll
        for (int c = 0; c < STRIP; ++c) {
            const uint32_t k = key[r][c];
            const bool inr = k != kNoKey;
            const bool below = inr && (k & pmask) < prefix;
            const bool inb = inr && (k & pmask) == prefix;
            const bool take = below || (whole && inb);
            const uint64_t tm = __ballot(take);
            if (take) {
                const uint32_t sl = nsel + lanes_below(tm);
                selk[sl] = k;
                seli[sl] = (uint32_t)(ibase[r] + c);
            }
            nsel += (uint32_t)__popcll(tm);
            if (rank) {
                const uint64_t mb = __ballot(inb);
                if (inb) {
                    const uint32_t q = nm + lanes_below(mb);
                    memk[q] = k;
                    memi[q] = (uint32_t)(ibase[r] + c);
                }
                nm += (uint32_t)__popcll(mb);
            }
        }
    if (rank) {
        // at most 64 members, one per lane: exact rank by (key, index)
        wave_sync();
        const bool v = (uint32_t)lane < nm;
        uint32_t kq = 0, iq = 0, rk = 0;
        if (v) {
            kq = memk[lane];
            iq = memi[lane];
        }
        for (uint32_t q = 0; q < nm; ++q) {
            const uint32_t k2 = lane_value(kq, q), i2 = lane_value(iq, q);
            rk += (uint32_t)((k2 < kq) | ((k2 == kq) & (i2 < iq)));
        }
        const bool take = v && rk < need;
        const uint64_t tm = __ballot(take);
        if (take) {
            const uint32_t sl = nsel + lanes_below(tm);
            selk[sl] = kq;
            seli[sl] = iq;
        }
        nsel += (uint32_t)__popcll(tm);
    } else if (!whole) {
        // more than 64 members, every key bit decided: one key, so the
        // `need` smallest indices, which come first in (r, lane, c) order
        uint32_t before = 0;
#pragma unroll
        for (int r = 0; r < RND; ++r) {
            uint32_t cl = 0;
#pragma unroll
            for (int c = 0; c < STRIP; ++c) cl += key[r][c] != kNoKey && (key[r][c] & pmask) == prefix;
            const uint32_t incl = wave_incl_scan(cl, lane);
            uint32_t p = before + incl - cl;
#pragma unroll
            for (int c = 0; c < STRIP; ++c) {
                const uint32_t k = key[r][c];
                if (k != kNoKey && (k & pmask) == prefix) {
                    if (p < need) {
                        selk[nsel + p] = k;
                        seli[nsel + p] = (uint32_t)(ibase[r] + c);
                    }
                    ++p;
                }
            }
            before += lane_value(incl, 63);
        }
        nsel += need;
    }
    wave_sync();
    return nsel;
}

// 4 waves per SIMD: at most 128 registers
// RND rounds of 64 lane strips of STRIP offsets: (2, 5) or, where 64 strips
// of 6 cover the window, (1, 6).  MULTI: a window of more than 128 strips,
// walked in chunks of 128; after each chunk a block keeps its K best
// (key, index), in (key, index) order, in LDS, and the next chunk's selection
// takes them as candidates ahead of its own (earlier chunks hold the smaller
// scan indices, so the order of ties is kept).
template <typename GT, bool ST1, bool BATCHED, int kWalkSlots, int RND, int STRIP, bool MULTI>
__global__ __launch_bounds__(kThreads, 4) void match_walk_kernel(MatchArgs a0, const uint32_t *__restrict__ bmap,
                                                              const MatchArgs *__restrict__ bargs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int tx, ty;
    const MatchArgs &a = tile_unit<BATCHED>(a0, bargs, bmap, tx, ty);
    const GT *__restrict__ guide = static_cast<const GT *>(a.guide);
    const int gpitch = a.gpitch, pitch = a.tilePitch;
    const int trx = a.trx, try_ = a.try_;
    const int tx0 = tx * trx, ty0 = ty * try_;
    const int rx0 = grid_pos(tx0, a.bstep, a.W - kBlock), ry0 = grid_pos(ty0, a.bstep, a.H - kBlock);
    const int R = a.range, st = a.step;
    const int ox = rx0 - R, oy = ry0 - R;

    float *tile = reinterpret_cast<float *>(smem);
    const size_t tileBytes = align16((size_t)a.tileH * pitch * sizeof(float));
    uint32_t *hist = reinterpret_cast<uint32_t *>(smem + tileBytes + (size_t)wave * walk_wave_bytes());
    uint32_t *memk = hist + 256, *memi = memk + 64, *selk = memi + 64, *seli = selk + 32;
    // MULTI: each block's kept (key, index) pairs [try_][2][32] and their counts [try_]
    uint32_t *prior = reinterpret_cast<uint32_t *>(smem + tileBytes + (size_t)kWaves * walk_wave_bytes() +
                                                   (size_t)wave * walk_prior_bytes(try_));
    uint32_t *pcnt = prior + (size_t)try_ * 64;

    for (int k = tid; k < a.tileH * a.tileW; k += kThreads) {
        const int yy0 = k / a.tileW, xx0 = k - yy0 * a.tileW;
        const int yy = oy + yy0, xx = ox + xx0;
        float v = 0.f;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) v = (float)guide[(size_t)yy * gpitch + xx];
        tile[yy0 * pitch + xx0] = v;
    }
    __syncthreads();

    // this lane's strips: offsets (dy[r], dx[r] + c * st), c < STRIP; a strip
    // past the lattice reads offsets -R (inside the tile) and keeps no key
    // lattice indices: offset (dy, dx) = (-R + my * st, -R + mx * st), my, mx < NM
    const int NM = 2 * R / st + 1, Rs = R / st;
    const int nsx = (NM + STRIP - 1) / STRIP, nstrips = NM * nsx;
    const int nchunks = MULTI ? (nstrips + RND * 64 - 1) / (RND * 64) : 1;
    const float thSSE = (float)a.thSSE, distMul = (float)a.distMul;
    const uint32_t K = (uint32_t)(a.G - 1);
    const int nk = min(try_, a.ngy - ty0);  // blocks of each column in this tile

    for (int col = wave; col < trx; col += kWaves)
    for (int h = 0; h < nchunks; ++h) {
        const int gx = tx0 + col;
        if (gx >= a.ngx || nk <= 0) continue;
        if (MULTI && a.bypass && h > 0) continue;
        // this chunk's strips of this lane
        int dy[RND], dx[RND], my[RND], mx0[RND];
#pragma unroll
        for (int r = 0; r < RND; ++r) {
            const int sidx = (h * RND + r) * 64 + lane;
            const bool sv = sidx < nstrips;
            // a strip past the lattice gets my = -1: never inside a window
            my[r] = sv ? sidx / nsx : -1;
            mx0[r] = sv ? (sidx - my[r] * nsx) * STRIP : 0;
            dy[r] = sv ? -R + my[r] * st : -R;
            dx[r] = -R + mx0[r] * st;
        }
        const int i = grid_pos(gx, a.bstep, a.W - kBlock);
        if (a.bypass) {
            for (int k = lane; k < nk; k += 64) {
                const int j = grid_pos(ty0 + k, a.bstep, a.H - kBlock);
                const size_t refIdx = (size_t)(ty0 + k) * a.ngx + gx;
                a.cnt_out[refIdx] = 1;
                a.pos_out[refIdx * a.G] = (j << 16) | i;
                if (a.key_out) a.key_out[refIdx * a.G] = 0.0;
            }
            continue;
        }
        // The column's window spans lattice columns [m0x, m0x + nx); each
        // block's spans rows [m0y, m0y + ny).  Scan index = (my - m0y) * nx + mx - m0x.
        const int l = search_boundary(i, 0, R, st), rr = search_boundary(i, a.W - kBlock, R, st);
        const int nx = (rr - l) / st + 1, m0x = (l - i + R) / st;
        uint32_t vxm[RND];  // strip positions inside the column's window, as bits
#pragma unroll
        for (int r = 0; r < RND; ++r) {
            vxm[r] = 0;
#pragma unroll
            for (int c = 0; c < STRIP; ++c)
                vxm[r] |= (uint32_t)((unsigned)(mx0[r] + c - m0x) < (unsigned)nx) << c;
            // the block itself (offset 0, 0) is no candidate
            const int sc = Rs - mx0[r];
            if (my[r] == Rs && sc >= 0 && sc < STRIP) vxm[r] &= ~(1u << sc);
        }

        float D[RND][kWalkSlots][STRIP];
#pragma unroll
        for (int r = 0; r < RND; ++r)
#pragma unroll
            for (int q = 0; q < kWalkSlots; ++q)
#pragma unroll
                for (int c = 0; c < STRIP; ++c) D[r][q][c] = 0.f;
        const float *tref = tile + (i - ox);
        int y = grid_pos(ty0, a.bstep, a.H - kBlock);
        for (int m = 0; m < nk; m += kWalkSlots) {
#pragma unroll
            for (int ph = 0; ph < kWalkSlots; ++ph) {
                const int k = m + ph;
                if (k >= nk) break;
                // blocks k .. k + kWalkSlots - 1 (slot (ph + d) % kWalkSlots)
                int jd[kWalkSlots];
#pragma unroll
                for (int d = 0; d < kWalkSlots; ++d)
                    jd[d] = k + d < nk ? grid_pos(ty0 + k + d, a.bstep, a.H - kBlock) : 0x3FFFFFFF;
                const int j = jd[0];

                // ---- the rows up to block k's last, added to the blocks covering them ----
                for (; y < j + kBlock; ++y) {
                    float rv[kBlock];
#pragma unroll
                    for (int v = 0; v < kBlock; ++v) rv[v] = tref[(y - oy) * pitch + v];
#pragma unroll
                    for (int r = 0; r < RND; ++r) {
                        const float *tc = tile + (y + dy[r] - oy) * pitch + (i + dx[r] - ox);
                        float rs[STRIP];
                        if constexpr (ST1) {
                            float cv[kBlock + STRIP - 1];
#pragma unroll
                            for (int q = 0; q < kBlock + STRIP - 1; ++q) cv[q] = tc[q];
#pragma unroll
                            for (int c = 0; c < STRIP; ++c) {
                                rs[c] = 0.f;
#pragma unroll
                                for (int v = 0; v < kBlock; ++v) {
                                    const float dd = rv[v] - cv[c + v];
                                    rs[c] = fmaf(dd, dd, rs[c]);
                                }
                            }
                        } else {
#pragma unroll
                            for (int c = 0; c < STRIP; ++c) {
                                rs[c] = 0.f;
#pragma unroll
                                for (int v = 0; v < kBlock; ++v) {
                                    const float dd = rv[v] - tc[c * st + v];
                                    rs[c] = fmaf(dd, dd, rs[c]);
                                }
                            }
                        }
#pragma unroll
                        for (int d = 0; d < kWalkSlots; ++d)
                            if (y >= jd[d] && y < jd[d] + kBlock) {
#pragma unroll
                                for (int c = 0; c < STRIP; ++c)
                                    D[r][(ph + d) % kWalkSlots][c] += rs[c];
                            }
                    }
                }

                // ---- block k is complete: keys, selection of the K = G-1 smallest, the group ----
                const int t = search_boundary(j, 0, R, st), b = search_boundary(j, a.H - kBlock, R, st);
                const int ny = (b - t) / st + 1, m0y = (t - j + R) / st;
                // MULTI: round 0 holds the block's kept pairs from earlier chunks
                constexpr int R0 = MULTI ? 1 : 0;
                uint32_t key[RND + R0][STRIP];
                int ibase[RND + R0];
                if constexpr (MULTI) {
                    const uint32_t np = h > 0 ? pcnt[k] : 0u;
                    key[0][0] = (uint32_t)lane < np ? prior[k * 64 + lane] : kNoKey;
                    ibase[0] = (uint32_t)lane < np ? (int)prior[k * 64 + 32 + lane] : 0;
#pragma unroll
                    for (int c = 1; c < STRIP; ++c) key[0][c] = kNoKey;
                }
#pragma unroll
                for (int r = 0; r < RND; ++r) {
                    const uint32_t vm = (unsigned)(my[r] - m0y) < (unsigned)ny ? vxm[r] : 0u;
                    ibase[R0 + r] = (my[r] - m0y) * nx + mx0[r] - m0x;
#pragma unroll
                    for (int c = 0; c < STRIP; ++c) {
                        const float dd = D[r][ph][c];
                        const bool ok = ((vm >> c) & 1u) && dd <= thSSE && dd != 0.f;
                        key[R0 + r][c] = ok ? __float_as_uint(dd * distMul) : kNoKey;
                        D[r][ph][c] = 0.f;  // the slot takes block k + kWalkSlots
                    }
                }
                const uint32_t nsel =
                    select_walk<RND + R0, STRIP>(key, ibase, K, selk, seli, hist, memk, memi, lane);

                const size_t refIdx = (size_t)(ty0 + k) * a.ngx + gx;
                const size_t base = refIdx * a.G;
                // order of the selection: rank by (key, index) as one 48-bit value
                // (indices are below 2^16: at most 640 offsets per chunk, and
                // MULTI windows are below 2^16 offsets)
                const uint64_t vq = (uint32_t)lane < nsel ? ((uint64_t)selk[lane] << 16) | seli[lane] : ~0ull;
                uint32_t rank = 0;
                for (uint32_t q = 0; q < nsel; ++q) rank += lane_value(vq, q) < vq;
                if (MULTI && h + 1 < nchunks) {
                    // kept for the next chunk, in (key, index) order
                    if ((uint32_t)lane < nsel) {
                        prior[k * 64 + rank] = (uint32_t)(vq >> 16);
                        prior[k * 64 + 32 + rank] = (uint32_t)vq & 0xFFFFu;
                    }
                    if (lane == 0) pcnt[k] = nsel;
                    wave_sync();
                    continue;
                }
                if (lane == 0) {
                    a.cnt_out[refIdx] = (int)(1 + nsel);
                    a.pos_out[base] = (j << 16) | i;
                    if (a.key_out) a.key_out[base] = 0.0;
                }
                if ((uint32_t)lane < nsel) {
                    const uint32_t iq = (uint32_t)vq & 0xFFFFu;
                    const int cy = (int)(iq / (uint32_t)nx), cx = (int)(iq - (uint32_t)cy * nx);
                    a.pos_out[base + 1 + rank] = ((t + cy * st) << 16) | (l + cx * st);
                    if (a.key_out) a.key_out[base + 1 + rank] = (double)__uint_as_float((uint32_t)(vq >> 16));
                }
                wave_sync();
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Collaborative filtering, FP32.  The unnormalised FFTW transforms of the
// reference are replaced by orthonormal DCT-II matrices; with them the
// reference's thresholds become a uniform lambda*sigma and its Wiener noise
// sigma^2*finalAMP becomes sigma^2/2^zeta (zeta = number of zero indices), and
// the reconstruction needs no finalAMP division.
//
// A group is staged per wave in LDS as rows of kRow = 9 floats (element
// (z, u, v) at z*ZS + u*kRow + v): the odd row stride keeps the row pass, the
// column pass and the per-pixel group-axis pass free of bank conflicts.
constexpr int kRow = kBlock + 1;
constexpr int ZS = kBlock * kRow;  // 72 floats per block

__device__ __forceinline__ int gidx(int z, int pix) { return z * ZS + (pix >> 3) * kRow + (pix & 7); }

// Read-only tables seen through the constant address space: the kernel never
// writes them, so wave-uniform reads become scalar loads (through a generic
// pointer, the flush's global atomics count as possible writes and every
// coefficient read stays a vector load per lane).
typedef const __attribute__((address_space(4))) float cfloat;
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *kconst(const T *p)
{
    return (const __attribute__((address_space(4))) T *)p;
}

__device__ __forceinline__ void rows_pass(float *g, int N, cfloat *__restrict__ c8, int lane,
                                          bool inverse)
{
    for (int r = lane; r < N * kBlock; r += 64) {
        float *p = g + r * kRow;  // row (z, u) = r
        float in[kBlock];
#pragma unroll
        for (int v = 0; v < kBlock; ++v) in[v] = p[v];
#pragma unroll
        for (int k = 0; k < kBlock; ++k) {
            float s = 0.f;
#pragma unroll
            for (int v = 0; v < kBlock; ++v)
                s = fmaf(inverse ? c8[v * kBlock + k] : c8[k * kBlock + v], in[v], s);
            p[k] = s;
        }
    }
}

__device__ __forceinline__ void cols_pass(float *g, int N, cfloat *__restrict__ c8, int lane,
                                          bool inverse)
{
    for (int r = lane; r < N * kBlock; r += 64) {
        float *p = g + (r >> 3) * ZS + (r & 7);  // column (z, v)
        float in[kBlock];
#pragma unroll
        for (int u = 0; u < kBlock; ++u) in[u] = p[u * kRow];
#pragma unroll
        for (int k = 0; k < kBlock; ++k) {
            float s = 0.f;
#pragma unroll
            for (int u = 0; u < kBlock; ++u)
                s = fmaf(inverse ? c8[u * kBlock + k] : c8[k * kBlock + u], in[u], s);
            p[k * kRow] = s;
        }
    }
}

// Group-axis DCT of the lane's pixel column, in place in LDS: the column is
// held in registers and each output row is written back as soon as it is
// formed.  inverse applies the transpose.
template <int NP>
__device__ __forceinline__ void zpass(float *g, int N, int lane, cfloat *__restrict__ cn, bool inverse)
{
    float x[NP];
    const int o = gidx(0, lane);
#pragma unroll
    for (int z = 0; z < NP; ++z) x[z] = z < N ? g[o + z * ZS] : 0.f;
#pragma unroll 1
    for (int k = 0; k < N; ++k) {
        float s = 0.f;
#pragma unroll
        for (int z = 0; z < NP; ++z) s = fmaf(inverse ? cn[z * NP + k] : cn[k * NP + z], x[z], s);
        g[o + k * ZS] = s;
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One 3D orthonormal DCT coefficient (k, u, v) of a group, in FP64 from the
// original samples, computed by the whole wave (lane = pixel (y, x)):
// sum_{z,y,x} CN[k][z] C8[u][y] C8[v][x] X[z][y][x].  Result in every lane.
__device__ __noinline__ double exact_coef_wave(int k, int u, int v, int N, int lane,
                                               const int *__restrict__ pos, const float *__restrict__ src,
                                               int pitch, const double *__restrict__ c8d,
                                               const double *__restrict__ cnd)
{
    const int y = lane >> 3, x = lane & 7;
    double s = 0.0;
    for (int z = 0; z < N; ++z) {
        const int pz = pos[z];
        s = fma(cnd[k * N + z], (double)src[(size_t)((pz >> 16) + y) * pitch + (pz & 0xFFFF) + x], s);
    }
    return wave_sum(s * c8d[u * kBlock + y] * c8d[v * kBlock + x]);
}

// The same for multi-frame groups: member z's plane is src + slot[z] * fstride.
__device__ __noinline__ double exact_coef_wave_slots(int k, int u, int v, int N, int lane,
                                                     const int *__restrict__ pos, const float *__restrict__ src,
                                                     int pitch, const double *__restrict__ c8d,
                                                     const double *__restrict__ cnd, const int *__restrict__ slot,
                                                     size_t fstride)
{
    const int y = lane >> 3, x = lane & 7;
    double s = 0.0;
    for (int z = 0; z < N; ++z) {
        const int pz = pos[z];
        const float *pl = src + (size_t)slot[z] * fstride;
        s = fma(cnd[k * N + z], (double)pl[(size_t)((pz >> 16) + y) * pitch + (pz & 0xFFFF) + x], s);
    }
    return wave_sum(s * c8d[u * kBlock + y] * c8d[v * kBlock + x]);
}

struct ExactCtx {
    const int *pos;
    const float *src;
    int pitch;
    const double *c8d;
    const double *cnd;  // N x N for this group
    double thr64;
    float band;
    float wmax;         // cap of the Wiener group weight (FilterArgs::wmax)
    const int *slot;    // multi-frame groups: member frame slots (else null)
    size_t fstride;
};

// Members are read 8 at a time: lane z holds member z's position (and frame
// slot), each load's address comes from it by readlane, and the 8 loads are
// in flight together before their LDS stores (one dependent load per member,
// waited on one at a time, left the plain filter waiting on memory).
constexpr int kGatherBatch = 8;

__device__ __forceinline__ void gather(float *g, const float *__restrict__ plane, int pitch,
                                       const int *__restrict__ pos, int N, int lane)
{
    const int u = lane >> 3, v = lane & 7;
    const int o = gidx(0, lane);
    const int pl = lane < N ? pos[lane] : 0;
    for (int z0 = 0; z0 < N; z0 += kGatherBatch) {
        float t[kGatherBatch];
#pragma unroll
        for (int j = 0; j < kGatherBatch; ++j) {
            const int pz = __builtin_amdgcn_readlane(pl, min(z0 + j, N - 1));
            t[j] = plane[(size_t)((pz >> 16) + u) * pitch + (pz & 0xFFFF) + v];
        }
#pragma unroll
        for (int j = 0; j < kGatherBatch; ++j)
            if (z0 + j < N) g[o + (z0 + j) * ZS] = t[j];
    }
}

// multi-frame groups: member z's plane is plane + slot[z] * fstride
__device__ __forceinline__ void gather_slots(float *g, const float *__restrict__ plane, int pitch,
                                             const int *__restrict__ pos, int N, int lane,
                                             const int *__restrict__ slot, size_t fstride)
{
    const int u = lane >> 3, v = lane & 7;
    const int o = gidx(0, lane);
    const int pl = lane < N ? pos[lane] : 0, sl = lane < N ? slot[lane] : 0;
    for (int z0 = 0; z0 < N; z0 += kGatherBatch) {
        float t[kGatherBatch];
#pragma unroll
        for (int j = 0; j < kGatherBatch; ++j) {
            const int z = min(z0 + j, N - 1);
            const int pz = __builtin_amdgcn_readlane(pl, z);
            const size_t base = (size_t)__builtin_amdgcn_readlane(sl, z) * fstride;
            t[j] = plane[base + (size_t)((pz >> 16) + u) * pitch + (pz & 0xFFFF) + v];
        }
#pragma unroll
        for (int j = 0; j < kGatherBatch; ++j)
            if (z0 + j < N) g[o + (z0 + j) * ZS] = t[j];
    }
}

// Forward transform, shrinkage and inverse of one group of one plane; the
// filtered blocks are left in g.  Returns the group's aggregation weight.
template <int NP, bool WIENER, bool SLOTS = false>
__device__ __forceinline__ float group_core(float *g, int N, int lane, const float *__restrict__ c8g,
                                            const float *__restrict__ cng, const float *__restrict__ src,
                                            const float *__restrict__ guide, int pitch, float thr, float sig2,
                                            const ExactCtx &ex)
{
    cfloat *c8 = kconst(c8g), *cn = kconst(cng);
    const int o = gidx(0, lane);
    float yr[WIENER ? NP : 1];
    if (WIENER) {
        // the guide group first; its coefficients stay in registers
        if constexpr (SLOTS) gather_slots(g, guide, pitch, ex.pos, N, lane, ex.slot, ex.fstride);
        else gather(g, guide, pitch, ex.pos, N, lane);
        wave_sync();
        rows_pass(g, N, c8, lane, false);
        wave_sync();
        cols_pass(g, N, c8, lane, false);
        wave_sync();
        zpass<NP>(g, N, lane, cn, false);
#pragma unroll
        for (int k = 0; k < NP; ++k) yr[k] = k < N ? g[o + k * ZS] : 0.f;
        wave_sync();
    }
    if constexpr (SLOTS) gather_slots(g, src, pitch, ex.pos, N, lane, ex.slot, ex.fstride);
    else gather(g, src, pitch, ex.pos, N, lane);
    wave_sync();
    rows_pass(g, N, c8, lane, false);
    wave_sync();
    cols_pass(g, N, c8, lane, false);
    wave_sync();
    zpass<NP>(g, N, lane, cn, false);

    // shrinkage on the lane's own coefficients (no cross-lane dependency)
    float weight;
    if (!WIENER) {
        int kept = 0;
        for (int k = 0; k < N; ++k) {
            const float ay = fabsf(g[o + k * ZS]);
            bool zero = ay <= thr;
            // decisions too close to call in FP32 are re-made in FP64
            uint64_t close = __ballot(fabsf(ay - thr) <= ex.band);
            while (close) {
                const int L = __ffsll((unsigned long long)close) - 1;
                close &= close - 1;
                const double oc = SLOTS ? exact_coef_wave_slots(k, L >> 3, L & 7, N, lane, ex.pos, ex.src, ex.pitch,
                                                                ex.c8d, ex.cnd, ex.slot, ex.fstride)
                                        : exact_coef_wave(k, L >> 3, L & 7, N, lane, ex.pos, ex.src, ex.pitch,
                                                          ex.c8d, ex.cnd);
                if (lane == L) zero = fabs(oc) <= ex.thr64;
            }
            if (zero) g[o + k * ZS] = 0.f;
            else ++kept;
        }
        kept = wave_sum(kept);
        weight = kept < 1 ? 1.f : 1.f / (float)kept;
    } else {
        const int u = lane >> 3, v = lane & 7;
        const float s2 = sig2 * (u == 0 ? 0.5f : 1.f) * (v == 0 ? 0.5f : 1.f);
        float l2 = 0.f;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            if (k < N) {
                const float yy = yr[k] * yr[k];
                const float wc = yy / (yy + (k == 0 ? 0.5f * s2 : s2));
                g[o + k * ZS] *= wc;
                l2 = fmaf(wc, wc, l2);
            }
        }
        l2 = wave_sum(l2);
        weight = l2 <= 0.f ? 1.f : fminf(1.f / l2, ex.wmax);
    }
    zpass<NP>(g, N, lane, cn, true);
    wave_sync();
    cols_pass(g, N, c8, lane, true);
    wave_sync();
    rows_pass(g, N, c8, lane, true);
    wave_sync();
    return weight;
}

// LDS of the filter kernel: accumulators of the tile footprint (num per pixel,
// den per block origin) and one group buffer per wave.
// accumulators as 32-bit words: num per pixel and den per origin, 64 bits each
__host__ __device__ inline size_t filter_acc_floats(int FW, int FH)
{
    const int DW = FW - kBlock + 1, DH = FH - kBlock + 1;
    return ((((size_t)FW * FH + (size_t)DW * DH) * 2) + 3) & ~size_t(3);
}

__host__ __device__ inline size_t filter_lds(int FW, int FH, int G)
{
    return (filter_acc_floats(FW, FH) + (size_t)kWaves * G * ZS) * 4;
}


// Adds a tile's LDS accumulators into the planes of plane c, in 64-bit fixed
// point (the sums do not depend on the order of the adds): num per pixel of
// the footprint, den per block origin (finalize_kernel sums a pixel's den over
// the 8x8 origins whose block covers it).  Zero entries are not added.  Rows
// go to the waves and columns to the lanes (no division by the footprint
// width).
__device__ __forceinline__ void flush_tile(const FilterArgs &a, int c, const u64 *accNum, const u64 *accDen, int ox,
                                           int oy, int tid, int nthreads = kThreads, size_t dst_off = 0)
{
    const int FW = a.FW, FH = a.FH, DW = FW - kBlock + 1, DH = FH - kBlock + 1;
    const int lane = tid & 63, wave = tid >> 6, nw = nthreads >> 6;
    u64 *num = a.num[c] + dst_off, *den = a.den[c] + dst_off;
    for (int py = wave; py < FH; py += nw) {
        const int yy = oy + py;
        if (yy < 0 || yy >= a.H) continue;
        for (int px = lane; px < FW; px += 64) {
            const int xx = ox + px;
            const u64 v = accNum[py * FW + px];
            if (v != 0 && xx >= 0 && xx < a.W) atomicAdd(&num[(size_t)yy * a.pitch + xx], v);
        }
    }
    for (int dy = wave; dy < DH; dy += nw) {
        const int yy = oy + dy;
        if (yy < 0 || yy > a.H - kBlock) continue;
        for (int dx = lane; dx < DW; dx += 64) {
            const int xx = ox + dx;
            const u64 d = accDen[dy * DW + dx];
            if (d != 0 && xx >= 0 && xx <= a.W - kBlock) atomicAdd(&den[(size_t)yy * a.pitch + xx], d);
        }
    }
}

template <bool WIENER, bool BATCHED>
__global__ __launch_bounds__(kThreads) void filter_kernel(FilterArgs a0, const uint32_t *__restrict__ bmap,
                                                          const FilterArgs *__restrict__ bargs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int tx, ty;
    const FilterArgs &a = tile_unit<BATCHED>(a0, bargs, bmap, tx, ty);
    const int trx = a.trx, tnum = a.trx * a.try_;
    const int tx0 = tx * trx, ty0 = ty * a.try_;
    const int ox = grid_pos(tx0, a.bstep, a.W - kBlock) - a.range;
    const int oy = grid_pos(ty0, a.bstep, a.H - kBlock) - a.range;
    const int FW = a.FW, FH = a.FH, DW = FW - kBlock + 1, DH = FH - kBlock + 1;

    u64 *accNum = reinterpret_cast<u64 *>(smem);
    u64 *accDen = accNum + FW * FH;
    float *gbase = reinterpret_cast<float *>(smem) + filter_acc_floats(FW, FH);
    float *g = gbase + (size_t)wave * a.G * ZS;
    const int u = lane >> 3, v = lane & 7;

    for (int c = 0; c < 3; ++c) {
        if (!a.active[c]) continue;
        for (int k = tid; k < FW * FH; k += kThreads) accNum[k] = 0;
        for (int k = tid; k < DW * DH; k += kThreads) accDen[k] = 0;
        __syncthreads();
        const float *src = a.src[c];
        const float *gsrc = a.guide[c];
        for (int r = wave; r < tnum; r += kWaves) {
            const int gxi = tx0 + (r % trx), gyi = ty0 + (r / trx);
            if (gxi >= a.ngx || gyi >= a.ngy) continue;
            const size_t refIdx = (size_t)gyi * a.ngx + gxi;
            const int *pos = a.pos + refIdx * a.G;
            const int N = kconst(a.cnt)[refIdx];
            const float thr = a.thr[c], sig2 = a.sig2[c];
            const ExactCtx ex{pos, src, a.pitch, a.c8d, a.cnd + a.cnd_off[N], a.thr64[c], a.band[c], a.wmax};
            const float *cn = a.cn + a.cn_off[N];
            float w;
            if (N <= 1) w = group_core<1, WIENER>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
            else if (N <= 2) w = group_core<2, WIENER>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
            else if (N <= 4) w = group_core<4, WIENER>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
            else if (N <= 8) w = group_core<8, WIENER>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
            else if (N <= 16) w = group_core<16, WIENER>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
            else w = group_core<32, WIENER>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
            const int o = gidx(0, lane);
            const double ws = fix_weight(w);
            const int pl = lane < N ? pos[lane] : 0;
            for (int z = 0; z < N; ++z) {
                const int pz = __builtin_amdgcn_readlane(pl, z);
                const int py = (pz >> 16) - oy, px = (pz & 0xFFFF) - ox;
                fix_add_scaled(&accNum[(py + u) * FW + px + v], g[o + z * ZS], ws);
            }
            if (lane < N) fix_add(&accDen[((pl >> 16) - oy) * DW + (pl & 0xFFFF) - ox], w);
            wave_sync();
        }
        __syncthreads();
        flush_tile(a, c, accNum, accDen, ox, oy, tid);
        __syncthreads();
    }
}


// STACK: multi-frame groups over a whole stack (see FilterArgs): the tile's
// accumulators form a ring of frame slots; frame f is flushed into its planes
// when reference frame f + radius is done.
template <bool WIENER, bool BATCHED, bool STACK>
__global__ __launch_bounds__(kThreads) void filter_stack_kernel(FilterArgs a0, const uint32_t *__restrict__ bmap,
                                                                const FilterArgs *__restrict__ bargs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int tx, ty;
    const FilterArgs &a = tile_unit<BATCHED>(a0, bargs, bmap, tx, ty);
    const int trx = a.trx, tnum = a.trx * a.try_;
    const int tx0 = tx * trx, ty0 = ty * a.try_;
    const int ox = grid_pos(tx0, a.bstep, a.W - kBlock) - a.range;
    const int oy = grid_pos(ty0, a.bstep, a.H - kBlock) - a.range;
    const int FW = a.FW, FH = a.FH, DW = FW - kBlock + 1, DH = FH - kBlock + 1;
    const int nslots = STACK ? a.nslots : 1, T = STACK ? a.frames : 1, rad = STACK ? a.radius : 0;
    const size_t accw = filter_acc_floats(FW, FH);   // 32-bit words per slot
    const size_t refs = (size_t)a.ngx * a.ngy;

    u64 *accNum0 = reinterpret_cast<u64 *>(smem);
    float *gbase = reinterpret_cast<float *>(smem) + accw * nslots;
    float *g = gbase + (size_t)wave * a.G * ZS;
    const int u = lane >> 3, v = lane & 7;
    auto slot_num = [&](int f) { return accNum0 + (accw / 2) * (STACK ? f % nslots : 0); };
    auto zero_slot = [&](int f) {
        u64 *an = slot_num(f), *ad = an + FW * FH;
        for (int k = tid; k < FW * FH; k += kThreads) an[k] = 0;
        for (int k = tid; k < DW * DH; k += kThreads) ad[k] = 0;
    };
    auto flush_frame = [&](int c, int f) {
        const u64 *an = slot_num(f);
        flush_tile(a, c, an, an + FW * FH, ox, oy, tid, kThreads, STACK ? a.fstride * f : 0);
        __syncthreads();
    };

    for (int c = 0; c < 3; ++c) {
        if (!a.active[c]) continue;
        for (int f = 0; f < nslots; ++f) zero_slot(f);
        __syncthreads();
        const float *src = a.src[c];
        const float *gsrc = a.guide[c];
        for (int t = 0; t < T; ++t) {
            if (STACK && t > rad) {
                // frame t - rad - 1 is complete; its slot takes frame t + rad
                flush_frame(c, t - rad - 1);
                zero_slot(t - rad - 1);
                __syncthreads();
            }
            for (int r = wave; r < tnum; r += kWaves) {
                const int gxi = tx0 + (r % trx), gyi = ty0 + (r / trx);
                if (gxi >= a.ngx || gyi >= a.ngy) continue;
                const size_t refIdx = (size_t)t * refs + (size_t)gyi * a.ngx + gxi;
                const int *pos = a.pos + refIdx * a.G;
                const int *slot = STACK ? a.slot + refIdx * a.G : nullptr;
                const int N = kconst(a.cnt)[refIdx];
                const float thr = a.thr[c], sig2 = a.sig2[c];
                const ExactCtx ex{pos, src, a.pitch, a.c8d, a.cnd + a.cnd_off[N], a.thr64[c], a.band[c], a.wmax, slot,
                                  STACK ? a.fstride : 0};
                const float *cn = a.cn + a.cn_off[N];
                float w;
                if (N <= 1) w = group_core<1, WIENER, STACK>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
                else if (N <= 2) w = group_core<2, WIENER, STACK>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
                else if (N <= 4) w = group_core<4, WIENER, STACK>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
                else if (N <= 8) w = group_core<8, WIENER, STACK>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
                else if (N <= 16)
                    w = group_core<16, WIENER, STACK>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
                else w = group_core<32, WIENER, STACK>(g, N, lane, a.c8, cn, src, gsrc, a.pitch, thr, sig2, ex);
                const int o = gidx(0, lane);
                const double ws = fix_weight(w);
                const int pl = lane < N ? pos[lane] : 0;
                const int sl = STACK && lane < N ? slot[lane] : 0;
                for (int z = 0; z < N; ++z) {
                    const int pz = __builtin_amdgcn_readlane(pl, z);
                    const int py = (pz >> 16) - oy, px = (pz & 0xFFFF) - ox;
                    u64 *an = slot_num(STACK ? __builtin_amdgcn_readlane(sl, z) : 0);
                    fix_add_scaled(&an[(py + u) * FW + px + v], g[o + z * ZS], ws);
                }
                if (lane < N) {
                    u64 *ad = slot_num(sl) + FW * FH;
                    fix_add(&ad[((pl >> 16) - oy) * DW + (pl & 0xFFFF) - ox], w);
                }
                wave_sync();
            }
            __syncthreads();
        }
        for (int f = T - rad - 1 < 0 ? 0 : T - rad - 1; f < T; ++f) flush_frame(c, f);
    }
}


// ---------------------------------------------------------------------------
// Collaborative filtering on the matrix cores (FP16 or BF16 operands, FP32
// accumulation).  A group is the 32 x 64 matrix X (block z, pixel p = 8y+x),
// zero-padded to 32 blocks.  With K64 = C8 (x) C8 (the 2D DCT on p) and C_N
// padded to 32 x 32:
//   forward  T = X K64,   Y = C_N T                  (12 MFMAs 32x32x16)
//   inverse  V^T = Y'^T C_N,   Xhat^T = K64 V^T       (12 MFMAs)
// Every product takes the previous accumulator straight from registers: an
// accumulator 32x32 tile holds rows 8i+4h+j (h = lane/32) of column lane%32,
// which is an operand fragment of the next product for a fixed permutation of
// its 16-long k chunks; the constant operand (C_N or K64) is stored with the
// same permutation, so no data moves between the stages.
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename T>
struct MfmaT;
template <>
struct MfmaT<_Float16> {
    typedef f16x8 v8;
    static __device__ __forceinline__ f32x16 mma(v8 a, v8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaT<__bf16> {
    typedef bf16x8 v8;
    static __device__ __forceinline__ f32x16 mma(v8 a, v8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

template <typename T>
__device__ __forceinline__ typename MfmaT<T>::v8 load_frag(const uint16_t *__restrict__ tab, int frag, int lane)
{
    return *reinterpret_cast<const typename MfmaT<T>::v8 *>(tab + ((size_t)frag * 64 + lane) * 8);
}

// accumulator registers 8m..8m+7 as an operand fragment
template <typename T>
__device__ __forceinline__ typename MfmaT<T>::v8 acc_frag(const f32x16 &acc, int m)
{
    typename MfmaT<T>::v8 f;
#pragma unroll
    for (int e = 0; e < 8; ++e) f[e] = (T)acc[8 * m + e];
    return f;
}

// Forward 3D DCT of one group: y[t] = (C_N X K64) columns 32t..32t+31.
template <typename T>
__device__ __forceinline__ typename MfmaT<T>::v8 lds_frag(const uint16_t *tab, int frag, int lane)
{
    return *reinterpret_cast<const typename MfmaT<T>::v8 *>(tab + ((size_t)frag * 64 + lane) * 8);
}

// plane: the tile footprint staged in LDS (row pitch `pitch`, origin ox, oy).
// slot (multi-frame groups, else null): member c's footprint is plane + slot[c] * slotsz.
template <typename T>
__device__ __forceinline__ void mfma_forward(const float *plane, int pitch, int ox, int oy,
                                             const int *__restrict__ pos, int N, int lane, const uint16_t *kb,
                                             const typename MfmaT<T>::v8 (&cna)[2], f32x16 (&y)[2],
                                             const int *__restrict__ slot = nullptr, size_t slotsz = 0)
{
    typedef typename MfmaT<T>::v8 v8;
    const int c = lane & 31, h = lane >> 5;
    // A fragments of X: block z = c, pixel rows y = 2s + h
    v8 xa[4];
    if (c < N) {
        const int pz = pos[c];
        const float *b = plane + ((pz >> 16) - oy) * pitch + (pz & 0xFFFF) - ox + (slot ? slot[c] * slotsz : 0);

#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float *r = b + (2 * s + h) * pitch;
#pragma unroll
            for (int e = 0; e < 8; ++e) xa[s][e] = (T)r[e];
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e) xa[s][e] = (T)0.f;
    }
    f32x16 t[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        t[q] = (f32x16){};
#pragma unroll
        for (int s = 0; s < 4; ++s) t[q] = MfmaT<T>::mma(xa[s], lds_frag<T>(kb, s * 2 + q, lane), t[q]);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        y[q] = (f32x16){};
#pragma unroll
        for (int m = 0; m < 2; ++m) y[q] = MfmaT<T>::mma(cna[m], acc_frag<T>(t[q], m), y[q]);
    }
}

// The forward and inverse one column half q at a time (the same products as
// mfma_forward / mfma_inverse, in the same order), so that fewer accumulators
// are live at once: X's fragments, y_q = columns 32q..32q+31 of C_N X K64,
// vt_q = y_q^T C_N, and the remaining inverse product from vt_0, vt_1.
template <typename T>
__device__ __forceinline__ void mfma_load_x(const float *plane, int pitch, int ox, int oy,
                                            const int *__restrict__ pos, int N, int lane,
                                            typename MfmaT<T>::v8 (&xa)[4], const int *__restrict__ slot = nullptr,
                                            size_t slotsz = 0)
{
    const int c = lane & 31, h = lane >> 5;
    if (c < N) {
        const int pz = pos[c];
        const float *b = plane + ((pz >> 16) - oy) * pitch + (pz & 0xFFFF) - ox + (slot ? slot[c] * slotsz : 0);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float *r = b + (2 * s + h) * pitch;
#pragma unroll
            for (int e = 0; e < 8; ++e) xa[s][e] = (T)r[e];
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e) xa[s][e] = (T)0.f;
    }
}

// The staged footprint of the single-frame MFMA filter holds, for every
// pixel x of a row, the 32-bit word (T(p[x]), T(p[x + 1])): a block row of 8
// samples at any x is the words x, x + 2, x + 4, x + 6, read without
// conversion (the samples are converted once, when the footprint is staged).
template <typename T>
__device__ __forceinline__ uint32_t pack_pair(float a, float b)
{
    typedef T t2 __attribute__((ext_vector_type(2)));
    const t2 v = {(T)a, (T)b};
    return __builtin_bit_cast(uint32_t, v);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// The lane's rows of one member block from a paired-word footprint (packed
// position pz, or zeros when pz < 0).
template <typename T>
__device__ __forceinline__ void mfma_load_pairs(const uint32_t *plane, int pitch, int ox, int oy, int pz, int h,
                                                typename MfmaT<T>::v8 (&xa)[4])
{
    typedef typename MfmaT<T>::v8 v8;
    if (pz >= 0) {
        const uint32_t *b = plane + ((pz >> 16) - oy) * pitch + (pz & 0xFFFF) - ox;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t *r = b + (2 * s + h) * pitch;
            const u32x4 q = {r[0], r[2], r[4], r[6]};
            xa[s] = __builtin_bit_cast(v8, q);
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) xa[s] = __builtin_bit_cast(v8, (u32x4){0u, 0u, 0u, 0u});
    }
}

template <typename T>
__device__ __forceinline__ f32x16 mfma_forward_half(const typename MfmaT<T>::v8 (&xa)[4], int q, const uint16_t *kb,
                                                    int lane, const typename MfmaT<T>::v8 (&cna)[2])
{
    f32x16 t = (f32x16){};
#pragma unroll
    for (int s = 0; s < 4; ++s) t = MfmaT<T>::mma(xa[s], lds_frag<T>(kb, s * 2 + q, lane), t);
    f32x16 y = (f32x16){};
#pragma unroll
    for (int m = 0; m < 2; ++m) y = MfmaT<T>::mma(cna[m], acc_frag<T>(t, m), y);
    return y;
}

template <typename T>
__device__ __forceinline__ f32x16 mfma_inverse_half(const f32x16 &yq, const typename MfmaT<T>::v8 (&cnb)[2])
{
    f32x16 vt = (f32x16){};
#pragma unroll
    for (int m = 0; m < 2; ++m) vt = MfmaT<T>::mma(acc_frag<T>(yq, m), cnb[m], vt);
    return vt;
}

template <typename T>
__device__ __forceinline__ void mfma_inverse_rest(const f32x16 (&vt)[2], const uint16_t *ka, int lane,
                                                  f32x16 (&xh)[2])
{
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        xh[r] = (f32x16){};
#pragma unroll
        for (int s = 0; s < 4; ++s)
            xh[r] = MfmaT<T>::mma(lds_frag<T>(ka, r * 4 + s, lane), acc_frag<T>(vt[s >> 1], s & 1), xh[r]);
    }
}

// Inverse: xh[r] = rows 32r..32r+31 (pixels) of Xhat^T = K64 (Y'^T C_N).
template <typename T>
__device__ __forceinline__ void mfma_inverse(const f32x16 (&y)[2], const uint16_t *ka, int lane,
                                             const typename MfmaT<T>::v8 (&cnb)[2], f32x16 (&xh)[2])
{
    f32x16 vt[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        vt[q] = (f32x16){};
#pragma unroll
        for (int m = 0; m < 2; ++m) vt[q] = MfmaT<T>::mma(acc_frag<T>(y[q], m), cnb[m], vt[q]);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        xh[r] = (f32x16){};
#pragma unroll
        for (int s = 0; s < 4; ++s)
            xh[r] = MfmaT<T>::mma(lds_frag<T>(ka, r * 4 + s, lane), acc_frag<T>(vt[s >> 1], s & 1), xh[r]);
    }
}

// FP32-accurate products on the FP16 matrix cores (the FP32_MFMA form).  An
// operand a splits into a_hi = fp16(a) and a_lo = fp16(a - a_hi), 22
// significant bits together, and A B = A_hi B_hi + A_hi B_lo + A_lo B_hi in
// FP32 accumulation, less A_lo B_lo (about 2^-22 of the product).  The small
// terms go in first.
// A pair of FP32 values as packed FP16 high halves h and low halves
// l = fp16(v - hi) (the subtraction is exact).  The low half is written as
// fma(v, one, -hi) with `one` opaque to the compiler (an empty asm emits no
// instruction), so that it becomes a mixed-precision FMA (v_fma_mix) rather
// than a conversion back to FP32, a subtraction and a second conversion.
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_pair(float v0, float v1, uint32_t &h, uint32_t &l)
{
    float one = 1.0f;
    asm("" : "+v"(one));
    const h16x2 hh = __builtin_convertvector((f32x2v){v0, v1}, h16x2);
    const h16x2 ll = {(_Float16)__builtin_fmaf(v0, one, -(float)hh[0]),
                      (_Float16)__builtin_fmaf(v1, one, -(float)hh[1])};
    h = __builtin_bit_cast(uint32_t, hh);
    l = __builtin_bit_cast(uint32_t, ll);
}

__device__ __forceinline__ void split_frag(const f32x16 &acc, int m, f16x8 &hi, f16x8 &lo)
{
    u32x4 H, L;
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        uint32_t h, l;
        split_pair(acc[8 * m + e], acc[8 * m + e + 1], h, l);
        H[e / 2] = h;
        L[e / 2] = l;
    }
    hi = __builtin_bit_cast(f16x8, H);
    lo = __builtin_bit_cast(f16x8, L);
}

// mfma_load_x in high and low halves (the FP32_MFMA / plain forms' stacks)
__device__ __forceinline__ void mfma_load_x_split(const float *plane, int pitch, int ox, int oy,
                                                  const int *__restrict__ pos, int N, int lane, f16x8 (&xh)[4],
                                                  f16x8 (&xl)[4], const int *__restrict__ slot = nullptr,
                                                  size_t slotsz = 0)
{
    const int c = lane & 31, h = lane >> 5;
    if (c < N) {
        const int pz = pos[c];
        const float *b = plane + ((pz >> 16) - oy) * pitch + (pz & 0xFFFF) - ox + (slot ? slot[c] * slotsz : 0);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float *r = b + (2 * s + h) * pitch;
            u32x4 H, L;
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                uint32_t hh, ll;
                split_pair(r[e], r[e + 1], hh, ll);
                H[e / 2] = hh;
                L[e / 2] = ll;
            }
            xh[s] = __builtin_bit_cast(f16x8, H);
            xl[s] = __builtin_bit_cast(f16x8, L);
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e) xh[s][e] = xl[s][e] = (_Float16)0.f;
    }
}

__device__ __forceinline__ f32x16 mma3(const f16x8 &ah, const f16x8 &al, const f16x8 &bh, const f16x8 &bl,
                                       f32x16 c)
{
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, c, 0, 0, 0);
}

// the low halves of a pair of samples, as mfma_load_pairs reads them
__device__ __forceinline__ uint32_t pack_lo_pair(float a, float b)
{
    uint32_t h, l;
    split_pair(a, b, h, l);
    return l;
}

// mfma_forward_half with every product split
__device__ __forceinline__ f32x16 split_forward_half(const f16x8 (&xh)[4], const f16x8 (&xl)[4], int q,
                                                     const uint16_t *kb, const uint16_t *kbl, int lane,
                                                     const f16x8 (&cah)[2], const f16x8 (&cal)[2])
{
    f32x16 t = (f32x16){};
#pragma unroll
    for (int s = 0; s < 4; ++s)
        t = mma3(xh[s], xl[s], lds_frag<_Float16>(kb, s * 2 + q, lane), lds_frag<_Float16>(kbl, s * 2 + q, lane), t);
    f32x16 y = (f32x16){};
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        f16x8 th, tl;
        split_frag(t, m, th, tl);
        y = mma3(cah[m], cal[m], th, tl, y);
    }
    return y;
}

// mfma_inverse_half with every product split
__device__ __forceinline__ f32x16 split_inverse_half(const f32x16 &yq, const f16x8 (&cbh)[2], const f16x8 (&cbl)[2])
{
    f32x16 vt = (f32x16){};
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        f16x8 yh, yl;
        split_frag(yq, m, yh, yl);
        vt = mma3(yh, yl, cbh[m], cbl[m], vt);
    }
    return vt;
}

// mfma_inverse_rest with every product split
__device__ __forceinline__ void split_inverse_rest(const f32x16 (&vt)[2], const uint16_t *ka, const uint16_t *kal,
                                                   int lane, f32x16 (&xh)[2])
{
    f16x8 vh[4], vl[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) split_frag(vt[s >> 1], s & 1, vh[s], vl[s]);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        xh[r] = (f32x16){};
#pragma unroll
        for (int s = 0; s < 4; ++s)
            xh[r] = mma3(lds_frag<_Float16>(ka, r * 4 + s, lane), lds_frag<_Float16>(kal, r * 4 + s, lane), vh[s],
                         vl[s], xh[r]);
    }
}

// LDS of the MFMA filter: accumulators, the K64 operand fragments (forward B
// and inverse A, 16 x 64 lanes x 8 elements of 16 bits), then the staged
// source (and guide) footprint, one 32-bit word per pixel.
constexpr int kK64Elems = 16 * 64 * 8;
// stack: nslots accumulator slots and no staged footprint (member blocks are
// read from HBM).
__host__ __device__ inline size_t filter_mfma_lds(int FW, int FH, bool wiener, int nslots = 1, bool stack = false,
                                                  bool split = false)
{
    // single frame: the staged footprint (split: its high then its low
    // halves), then the Wiener schedule (2 ints per unit + count)
    const size_t st = stack ? 0
                            : (size_t)FW * FH * (wiener ? 2 : 1) * (split ? 2 : 1) +
                                  (wiener ? 2 * kMaxTileRefs + 4 : 0);
    return filter_acc_floats(FW, FH) * 4 * nslots + kK64Elems * 2 * (split ? 2 : 1) + st * 4;
}

// 4 waves per SIMD: two 8-wave workgroups per CU (at most 128 registers)
// NW waves per workgroup: 8 (two workgroups per CU) or 16 (one workgroup of a
// larger tile per CU); 4 waves per SIMD either way, 128 registers.
// SPLIT (T = _Float16): the FP32_MFMA form, every product split in high and
// low halves (split_frag, mma3); the K64 fragments and the staged footprint
// hold both halves; 2 waves per SIMD (256 registers).
template <typename T, bool WIENER, bool BATCHED, int NW = kMWaves, bool SPLIT = false>
__global__ __launch_bounds__(NW * 64, SPLIT ? 2 : 4) void filter_mfma_kernel(FilterArgs a0,
                                                                            const uint32_t *__restrict__ bmap,
                                                                            const FilterArgs *__restrict__ bargs)
{
    static_assert(!SPLIT || std::is_same<T, _Float16>::value, "the split form runs on FP16 operands");
    typedef typename MfmaT<T>::v8 v8;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = lane & 31, h = lane >> 5;
    int tx, ty;
    const FilterArgs &a = tile_unit<BATCHED>(a0, bargs, bmap, tx, ty);
    const int trx = a.trx, tnum = a.trx * a.try_;
    const int tx0 = tx * trx, ty0 = ty * a.try_;
    const int ox = grid_pos(tx0, a.bstep, a.W - kBlock) - a.range;
    const int oy = grid_pos(ty0, a.bstep, a.H - kBlock) - a.range;
    const int FW = a.FW, FH = a.FH, DW = FW - kBlock + 1, DH = FH - kBlock + 1;

    u64 *accNum = reinterpret_cast<u64 *>(smem);
    u64 *accDen = accNum + FW * FH;
    uint16_t *kb = reinterpret_cast<uint16_t *>(reinterpret_cast<float *>(smem) + filter_acc_floats(FW, FH));
    uint16_t *ka = kb + kK64Elems / 2;
    uint16_t *kbl = kb + kK64Elems, *kal = kbl + kK64Elems / 2;  // SPLIT: the low halves
    uint32_t *stage = reinterpret_cast<uint32_t *>(kb + kK64Elems * (SPLIT ? 2 : 1));
    uint32_t *lsrc = stage;  // paired words (mfma_load_pairs)
    uint32_t *lgd = stage + FW * FH;
    uint32_t *lsrcl = stage + (WIENER ? 2 : 1) * FW * FH, *lgdl = lsrcl + FW * FH;  // SPLIT: low halves
    {
        // the K64 fragments (kb then ka are contiguous in the table)
        const uint4 *srcq = reinterpret_cast<const uint4 *>(a.kb);
        uint4 *dstq = reinterpret_cast<uint4 *>(kb);
        for (int k = tid; k < kK64Elems / 8; k += NW * 64) dstq[k] = srcq[k];
        if constexpr (SPLIT) {
            const uint4 *srcl = reinterpret_cast<const uint4 *>(a.kbl);
            uint4 *dstl = reinterpret_cast<uint4 *>(kbl);
            for (int k = tid; k < kK64Elems / 8; k += NW * 64) dstl[k] = srcl[k];
        }
    }

    // Wiener step: the tile's schedule of units, consecutive reference blocks
    // (2p, 2p+1) as one unit when both groups have at most 16 blocks, else one
    // unit each; sched[2u], sched[2u+1] = (refA, refB or -1), sched[2 * kMaxTileRefs]
    // = the number of units.
    int *sched = reinterpret_cast<int *>(stage + (WIENER ? 2 : 1) * FW * FH * (SPLIT ? 2 : 1));
    if (WIENER && wave == 0) {
        const int np = (tnum + 1) / 2;
        for (int p0 = 0; p0 < np; p0 += 64) {
            const int p = p0 + lane;
            int refA = -1, refB = -1, NA = 0, NB = 0;
            if (p < np) {
                const int rA = 2 * p, rB = rA + 1;
                const int gxA = tx0 + (rA % trx), gyA = ty0 + (rA / trx);
                if (gxA < a.ngx && gyA < a.ngy) {
                    refA = gyA * a.ngx + gxA;
                    NA = a.cnt[refA];
                }
                const int gxB = tx0 + (rB % trx), gyB = ty0 + (rB / trx);
                if (rB < tnum && gxB < a.ngx && gyB < a.ngy) {
                    refB = gyB * a.ngx + gxB;
                    NB = a.cnt[refB];
                }
            }
            const bool pr = refA >= 0 && refB >= 0 && NA <= 16 && NB <= 16;
            const uint32_t nu = pr ? 1u : (uint32_t)(refA >= 0) + (uint32_t)(refB >= 0);
            const uint32_t incl = wave_incl_scan(nu, lane);
            const int base = (p0 == 0 ? 0 : sched[2 * kMaxTileRefs]) + (int)(incl - nu);
            if (pr) {
                sched[2 * base] = refA;
                sched[2 * base + 1] = refB;
            } else {
                int u = base;
                if (refA >= 0) {
                    sched[2 * u] = refA;
                    sched[2 * u + 1] = -1;
                    ++u;
                }
                if (refB >= 0) {
                    sched[2 * u] = refB;
                    sched[2 * u + 1] = -1;
                }
            }
            const uint32_t total = __shfl(incl, 63);
            wave_sync();
            if (lane == 0) sched[2 * kMaxTileRefs] = (p0 == 0 ? 0 : sched[2 * kMaxTileRefs]) + (int)total;
            wave_sync();
        }
    }

    for (int pl = 0; pl < 3; ++pl) {
        if (!a.active[pl]) continue;
        const float *src = a.src[pl];
        const float *gsrc = a.guide[pl];
        // rows to the waves, columns to the lanes
        for (int py = wave; py < FH; py += NW) {
            const int yy = oy + py;
            const bool rin = yy >= 0 && yy < a.H;
            for (int px = lane; px < FW; px += 64) {
                const int k = py * FW + px, xx = ox + px;
                accNum[k] = 0;
                const bool in0 = rin && xx >= 0 && xx < a.W;
                const bool in1 = rin && px + 1 < FW && xx + 1 >= 0 && xx + 1 < a.W;
                const size_t o = (size_t)yy * a.pitch + xx;
                const float s0 = in0 ? src[o] : 0.f, s1 = in1 ? src[o + 1] : 0.f;
                lsrc[k] = pack_pair<T>(s0, s1);
                if constexpr (SPLIT) lsrcl[k] = pack_lo_pair(s0, s1);
                if (WIENER) {
                    const float g0 = in0 ? gsrc[o] : 0.f, g1 = in1 ? gsrc[o + 1] : 0.f;
                    lgd[k] = pack_pair<T>(g0, g1);
                    if constexpr (SPLIT) lgdl[k] = pack_lo_pair(g0, g1);
                }
            }
        }
        for (int k = tid; k < DW * DH; k += NW * 64) accDen[k] = 0;
        __syncthreads();
        const float thr = a.thr[pl], sig2 = a.sig2[pl];
        if constexpr (WIENER) {
            // Units of work from the tile's schedule: a pair of groups of at
            // most 16 blocks sharing one operand (as in step 1), or one group.
            const int nunits = sched[2 * kMaxTileRefs];
            for (int un = wave; un < nunits; un += NW) {
                const int2 unit = reinterpret_cast<const int2 *>(sched)[un];
                const size_t refA = (size_t)unit.x;
                const bool paired = unit.y >= 0;
                const size_t refB = paired ? (size_t)unit.y : 0;
                const int N0 = a.cnt[refA], N1 = paired ? a.cnt[refB] : 0;
                const int *pos0 = a.pos + refA * a.G, *pos1 = a.pos + refB * a.G;
                const int g = paired ? (c >> 4) : 0, mi = paired ? (c & 15) : c;
                const int Ng = g ? N1 : N0;
                const int pz = mi < Ng ? (g ? pos1 : pos0)[mi] : -1;
                v8 cna[2], cnb[2], cnal[2], cnbl[2];
                {
                    const uint16_t *ta = paired ? a.cnap : a.cna, *tb = paired ? a.cnbp : a.cnb;
                    const int fi = paired ? (N0 - 1) * 17 + N1 : N0 - 1;
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        cna[m] = load_frag<T>(ta, fi * 2 + m, lane);
                        cnb[m] = load_frag<T>(tb, fi * 2 + m, lane);
                    }
                    if constexpr (SPLIT) {
                        const uint16_t *tal = paired ? a.cnapl : a.cnal, *tbl = paired ? a.cnbpl : a.cnbl;
#pragma unroll
                        for (int m = 0; m < 2; ++m) {
                            cnal[m] = load_frag<T>(tal, fi * 2 + m, lane);
                            cnbl[m] = load_frag<T>(tbl, fi * 2 + m, lane);
                        }
                    }
                }
                // the guide's spectrum first, one column half at a time, and
                // from it the gains (the guide's operand is dead before the
                // source's is loaded); then the source's halves, shrunk and
                // taken through the first inverse product
                f32x16 wg[2], vt[2];
                float l20 = 0.f, l21 = 0.f;
                {
                    v8 xg[4], xgl[4];
                    mfma_load_pairs<T>(lgd, FW, ox, oy, pz, h, xg);
                    if constexpr (SPLIT) mfma_load_pairs<T>(lgdl, FW, ox, oy, pz, h, xgl);
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        if constexpr (SPLIT) wg[q] = split_forward_half(xg, xgl, q, kb, kbl, lane, cna, cnal);
                        else wg[q] = mfma_forward_half<T>(xg, q, kb, lane, cna);
                        // column 32q + c = frequency (u, v): u = 4q + c/8, v = c%8
                        const float su =
                            sig2 * ((4 * q + (c >> 3)) == 0 ? 0.5f : 1.f) * ((c & 7) == 0 ? 0.5f : 1.f);
#pragma unroll
                        for (int k = 0; k < 16; ++k) {
                            // group-axis frequency: the row within the group
                            // group-axis DC: row 8(k/4) + 4h + k%4 is 0 (or 16 when paired),
                            // possible only where k % 4 == 0 (static) and h == 0
                            const bool kz0 = (k & 3) == 0 && h == 0 && (paired ? ((k >> 2) & 1) == 0 : k == 0);
                            const float yy = wg[q][k] * wg[q][k];
                            // 1-ulp reciprocal: the operands already carry 16-bit rounding
                            const float wc = yy * __builtin_amdgcn_rcpf(yy + (kz0 ? 0.5f * su : su));
                            wg[q][k] = wc;
                            // rows 0..15 (elements k < 8) and 16..31 apart; one group
                            // adds them below (the order of the stack kernel's sum)
                            if (k < 8) l20 = fmaf(wc, wc, l20);
                            else l21 = fmaf(wc, wc, l21);
                        }
                    }
                }
                // keep the source's loads after the gains (register pressure)
                __builtin_amdgcn_sched_barrier(0);
                {
                    v8 xs[4], xsl[4];
                    mfma_load_pairs<T>(lsrc, FW, ox, oy, pz, h, xs);
                    if constexpr (SPLIT) mfma_load_pairs<T>(lsrcl, FW, ox, oy, pz, h, xsl);
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        f32x16 yq;
                        if constexpr (SPLIT) yq = split_forward_half(xs, xsl, q, kb, kbl, lane, cna, cnal);
                        else yq = mfma_forward_half<T>(xs, q, kb, lane, cna);
#pragma unroll
                        for (int k = 0; k < 16; ++k) yq[k] *= wg[q][k];
                        if constexpr (SPLIT) vt[q] = split_inverse_half(yq, cnb, cnbl);
                        else vt[q] = mfma_inverse_half<T>(yq, cnb);
                    }
                }
                if (!paired) {
                    l20 += l21;
                    l21 = 0.f;
                }
                l20 = wave_sum(l20);
                l21 = wave_sum(l21);
                f32x16 xh[2];
                if constexpr (SPLIT) split_inverse_rest(vt, ka, kal, lane, xh);
                else mfma_inverse_rest<T>(vt, ka, lane, xh);
                const float l2 = g ? l21 : l20;
                const float w = l2 <= 0.f ? 1.f : fminf(1.f / l2, a.wmax);
                const double ws = fix_weight(w);
                if (pz >= 0) {
                    const int py = (pz >> 16) - oy, px = (pz & 0xFFFF) - ox + 4 * h;
#pragma unroll
                    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                fix_add_scaled(&accNum[(py + 4 * rr + i) * FW + px + j], xh[rr][4 * i + j], ws);
                    if (h == 0) fix_add(&accDen[py * DW + (pz & 0xFFFF) - ox], w);
                }
            }
        } else {
        // Reference blocks in pairs.  Two groups of at most 16 blocks share
        // one 32-row operand (group 0 on rows 0..15, group 1 on 16..31, the
        // group-axis DCT block-diagonal); otherwise each group takes a pass
        // of its own on rows 0..N-1.
        const int npairs = (tnum + 1) / 2;
        for (int pi = wave; pi < npairs; pi += NW) {
            int NA = 0, NB = 0;
            size_t refA = 0, refB = 0;
            {
                const int rA = 2 * pi, rB = rA + 1;
                const int gxA = tx0 + (rA % trx), gyA = ty0 + (rA / trx);
                if (gxA < a.ngx && gyA < a.ngy) {
                    refA = (size_t)gyA * a.ngx + gxA;
                    NA = a.cnt[refA];
                }
                const int gxB = tx0 + (rB % trx), gyB = ty0 + (rB / trx);
                if (rB < tnum && gxB < a.ngx && gyB < a.ngy) {
                    refB = (size_t)gyB * a.ngx + gxB;
                    NB = a.cnt[refB];
                }
            }
            const bool paired = NA > 0 && NB > 0 && NA <= 16 && NB <= 16;
            const int passes = paired ? 1 : (NA > 0) + (NB > 0);
            for (int ps = 0; ps < passes; ++ps) {
                const bool first = paired || (ps == 0 && NA > 0);
                const int N0 = first ? NA : NB, N1 = paired ? NB : 0;
                const int *pos0 = a.pos + (first ? refA : refB) * a.G, *pos1 = a.pos + refB * a.G;
                // this lane's member: group g, index mi
                const int g = paired ? (c >> 4) : 0, mi = paired ? (c & 15) : c;
                const int Ng = g ? N1 : N0;
                const int pz = mi < Ng ? (g ? pos1 : pos0)[mi] : -1;
                v8 cna[2], cnb[2], cnal[2], cnbl[2];
                {
                    const uint16_t *ta = paired ? a.cnap : a.cna, *tb = paired ? a.cnbp : a.cnb;
                    const int fi = paired ? (N0 - 1) * 17 + N1 : N0 - 1;
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        cna[m] = load_frag<T>(ta, fi * 2 + m, lane);
                        cnb[m] = load_frag<T>(tb, fi * 2 + m, lane);
                    }
                    if constexpr (SPLIT) {
                        const uint16_t *tal = paired ? a.cnapl : a.cnal, *tbl = paired ? a.cnbpl : a.cnbl;
#pragma unroll
                        for (int m = 0; m < 2; ++m) {
                            cnal[m] = load_frag<T>(tal, fi * 2 + m, lane);
                            cnbl[m] = load_frag<T>(tbl, fi * 2 + m, lane);
                        }
                    }
                }
                // accumulator element k of this lane lies on row 8(k/4) + 4h + k%4,
                // of group 1 when paired and k >= 8
                // (the Wiener step takes the schedule above: here hard thresholds only)
                f32x16 xh[2];
                v8 xs[4], xsl[4];
                mfma_load_pairs<T>(lsrc, FW, ox, oy, pz, h, xs);
                if constexpr (SPLIT) mfma_load_pairs<T>(lsrcl, FW, ox, oy, pz, h, xsl);
                f32x16 y[2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    if constexpr (SPLIT) y[q] = split_forward_half(xs, xsl, q, kb, kbl, lane, cna, cnal);
                    else y[q] = mfma_forward_half<T>(xs, q, kb, lane, cna);
                }
                // SPLIT with band >= 0 (the plain form): decisions within
                // band of the threshold are re-made from the coefficient in
                // FP64 (exact_coef_wave, the whole wave per coefficient);
                // bit 16q + k of a lane's masks stands for its element k of half q
                uint32_t nearm = 0, zerom = 0;
                if constexpr (SPLIT) {
                    const float band = a.band[pl];
                    if (band >= 0.f) {
#pragma unroll
                        for (int q = 0; q < 2; ++q)
#pragma unroll
                            for (int k = 0; k < 16; ++k)
                                if (fabsf(fabsf(y[q][k]) - thr) <= band) nearm |= 1u << (16 * q + k);
                        uint64_t lanes = __ballot(nearm != 0);
                        while (lanes) {
                            const int L = __ffsll((unsigned long long)lanes) - 1;
                            lanes &= lanes - 1;
                            uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)nearm, L);
                            while (m) {
                                const int bit = __ffs(m) - 1;
                                m &= m - 1;
                                const int q = bit >> 4, k = bit & 15, cL = L & 31, hL = L >> 5;
                                // element k of lane L: group-axis row 8(k/4) + 4h + k%4, column 32q + c
                                const int row = 8 * (k >> 2) + 4 * hL + (k & 3);
                                const int gg = paired ? (row >> 4) : 0, kz = paired ? (row & 15) : row;
                                const int Ng = gg ? N1 : N0;
                                const double oc = exact_coef_wave(kz, 4 * q + (cL >> 3), cL & 7, Ng, lane,
                                                                  gg ? pos1 : pos0, src, a.pitch, a.c8d,
                                                                  a.cnd + a.cnd_off[Ng]);
                                if (lane == L && fabs(oc) <= a.thr64[pl]) zerom |= 1u << bit;
                            }
                        }
                    }
                }
                int kept = 0;  // group 0 in the low 16 bits, group 1 above
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        const uint32_t b = 1u << (16 * q + k);
                        const bool zero = (nearm & b) ? (zerom & b) != 0 : fabsf(y[q][k]) <= thr;
                        if (zero) y[q][k] = 0.f;
                        else kept += (paired && k >= 8) ? 65536 : 1;
                    }
                kept = wave_sum(kept);
                const int k0 = kept & 0xFFFF, k1 = kept >> 16;
                const float w0 = k0 < 1 ? 1.f : 1.f / (float)k0;
                const float w1 = k1 < 1 ? 1.f : 1.f / (float)k1;
                if constexpr (SPLIT) {
                    f32x16 vt[2];
#pragma unroll
                    for (int q = 0; q < 2; ++q) vt[q] = split_inverse_half(y[q], cnb, cnbl);
                    split_inverse_rest(vt, ka, kal, lane, xh);
                } else {
                    mfma_inverse<T>(y, ka, lane, cnb, xh);
                }
                const float w = g ? w1 : w0;
                const double ws = fix_weight(w);
                if (pz >= 0) {
                    const int py = (pz >> 16) - oy, px = (pz & 0xFFFF) - ox + 4 * h;
#pragma unroll
                    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                fix_add_scaled(&accNum[(py + 4 * rr + i) * FW + px + j], xh[rr][4 * i + j], ws);
                    if (h == 0) fix_add(&accDen[py * DW + (pz & 0xFFFF) - ox], w);
                }
            }
        }
        }
        __syncthreads();
        flush_tile(a, pl, accNum, accDen, ox, oy, tid, NW * 64);
        __syncthreads();
    }
}

// STACK: multi-frame groups over a whole stack (see FilterArgs and
// filter_stack_kernel): a ring of accumulator slots, member blocks read from HBM.
// NW waves per workgroup as in filter_mfma_kernel (16 where one workgroup
// fills a CU's LDS), at most 128 registers.
// SPLIT: as filter_mfma_kernel's (split products, the FP64 re-decision of
// the plain form), members read from HBM in high and low halves.
template <typename T, bool WIENER, bool BATCHED, bool STACK, int NW = kMWaves, bool SPLIT = false>
__global__ __launch_bounds__(NW * 64, SPLIT ? 2 : 4) void filter_mfma_stack_kernel(FilterArgs a0,
                                                                      const uint32_t *__restrict__ bmap,
                                                                      const FilterArgs *__restrict__ bargs)
{
    typedef typename MfmaT<T>::v8 v8;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 31, h = lane >> 5;
    int tx, ty;
    const FilterArgs &a = tile_unit<BATCHED>(a0, bargs, bmap, tx, ty);
    const int trx = a.trx, tnum = a.trx * a.try_;
    const int tx0 = tx * trx, ty0 = ty * a.try_;
    const int ox = grid_pos(tx0, a.bstep, a.W - kBlock) - a.range;
    const int oy = grid_pos(ty0, a.bstep, a.H - kBlock) - a.range;
    const int FW = a.FW, FH = a.FH, DW = FW - kBlock + 1, DH = FH - kBlock + 1;
    const int nslots = STACK ? a.nslots : 1, NF = STACK ? a.frames : 1, rad = STACK ? a.radius : 0;
    const size_t accw = filter_acc_floats(FW, FH);   // 32-bit words per slot
    const int fsz = FW * FH;                          // footprint pixels
    const size_t refs = (size_t)a.ngx * a.ngy;

    static_assert(!SPLIT || (STACK && std::is_same<T, _Float16>::value), "split stacks run on FP16 operands");
    u64 *accNum0 = reinterpret_cast<u64 *>(smem);
    uint16_t *kb = reinterpret_cast<uint16_t *>(reinterpret_cast<float *>(smem) + accw * nslots);
    uint16_t *ka = kb + kK64Elems / 2;
    uint16_t *kbl = kb + kK64Elems, *kal = kbl + kK64Elems / 2;  // SPLIT: the low halves
    float *stage = reinterpret_cast<float *>(kb + kK64Elems * (SPLIT ? 2 : 1));
    float *lsrc = stage;          // single frame: the source footprint staged in LDS
    float *lgd = stage + fsz;     // and the Wiener guide's
    {
        // the K64 fragments (kb then ka are contiguous in the table)
        const uint4 *srcq = reinterpret_cast<const uint4 *>(a.kb);
        uint4 *dstq = reinterpret_cast<uint4 *>(kb);
        for (int k = tid; k < kK64Elems / 8; k += NW * 64) dstq[k] = srcq[k];
        if constexpr (SPLIT) {
            const uint4 *srcl = reinterpret_cast<const uint4 *>(a.kbl);
            uint4 *dstl = reinterpret_cast<uint4 *>(kbl);
            for (int k = tid; k < kK64Elems / 8; k += NW * 64) dstl[k] = srcl[k];
        }
    }
    auto slot_num = [&](int f) { return accNum0 + (accw / 2) * (STACK ? f % nslots : 0); };
    auto zero_slot = [&](int f) {
        u64 *an = slot_num(f), *ad = an + fsz;
        for (int k = tid; k < fsz; k += NW * 64) an[k] = 0;
        for (int k = tid; k < DW * DH; k += NW * 64) ad[k] = 0;
    };
    auto flush_frame = [&](int pl, int f) {
        const u64 *an = slot_num(f);
        flush_tile(a, pl, an, an + fsz, ox, oy, tid, NW * 64, STACK ? a.fstride * f : 0);
        __syncthreads();
    };

    for (int pl = 0; pl < 3; ++pl) {
        if (!a.active[pl]) continue;
        const float *src = a.src[pl];
        const float *gsrc = a.guide[pl];
        if (!STACK) {
            for (int k = tid; k < fsz; k += NW * 64) {
                const int py = k / FW, px = k - py * FW;
                const int yy = oy + py, xx = ox + px;
                const bool in = yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
                const size_t o = (size_t)yy * a.pitch + xx;
                lsrc[k] = in ? src[o] : 0.f;
                if (WIENER) lgd[k] = in ? gsrc[o] : 0.f;
            }
        }
        for (int f = 0; f < nslots; ++f) zero_slot(f);
        __syncthreads();
        const float thr = a.thr[pl], sig2 = a.sig2[pl];
        // where member blocks are read: the staged footprint, or the frames in HBM
        const float *fsrc = STACK ? src : lsrc, *fgd = STACK ? gsrc : lgd;
        const int fp = STACK ? a.pitch : FW, fox = STACK ? 0 : ox, foy = STACK ? 0 : oy;
        for (int t = 0; t < NF; ++t) {
            if (STACK && t > rad) {
                // frame t - rad - 1 is complete; its slot takes frame t + rad
                flush_frame(pl, t - rad - 1);
                zero_slot(t - rad - 1);
                __syncthreads();
            }
            for (int r = wave; r < tnum; r += NW) {
                const int gxi = tx0 + (r % trx), gyi = ty0 + (r / trx);
                if (gxi >= a.ngx || gyi >= a.ngy) continue;
                const size_t refIdx = (size_t)t * refs + (size_t)gyi * a.ngx + gxi;
                const int *pos = a.pos + refIdx * a.G;
                const int *slot = STACK ? a.slot + refIdx * a.G : nullptr;
                const size_t sstride = STACK ? a.fstride : 0;
                const int N = kconst(a.cnt)[refIdx];
                v8 cna[2], cnb[2], cnal[2], cnbl[2];
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    cna[m] = load_frag<T>(a.cna, (N - 1) * 2 + m, lane);
                    cnb[m] = load_frag<T>(a.cnb, (N - 1) * 2 + m, lane);
                    if constexpr (SPLIT) {
                        cnal[m] = load_frag<T>(a.cnal, (N - 1) * 2 + m, lane);
                        cnbl[m] = load_frag<T>(a.cnbl, (N - 1) * 2 + m, lane);
                    }
                }
                float w;
                f32x16 xh[2];
                if (!WIENER) {
                    f32x16 y[2];
                    if constexpr (SPLIT) {
                        v8 xs[4], xsl[4];
                        mfma_load_x_split(fsrc, fp, fox, foy, pos, N, lane, xs, xsl, slot, sstride);
#pragma unroll
                        for (int q = 0; q < 2; ++q) y[q] = split_forward_half(xs, xsl, q, kb, kbl, lane, cna, cnal);
                    } else {
                        mfma_forward<T>(fsrc, fp, fox, foy, pos, N, lane, kb, cna, y, slot, sstride);
                    }
                    // SPLIT with band >= 0 (the plain form): near-threshold
                    // decisions re-made in FP64 (as in filter_mfma_kernel)
                    uint32_t nearm = 0, zerom = 0;
                    if constexpr (SPLIT) {
                        const float band = a.band[pl];
                        if (band >= 0.f) {
#pragma unroll
                            for (int q = 0; q < 2; ++q)
#pragma unroll
                                for (int k = 0; k < 16; ++k)
                                    if (fabsf(fabsf(y[q][k]) - thr) <= band) nearm |= 1u << (16 * q + k);
                            uint64_t lanes = __ballot(nearm != 0);
                            while (lanes) {
                                const int L = __ffsll((unsigned long long)lanes) - 1;
                                lanes &= lanes - 1;
                                uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)nearm, L);
                                while (m) {
                                    const int bit = __ffs(m) - 1;
                                    m &= m - 1;
                                    const int q = bit >> 4, k = bit & 15, cL = L & 31, hL = L >> 5;
                                    const int row = 8 * (k >> 2) + 4 * hL + (k & 3);
                                    const double oc = exact_coef_wave_slots(row, 4 * q + (cL >> 3), cL & 7, N, lane,
                                                                            pos, src, a.pitch, a.c8d,
                                                                            a.cnd + a.cnd_off[N], slot, sstride);
                                    if (lane == L && fabs(oc) <= a.thr64[pl]) zerom |= 1u << bit;
                                }
                            }
                        }
                    }
                    int kept = 0;
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int k = 0; k < 16; ++k) {
                            const uint32_t b = 1u << (16 * q + k);
                            const bool zero = (nearm & b) ? (zerom & b) != 0 : fabsf(y[q][k]) <= thr;
                            if (zero) y[q][k] = 0.f;
                            else ++kept;
                        }
                    kept = wave_sum(kept);
                    w = kept < 1 ? 1.f : 1.f / (float)kept;
                    if constexpr (SPLIT) {
                        f32x16 vt[2];
#pragma unroll
                        for (int q = 0; q < 2; ++q) vt[q] = split_inverse_half(y[q], cnb, cnbl);
                        split_inverse_rest(vt, ka, kal, lane, xh);
                    } else {
                        mfma_inverse<T>(y, ka, lane, cnb, xh);
                    }
                } else {
                    // as filter_mfma_kernel: the guide's spectrum and the gains
                    // first, then the source's halves
                    f32x16 wg[2], vt[2];
                    float l2 = 0.f, l2b = 0.f;
                    {
                        v8 xg[4], xgl[4];
                        if constexpr (SPLIT) mfma_load_x_split(fgd, fp, fox, foy, pos, N, lane, xg, xgl, slot, sstride);
                        else mfma_load_x<T>(fgd, fp, fox, foy, pos, N, lane, xg, slot, sstride);
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            if constexpr (SPLIT) wg[q] = split_forward_half(xg, xgl, q, kb, kbl, lane, cna, cnal);
                            else wg[q] = mfma_forward_half<T>(xg, q, kb, lane, cna);
                            // column 32q + c = frequency (u, v): u = 4q + c/8, v = c%8
                            const float su =
                                sig2 * ((4 * q + (c >> 3)) == 0 ? 0.5f : 1.f) * ((c & 7) == 0 ? 0.5f : 1.f);
#pragma unroll
                            for (int k = 0; k < 16; ++k) {
                                const bool kz0 = k == 0 && h == 0;  // group-axis DC (row 0)
                                const float yy = wg[q][k] * wg[q][k];
                                // 1-ulp reciprocal: the operands already carry 16-bit rounding
                                const float wc = yy * __builtin_amdgcn_rcpf(yy + (kz0 ? 0.5f * su : su));
                                wg[q][k] = wc;
                                // rows 0..15 and 16..31 apart, then added (the order of
                                // the single-frame kernel, whose operands pair groups)
                                if (k < 8) l2 = fmaf(wc, wc, l2);
                                else l2b = fmaf(wc, wc, l2b);
                            }
                        }
                    }
                    // keep the source's loads after the gains (register pressure)
                    __builtin_amdgcn_sched_barrier(0);
                    {
                        v8 xs[4], xsl[4];
                        if constexpr (SPLIT) mfma_load_x_split(fsrc, fp, fox, foy, pos, N, lane, xs, xsl, slot, sstride);
                        else mfma_load_x<T>(fsrc, fp, fox, foy, pos, N, lane, xs, slot, sstride);
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            f32x16 yq;
                            if constexpr (SPLIT) yq = split_forward_half(xs, xsl, q, kb, kbl, lane, cna, cnal);
                            else yq = mfma_forward_half<T>(xs, q, kb, lane, cna);
#pragma unroll
                            for (int k = 0; k < 16; ++k) yq[k] *= wg[q][k];
                            if constexpr (SPLIT) vt[q] = split_inverse_half(yq, cnb, cnbl);
                            else vt[q] = mfma_inverse_half<T>(yq, cnb);
                        }
                    }
                    l2 = wave_sum(l2 + l2b);
                    w = l2 <= 0.f ? 1.f : fminf(1.f / l2, a.wmax);
                    if constexpr (SPLIT) split_inverse_rest(vt, ka, kal, lane, xh);
                    else mfma_inverse_rest<T>(vt, ka, lane, xh);
                }
                const double ws = fix_weight(w);
                if (c < N) {
                    const int pz = pos[c];
                    const int py = (pz >> 16) - oy, px = (pz & 0xFFFF) - ox + 4 * h;
                    u64 *accNum = slot_num(STACK ? slot[c] : 0);
                    u64 *accDen = accNum + fsz;
#pragma unroll
                    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                fix_add_scaled(&accNum[(py + 4 * rr + i) * FW + px + j], xh[rr][4 * i + j], ws);
                    if (h == 0) fix_add(&accDen[py * DW + (pz & 0xFFFF) - ox], w);
                }
            }
            __syncthreads();
        }
        for (int f = NF - rad - 1 < 0 ? 0 : NF - rad - 1; f < NF; ++f) flush_frame(pl, f);
    }
}

// ---------------------------------------------------------------------------
// num/den, then either the plane output (float, or int with RangeConvert
// rounding and clip, include/Conversion.hpp:241-256) or OPP -> RGB
// (MatrixConvert_YUV2RGB, include/Conversion.hpp:660-762) with clip, float or
// int.  FP64 arithmetic.  With den per block origin (den_box = 8), a pixel's
// den is the sum over the origins (oy, ox) with y - 8 < oy <= y and
// x - 8 < ox <= x: the workgroup sums each column of its row segment (and
// the 7 columns left of it) over those rows, 8 independent loads per column,
// then each pixel adds 8 column sums from LDS.  Integer sums: exact in any
// order.
template <bool BATCHED>
__global__ __launch_bounds__(256) void finalize_kernel(FinalizeArgs a0, const uint32_t *__restrict__ bmap,
                                                       const FinalizeArgs *__restrict__ bargs)
{
    // separate mul and add roundings, as the reference's double code
#pragma clang fp contract(off)
    constexpr int kCols = 256 + kBlock - 1;
    __shared__ u64 cols[3][kCols];
    int y;
    const FinalizeArgs &a = row_unit<BATCHED>(a0, bargs, bmap, y);
    const int x0 = blockIdx.x * 256;
    if (x0 >= a.W) return;  // the whole workgroup (batches launch the widest image's row)
    const int x = x0 + (int)threadIdx.x;
    const bool in = x < a.W;
    const size_t o = (size_t)y * a.pitch + x;
    if (a.den_box) {
        for (int c = 0; c < a.nplanes; ++c) {
            if (!a.active[c]) continue;
            const u64 *den = a.den[c];
            for (int j = threadIdx.x; j < kCols; j += 256) {
                const int ox = x0 - (kBlock - 1) + j;
                u64 sum = 0;
                if (ox >= 0 && ox <= a.W - kBlock) {
                    u64 v[kBlock];
#pragma unroll
                    for (int r = 0; r < kBlock; ++r) {
                        const int oy = y - (kBlock - 1) + r;
                        v[r] = oy >= 0 && oy <= a.H - kBlock ? den[(size_t)oy * a.pitch + ox] : 0;
                    }
#pragma unroll
                    for (int r = 0; r < kBlock; ++r) sum += v[r];
                }
                cols[c][j] = sum;
            }
        }
        __syncthreads();
    }
    if (!in) return;
    double val[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < a.nplanes; ++c) {
        if (!a.active[c]) {
            val[c] = (double)a.src[c][o];
            continue;
        }
        u64 d = 0;
        if (a.den_box) {
#pragma unroll
            for (int j = 0; j < kBlock; ++j) d += cols[c][threadIdx.x + j];
        } else {
            d = a.den[c][o];
        }
        val[c] = (double)(long long)a.num[c][o] / (double)(long long)d;
    }
    const size_t oo = (size_t)y * a.out_pitch + x;
    double res[3];
    int n = a.nplanes;
    if (a.to_rgb) {
        const double Y = val[0], U = val[1], V = val[2];
        const double Ry = 1.0 * a.gain_y, Ru = 1.0 * a.gain_c, Rv = (double)(2.0L / 3.0L) * a.gain_c;
        const double Gy = 1.0 * a.gain_y, Gv = (double)(-4.0L / 3.0L) * a.gain_c;
        const double By = 1.0 * a.gain_y, Bu = -1.0 * a.gain_c, Bv = (double)(2.0L / 3.0L) * a.gain_c;
        res[0] = Ry * Y + Ru * U + Rv * V + a.offset;
        res[1] = Gy * Y + Gv * V + a.offset;
        res[2] = By * Y + Bu * U + Bv * V + a.offset;
        n = 3;
    } else {
        for (int c = 0; c < n; ++c) res[c] = a.out_int ? val[c] * a.q_gain + a.offset : val[c];
    }
    for (int c = 0; c < n; ++c) {
        double v = res[c];
        if (a.clip) v = v < a.lo ? a.lo : (v > a.hi ? a.hi : v);
        if (a.out_int) static_cast<int *>(a.out[c])[oo] = (int)v;
        else static_cast<float *>(a.out[c])[oo] = (float)v;
    }
}

// MatrixConvert_RGB2YUV / ConvertToY, OPP (include/Conversion.hpp:418-432,
// 547-578) from float or int32 RGB, in FP64.
template <bool BATCHED>
__global__ void opp_kernel(OppArgs a0, const uint32_t *__restrict__ bmap, const OppArgs *__restrict__ bargs)
{
    // separate mul and add roundings, as the reference's double code
#pragma clang fp contract(off)
    int y;
    const OppArgs &a = row_unit<BATCHED>(a0, bargs, bmap, y);
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.W) return;
    const size_t i = (size_t)y * a.in_pitch + x;
    double R, G, B;
    if (a.in_int) {
        R = static_cast<const int *>(a.r)[i];
        G = static_cast<const int *>(a.g)[i];
        B = static_cast<const int *>(a.b)[i];
    } else {
        R = static_cast<const float *>(a.r)[i];
        G = static_cast<const float *>(a.g)[i];
        B = static_cast<const float *>(a.b)[i];
    }
    double Y = (R + G + B) * a.gainY + a.offsetY;
    if (a.clipY) Y = Y < 0.0 ? 0.0 : (Y > 1.0 ? 1.0 : Y);
    const size_t o = (size_t)y * a.pitch + x;
    if (a.y64) a.y64[(size_t)y * a.pitch64 + x] = Y;
    if (a.y) a.y[o] = (float)Y;
    if (a.u) a.u[o] = (float)((R - B) * a.gainU + a.offsetC);
    if (a.v) a.v[o] = (float)((R - G * 2.0 + B) * a.gainV + a.offsetC);
}

// MatrixConvert_YUV2RGB with the OPP matrix on float planes
// (include/Conversion.hpp:660-762), optional clip to [0, 1].
__global__ void opp_inverse_kernel(float *r, float *g, float *b, const float *y, const float *u, const float *v,
                                   int W, size_t pitch, int clip)
{
    // separate mul and add roundings, as the reference's double code
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= W) return;
    const size_t o = (size_t)blockIdx.y * pitch + x;
    const double Y = y[o], U = u[o], V = v[o];
    double res[3] = {Y + U + (double)(2.0L / 3.0L) * V, Y + (double)(-4.0L / 3.0L) * V,
                     Y - U + (double)(2.0L / 3.0L) * V};
    float *out[3] = {r, g, b};
    for (int c = 0; c < 3; ++c) {
        double t = res[c];
        if (clip) t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        out[c][o] = (float)t;
    }
}

__global__ void to_float_kernel(ToFloatArgs a)
{
    // separate mul and add roundings, as the reference's double code
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= a.W) return;
    double v = (double)a.src[(size_t)y * a.spitch + x] * a.gain + a.offset;
    if (a.clip) v = v < a.lo ? a.lo : (v > a.hi ? a.hi : v);
    a.dst[(size_t)y * a.dpitch + x] = (float)v;
    if (a.dst64) a.dst64[(size_t)y * a.dpitch + x] = v;
}

__global__ void fill_kernel(float *p, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void copy2d_kernel(float *dst, size_t dpitch, const float *src, size_t spitch, int W)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x < W) dst[(size_t)y * dpitch + x] = src[(size_t)y * spitch + x];
}

template <typename K>
hipError_t set_lds(K kernel, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename K>
hipError_t launch_tiled_m(K kernel, size_t lds, const MatchArgs &a, hipStream_t s, const BatchLaunch *b, dim3 grid)
{
    hipError_t e;
    if ((e = set_lds(kernel, lds)) != hipSuccess) return e;
    if (b) hipLaunchKernelGGL(kernel, dim3(b->units), dim3(kThreads), lds, s, a, b->map,
                              static_cast<const MatchArgs *>(b->args));
    else hipLaunchKernelGGL(kernel, grid, dim3(kThreads), lds, s, a, nullptr, nullptr);
    return hipGetLastError();
}

template <typename GT, typename DT, bool EXACT, bool ST1>
hipError_t launch_match_t(const MatchArgs &a, hipStream_t s, const BatchLaunch *b)
{
    const size_t lds = match_lds_bytes(a);
    hipError_t e;
    if (a.search) {
        if (b) return hipErrorInvalidValue;
        auto k = match_kernel<GT, DT, ST1, EXACT, false, true>;
        if ((e = set_lds(k, lds)) != hipSuccess) return e;
        const dim3 grid((a.ngx + a.trx - 1) / a.trx, (a.ngy + a.try_ - 1) / a.try_);
        hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a, nullptr, nullptr);
        return hipGetLastError();
    }
    if constexpr (!EXACT) {
        if (a.walk) {
            const dim3 grid((a.ngx + a.trx - 1) / a.trx, (a.ngy + a.try_ - 1) / a.try_);
            const int nm = 2 * (a.range / a.step) + 1;
            const bool one = nm * ((nm + 5) / 6) <= 64;    // one round of strips of 6
            const bool multi = nm * ((nm + 4) / 5) > 128;  // chunks of 128 strips of 5
            if (a.bstep == 3) {
                if (multi)
                    return b ? launch_tiled_m(match_walk_kernel<GT, ST1, true, 4, 2, 5, true>, lds, a, s, b, grid)
                             : launch_tiled_m(match_walk_kernel<GT, ST1, false, 4, 2, 5, true>, lds, a, s, b, grid);
                if (one)
                    return b ? launch_tiled_m(match_walk_kernel<GT, ST1, true, 4, 1, 6, false>, lds, a, s, b, grid)
                             : launch_tiled_m(match_walk_kernel<GT, ST1, false, 4, 1, 6, false>, lds, a, s, b, grid);
                return b ? launch_tiled_m(match_walk_kernel<GT, ST1, true, 4, 2, 5, false>, lds, a, s, b, grid)
                         : launch_tiled_m(match_walk_kernel<GT, ST1, false, 4, 2, 5, false>, lds, a, s, b, grid);
            }
            if (multi)
                return b ? launch_tiled_m(match_walk_kernel<GT, ST1, true, 3, 2, 5, true>, lds, a, s, b, grid)
                         : launch_tiled_m(match_walk_kernel<GT, ST1, false, 3, 2, 5, true>, lds, a, s, b, grid);
            if (one)
                return b ? launch_tiled_m(match_walk_kernel<GT, ST1, true, 3, 1, 6, false>, lds, a, s, b, grid)
                         : launch_tiled_m(match_walk_kernel<GT, ST1, false, 3, 1, 6, false>, lds, a, s, b, grid);
            return b ? launch_tiled_m(match_walk_kernel<GT, ST1, true, 3, 2, 5, false>, lds, a, s, b, grid)
                     : launch_tiled_m(match_walk_kernel<GT, ST1, false, 3, 2, 5, false>, lds, a, s, b, grid);
        }
    }
    if (b) {
        auto k = match_kernel<GT, DT, ST1, EXACT, true>;
        if ((e = set_lds(k, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3(b->units), dim3(kThreads), lds, s, a, b->map,
                           static_cast<const MatchArgs *>(b->args));
    } else {
        auto k = match_kernel<GT, DT, ST1, EXACT, false>;
        if ((e = set_lds(k, lds)) != hipSuccess) return e;
        const dim3 grid((a.ngx + a.trx - 1) / a.trx, (a.ngy + a.try_ - 1) / a.try_);
        hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a, nullptr, nullptr);
    }
    return hipGetLastError();
}

template <typename GT, typename DT, bool EXACT>
hipError_t launch_match_g(const MatchArgs &a, hipStream_t s, const BatchLaunch *b)
{
    return a.step == 1 ? launch_match_t<GT, DT, EXACT, true>(a, s, b)
                       : launch_match_t<GT, DT, EXACT, false>(a, s, b);
}

}  // namespace

// exact (the plain form): FP64 distances and keys; otherwise FP32
size_t match_lds_bytes(const MatchArgs &a)
{
    if (a.walk) {
        const int nm = 2 * (a.range / a.step) + 1;
        const bool multi = nm * ((nm + 4) / 5) > 128;
        return align16((size_t)a.tileH * a.tilePitch * 4) +
               (size_t)kWaves * (walk_wave_bytes() + (multi ? walk_prior_bytes(a.try_) : 0));
    }
    const int db = a.exact ? 8 : 4;
    const size_t rt = a.search ? align16((size_t)a.refH * a.refPitch * db) : 0;
    return align16((size_t)a.tileH * a.tilePitch * db) + rt + (size_t)kWaves * match_wave_bytes(a.maxCand, db);
}

size_t filter_lds_bytes(const FilterArgs &a, bool) { return filter_lds(a.FW, a.FH, a.G); }

hipError_t launch_match(const MatchArgs &a, bool guide64, hipStream_t s, const BatchLaunch *b)
{
    if (guide64)
        return a.exact ? launch_match_g<double, double, true>(a, s, b) : launch_match_g<double, float, false>(a, s, b);
    return a.exact ? launch_match_g<float, double, true>(a, s, b) : launch_match_g<float, float, false>(a, s, b);
}

template <typename K, typename A>
hipError_t launch_tiled(K kernel, unsigned threads, size_t lds, const A &a, hipStream_t s, const BatchLaunch *b,
                        dim3 grid)
{
    hipError_t e;
    if ((e = set_lds(kernel, lds)) != hipSuccess) return e;
    if (b) hipLaunchKernelGGL(kernel, dim3(b->units), dim3(threads), lds, s, a, b->map, static_cast<const A *>(b->args));
    else hipLaunchKernelGGL(kernel, grid, dim3(threads), lds, s, a, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_filter(const FilterArgs &a, bool wiener, hipStream_t s, const BatchLaunch *b)
{
    const size_t lds = filter_lds_bytes(a, wiener);
    const dim3 grid((a.ngx + a.trx - 1) / a.trx, (a.ngy + a.try_ - 1) / a.try_);
    if (wiener)
        return b ? launch_tiled(filter_kernel<true, true>, kThreads, lds, a, s, b, grid)
                 : launch_tiled(filter_kernel<true, false>, kThreads, lds, a, s, b, grid);
    return b ? launch_tiled(filter_kernel<false, true>, kThreads, lds, a, s, b, grid)
             : launch_tiled(filter_kernel<false, false>, kThreads, lds, a, s, b, grid);
}

size_t filter_mfma_lds_bytes(const FilterArgs &a, bool wiener)
{
    return filter_mfma_lds(a.FW, a.FH, wiener, 1, false, a.kbl != nullptr);
}

size_t filter_stack_lds_bytes(const FilterArgs &a, bool wiener, int precision)
{
    if (precision != 0) return filter_mfma_lds(a.FW, a.FH, wiener, a.nslots, true, precision == 3);
    return (filter_acc_floats(a.FW, a.FH) * a.nslots + (size_t)kWaves * a.G * ZS) * 4;
}

// LDS of one CU
constexpr size_t kLdsPerCu = 160 * 1024;

template <typename T, bool W>
hipError_t launch_mfma_stack_t(const FilterArgs &a, size_t lds, dim3 grid, hipStream_t s)
{
    // 16 waves where one workgroup fills a CU's LDS (as launch_mfma_t)
    if (2 * lds > kLdsPerCu)
        return launch_tiled(filter_mfma_stack_kernel<T, W, false, true, 2 * kMWaves>, 2 * kMThreads, lds, a, s,
                            nullptr, grid);
    return launch_tiled(filter_mfma_stack_kernel<T, W, false, true>, kMThreads, lds, a, s, nullptr, grid);
}

hipError_t launch_filter_stack(const FilterArgs &a, bool wiener, int precision, hipStream_t s)
{
    const size_t lds = filter_stack_lds_bytes(a, wiener, precision);
    const dim3 grid((a.ngx + a.trx - 1) / a.trx, (a.ngy + a.try_ - 1) / a.try_);
    if (precision == 0)
        return wiener ? launch_tiled(filter_stack_kernel<true, false, true>, kThreads, lds, a, s, nullptr, grid)
                      : launch_tiled(filter_stack_kernel<false, false, true>, kThreads, lds, a, s, nullptr, grid);
    if (precision == 1)
        return wiener ? launch_mfma_stack_t<_Float16, true>(a, lds, grid, s)
                      : launch_mfma_stack_t<_Float16, false>(a, lds, grid, s);
    if (precision == 3) {
        // split products (8-wave workgroups at 2 waves per SIMD)
        if (!a.kbl || !a.kal || !a.cnal || !a.cnbl) return hipErrorInvalidValue;
        return wiener ? launch_tiled(filter_mfma_stack_kernel<_Float16, true, false, true, kMWaves, true>, kMThreads,
                                     lds, a, s, nullptr, grid)
                      : launch_tiled(filter_mfma_stack_kernel<_Float16, false, false, true, kMWaves, true>, kMThreads,
                                     lds, a, s, nullptr, grid);
    }
    return wiener ? launch_mfma_stack_t<__bf16, true>(a, lds, grid, s) : launch_mfma_stack_t<__bf16, false>(a, lds, grid, s);
}

// 16-wave workgroups where only one workgroup fits a CU's 160 KB of LDS
// (large footprints: BMrange 19 at the default tile), so that a CU still
// holds 4 waves per SIMD, and for single launches of tiles of at least
// kMWideRefs reference blocks.  The workgroup size never changes results.
constexpr int kMWideRefs = 128;

template <typename T, bool W>
hipError_t launch_mfma_t(const FilterArgs &a, size_t lds, dim3 grid, hipStream_t s, const BatchLaunch *b)
{
    const bool wide = 2 * lds > kLdsPerCu || (!b && a.trx * a.try_ >= kMWideRefs);
    if (b)
        return wide ? launch_tiled(filter_mfma_kernel<T, W, true, 2 * kMWaves>, 2 * kMThreads, lds, a, s, b, grid)
                    : launch_tiled(filter_mfma_kernel<T, W, true>, kMThreads, lds, a, s, b, grid);
    return wide ? launch_tiled(filter_mfma_kernel<T, W, false, 2 * kMWaves>, 2 * kMThreads, lds, a, s, b, grid)
                : launch_tiled(filter_mfma_kernel<T, W, false>, kMThreads, lds, a, s, b, grid);
}

// The split (FP32_MFMA) form: 8-wave workgroups at 2 waves per SIMD, one
// workgroup per CU (its LDS holds both halves of the K64 fragments and of the
// staged footprint).
template <bool W>
hipError_t launch_mfma_split_t(const FilterArgs &a, size_t lds, dim3 grid, hipStream_t s, const BatchLaunch *b)
{
    if (b) return launch_tiled(filter_mfma_kernel<_Float16, W, true, kMWaves, true>, kMThreads, lds, a, s, b, grid);
    return launch_tiled(filter_mfma_kernel<_Float16, W, false, kMWaves, true>, kMThreads, lds, a, s, b, grid);
}

hipError_t launch_filter_mfma(const FilterArgs &a, bool wiener, int precision, hipStream_t s, const BatchLaunch *b)
{
    const size_t lds = filter_mfma_lds_bytes(a, wiener);
    const dim3 grid((a.ngx + a.trx - 1) / a.trx, (a.ngy + a.try_ - 1) / a.try_);
    if (precision == 3) {
        if (!a.kbl || !a.kal || !a.cnal || !a.cnbl || !a.cnapl || !a.cnbpl) return hipErrorInvalidValue;
        return wiener ? launch_mfma_split_t<true>(a, lds, grid, s, b) : launch_mfma_split_t<false>(a, lds, grid, s, b);
    }
    if (precision == 1)
        return wiener ? launch_mfma_t<_Float16, true>(a, lds, grid, s, b) : launch_mfma_t<_Float16, false>(a, lds, grid, s, b);
    return wiener ? launch_mfma_t<__bf16, true>(a, lds, grid, s, b) : launch_mfma_t<__bf16, false>(a, lds, grid, s, b);
}

template <typename K, typename A>
hipError_t launch_rows(K k1, K kb, const A &a, hipStream_t s, const BatchLaunch *b)
{
    if (b) hipLaunchKernelGGL(kb, dim3((b->maxW + 255) / 256, b->units), dim3(256), 0, s, a, b->map,
                              static_cast<const A *>(b->args));
    else hipLaunchKernelGGL(k1, dim3((a.W + 255) / 256, a.H), dim3(256), 0, s, a, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_finalize(const FinalizeArgs &a, hipStream_t s, const BatchLaunch *b)
{
    if (a.den_box != 0 && a.den_box != kBlock) return hipErrorInvalidValue;
    return launch_rows(finalize_kernel<false>, finalize_kernel<true>, a, s, b);
}

hipError_t launch_opp(const OppArgs &a, hipStream_t s, const BatchLaunch *b)
{
    return launch_rows(opp_kernel<false>, opp_kernel<true>, a, s, b);
}

hipError_t launch_opp_inverse(float *r, float *g, float *b, const float *y, const float *u, const float *v, int W,
                              int H, size_t pitch, int clip, hipStream_t s)
{
    hipLaunchKernelGGL(opp_inverse_kernel, dim3((W + 255) / 256, H), dim3(256), 0, s, r, g, b, y, u, v, W, pitch,
                       clip);
    return hipGetLastError();
}

hipError_t launch_to_float(const ToFloatArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(to_float_kernel, dim3((a.W + 255) / 256, a.H), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_copy2d(float *dst, size_t dpitch, const float *src, size_t spitch, int W, int H,
                         hipStream_t s)
{
    hipLaunchKernelGGL(copy2d_kernel, dim3((W + 255) / 256, H), dim3(256), 0, s, dst, dpitch, src, spitch, W);
    return hipGetLastError();
}

}  // namespace ispbm3d
