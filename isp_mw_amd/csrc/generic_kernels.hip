// Generic BM3D kernels: any BlockSize 1..kGenMaxBlock and GroupSize
// 1..kGenMaxGroup (G * B^2 <= kGenMaxElems), for the parameters the tuned
// kernels of bm3d_kernels.hip (BlockSize 8, GroupSize <= 32) do not take:
// profile vn's final step (BlockSize 11, include/BM3D.h:127-130), the CLI's
// -BS1 / -BS2 / -GS1 / -GS2 overrides (include/BM3D.h:416-452), and the
// reference's transform plans for every group length (source/BM3D.cpp:9-73).
//
//   match_generic_kernel  - BlockMatchingMulti (include/Block.h:578-711) with
//                           FP64 distances summed in the reference's row-major
//                           order, threshold, self first, the G-1 smallest by
//                           (key, scan index) from the shared radix selection.
//   filter_generic_kernel - one workgroup per group: gather, orthonormal 3D
//                           DCT-II in FP64 (separable line passes), hard
//                           threshold or Wiener shrinkage, inverse, and the
//                           weighted block added to the 64-bit fixed-point
//                           num / den planes with integer atomics
//                           (source/BM3D.cpp:339-385, 411-457;
//                           include/Block.h:1240-1357).
#include "bm3d_device.h"

namespace ispbm3d {
namespace {

// per wave: keys f64[maxCand], selk f64[64], idxs u32[maxCand],
// memb u32[maxCand], hist u32[256], seli u32[64]
__host__ __device__ inline size_t gen_match_wave_bytes(int maxCand)
{
    return align16((size_t)maxCand * 8 + 64 * 8 + (size_t)maxCand * 8 + 256 * 4 + 64 * 4);
}

template <typename GT>
__global__ __launch_bounds__(kThreads) void match_generic_kernel(MatchArgs a)
{
    typedef uint64_t KT;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int B = a.B;
    const GT *__restrict__ guide = static_cast<const GT *>(a.guide);
    const int trx = a.trx, tnum = a.trx * a.try_;
    const int tx0 = blockIdx.x * trx, ty0 = blockIdx.y * a.try_;
    const int rx0 = grid_pos(tx0, a.bstep, a.W - B), ry0 = grid_pos(ty0, a.bstep, a.H - B);
    const int ox = rx0 - a.range, oy = ry0 - a.range;

    double *tile = reinterpret_cast<double *>(smem);
    const size_t tileBytes = align16((size_t)a.tileH * a.tilePitch * 8);
    unsigned char *wb = smem + tileBytes + (size_t)wave * gen_match_wave_bytes(a.maxCand);
    KT *keys = reinterpret_cast<KT *>(wb);
    KT *selk = keys + a.maxCand;
    uint32_t *idxs = reinterpret_cast<uint32_t *>(selk + 64);
    uint32_t *memb = idxs + a.maxCand;
    uint32_t *hist = memb + a.maxCand;
    uint32_t *seli = hist + 256;

    for (int k = tid; k < a.tileH * a.tileW; k += kThreads) {
        const int yy0 = k / a.tileW, xx0 = k - yy0 * a.tileW;
        const int yy = oy + yy0, xx = ox + xx0;
        double v = 0;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) v = (double)guide[(size_t)yy * a.gpitch + xx];
        tile[yy0 * a.tilePitch + xx0] = v;
    }
    __syncthreads();

    const int tp = a.tilePitch;
    for (int r = wave; r < tnum; r += kWaves) {
        const int gxi = tx0 + (r % trx), gyi = ty0 + (r / trx);
        if (gxi >= a.ngx || gyi >= a.ngy) continue;
        const int i = grid_pos(gxi, a.bstep, a.W - B), j = grid_pos(gyi, a.bstep, a.H - B);
        const size_t refIdx = (size_t)gyi * a.ngx + gxi;
        const size_t base = refIdx * a.G;
        if (a.bypass) {
            if (lane == 0) {
                a.cnt_out[refIdx] = 1;
                a.pos_out[base] = (j << 16) | i;
                if (a.key_out) a.key_out[base] = 0.0;
            }
            continue;
        }
        const int st = a.step;
        const int l = search_boundary(i, 0, a.range, st);
        const int rr = search_boundary(i, a.W - B, a.range, st);
        const int t = search_boundary(j, 0, a.range, st);
        const int b = search_boundary(j, a.H - B, a.range, st);
        const int nx = (rr - l) / st + 1, ny = (b - t) / st + 1;
        const int selfx = (i - l) / st, selfy = (j - t) / st;
        const int nsx = (nx + kStrip - 1) / kStrip, nstrips = nsx * ny;
        const double *tref = tile + (j - oy) * tp + (i - ox);
        const double thSSE = a.thSSE, distMul = a.distMul;

        // as match_kernel: the list is reduced to its K best before a batch
        // of strips that could overflow it; kbound skips later candidates
        // that rank behind all K of them
        const uint32_t K = (uint32_t)(a.G - 1);
        const bool may_fill = (uint32_t)a.maxCand < (uint32_t)(nx * ny);
        uint32_t cnt = 0;
        KT kbound = ~(KT)0;
        for (int s0 = 0; s0 < nstrips; s0 += 64) {
            if (may_fill && cnt + 64 * kStrip > (uint32_t)a.maxCand) {
                wave_sync();
                const uint32_t nk = select_smallest<KT, 64>(keys, idxs, cnt, K, selk, seli, hist, memb, lane);
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
            double acc[kStrip];
#pragma unroll
            for (int c = 0; c < kStrip; ++c) acc[c] = 0;
            if (valid) {
                const double *tc = tile + (t - oy + ky * st) * tp + (l - ox + kx0 * st);
                // one sum over the B^2 terms in row-major order, each square
                // rounded before it is added (include/Block.h:589-605)
                for (int u = 0; u < B; ++u) {
                    const double *ru = tref + u * tp, *cu = tc + u * tp;
                    for (int v = 0; v < B; ++v) {
                        const double rv = ru[v];
#pragma unroll
                        for (int c = 0; c < kStrip; ++c) acc[c] = sq_acc_exact(rv - cu[c * st + v], acc[c]);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < kStrip; ++c) {
                const int kx = kx0 + c;
                const KT key = KeyOf<double>::key(acc[c] * distMul);
                const bool ok = valid && kx < nx && !(kx == selfx && ky == selfy) && acc[c] <= thSSE &&
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
        const uint32_t nsel = select_smallest<KT, 64>(keys, idxs, cnt, K, selk, seli, hist, memb, lane);
        if (lane == 0) {
            a.cnt_out[refIdx] = (int)(1 + nsel);
            a.pos_out[base] = (j << 16) | i;
            if (a.key_out) a.key_out[base] = 0.0;
        }
        if ((uint32_t)lane < nsel) {
            const KT kq = selk[lane];
            const uint32_t iq = seli[lane];
            uint32_t rank = 0;
            for (uint32_t q = 0; q < nsel; ++q) {
                const KT k2 = selk[q];
                rank += (k2 < kq) || (k2 == kq && seli[q] < iq);
            }
            const int ky = (int)(iq / (uint32_t)nx), kx = (int)(iq - (uint32_t)ky * nx);
            const int y = t + ky * st, x = l + kx * st;
            a.pos_out[base + 1 + rank] = (y << 16) | x;
            if (a.key_out) a.key_out[base + 1 + rank] = KeyOf<double>::value(kq);
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Collaborative filter of one group per workgroup.

constexpr int kGenThreads = 256;

// Coefficient tables through the constant address space: every lane of a
// wave reads the same entry at the same time, so these become scalar loads.
typedef const __attribute__((address_space(4))) double cdouble;
__device__ __forceinline__ cdouble *kconst64(const double *p) { return (cdouble *)(p); }

// One separable pass over lines of length L (L <= LMAX) of the group in
// LDS, in place: line q starts at line_base(q) with element stride `es`.
// Forward: y[k] = sum_n C[k][n] x[n]; inverse: x[n] = sum_k C[k][n] y[k].
template <int LMAX, typename Base>
__device__ __forceinline__ void line_pass(double *g, int nlines, int L, int es, Base line_base, cdouble *C,
                                          bool inverse)
{
    for (int q = threadIdx.x; q < nlines; q += kGenThreads) {
        double *p = g + line_base(q);
        double t[LMAX];
#pragma unroll
        for (int n = 0; n < LMAX; ++n)
            if (n < L) t[n] = p[n * es];
        for (int k = 0; k < L; ++k) {
            double s = 0;
#pragma unroll
            for (int n = 0; n < LMAX; ++n)
                if (n < L) s = fma(inverse ? C[n * L + k] : C[k * L + n], t[n], s);
            p[k * es] = s;
        }
    }
}

// 3D orthonormal DCT-II (or its inverse) of g[z][y][x], N x B x B.
__device__ __forceinline__ void dct3d(double *g, int N, int B, cdouble *cb, cdouble *cn, bool inverse)
{
    const int B2 = B * B;
    auto rows = [=](int q) { return q * B; };                                    // (z, y): x varies
    auto cols = [=](int q) { return (q / B) * B2 + (q % B); };                   // (z, x): y varies
    auto tube = [=](int q) { return q; };                                        // (y, x): z varies
    if (!inverse) {
        line_pass<kGenMaxBlock>(g, N * B, B, 1, rows, cb, false);
        __syncthreads();
        line_pass<kGenMaxBlock>(g, N * B, B, B, cols, cb, false);
        __syncthreads();
        line_pass<kGenMaxGroup>(g, B2, N, B2, tube, cn, false);
        __syncthreads();
    } else {
        line_pass<kGenMaxGroup>(g, B2, N, B2, tube, cn, true);
        __syncthreads();
        line_pass<kGenMaxBlock>(g, N * B, B, B, cols, cb, true);
        __syncthreads();
        line_pass<kGenMaxBlock>(g, N * B, B, 1, rows, cb, true);
        __syncthreads();
    }
}

// Sum over the workgroup (every thread gets the total).
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    T s = 0;
    for (int w = 0; w < kGenThreads / 64; ++w) s += red[w];
    return s;
}

// floor(v * 2^32) added to a 64-bit fixed-point accumulator in HBM.
__device__ __forceinline__ void fix_add_global(u64 *acc, double v)
{
    atomicAdd(acc, (u64)__double2ll_rd(v * (double)kFixScale));
}

template <bool WIENER>
__global__ __launch_bounds__(kGenThreads) void filter_generic_kernel(GenFilterArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int refIdx = blockIdx.x;
    if (refIdx >= a.ngx * a.ngy) return;
    const int B = a.B, B2 = B * B;
    int N = a.cnt[refIdx];
    if (a.haar) N = 1 << (31 - __clz(N));  // the largest power of two of the members
    const int E = N * B2;
    const int *pos = a.pos + (size_t)refIdx * a.G;
    double *X = reinterpret_cast<double *>(smem);
    double *R = X + (size_t)a.G * B2;  // Wiener guide group
    double *red = (WIENER ? R : X) + (size_t)a.G * B2;
    int *ired = reinterpret_cast<int *>(red);
    cdouble *cb = kconst64(a.cb);
    cdouble *cn = kconst64(a.cn + (size_t)(N - 1) * N * (2 * N - 1) / 6);

    for (int c = 0; c < 3; ++c) {
        if (!a.active[c]) continue;
        const float *src = a.src[c];
        for (int e = threadIdx.x; e < E; e += kGenThreads) {
            const int z = e / B2, yx = e - z * B2, y = yx / B, x = yx - y * B;
            const int pz = pos[z];
            const size_t o = (size_t)((pz >> 16) + y) * a.pitch + (pz & 0xFFFF) + x;
            X[e] = (double)src[o];
            if (WIENER) R[e] = (double)a.guide[c][o];
        }
        __syncthreads();
        dct3d(X, N, B, cb, cn, false);
        if (WIENER) dct3d(R, N, B, cb, cn, false);
        double weight;
        if (!WIENER) {
            // zero |O| <= lambda sigma, count the survivors (source/BM3D.cpp:361-371)
            const double thr = a.thr[c];
            int kept = 0;
            for (int e = threadIdx.x; e < E; e += kGenThreads) {
                const double v = X[e];
                if (fabs(v) <= thr) X[e] = 0.0;
                else ++kept;
            }
            kept = block_sum(kept, ired);
            weight = kept < 1 ? 1.0 : 1.0 / (double)kept;
        } else {
            // W = O_r^2 / (O_r^2 + sigma^2 / 2^zeta) (source/BM3D.cpp:437-443)
            const double s2 = a.sig2[c];
            double l2 = 0;
            for (int e = threadIdx.x; e < E; e += kGenThreads) {
                const int z = e / B2, yx = e - z * B2, y = yx / B, x = yx - y * B;
                const int zeta = (z == 0) + (y == 0) + (x == 0);
                const double rr = R[e] * R[e];
                const double wc = rr / (rr + s2 / (double)(1 << zeta));
                X[e] *= wc;
                l2 = fma(wc, wc, l2);
            }
            l2 = block_sum(l2, red);
            weight = l2 <= 0 ? 1.0 : fmin(1.0 / l2, (double)a.wmax);
        }
        __syncthreads();
        dct3d(X, N, B, cb, cn, true);
        u64 *num = a.num[c], *den = a.den[c];
        const u64 wfix = (u64)__double2ll_rd(weight * (double)kFixScale);
        for (int e = threadIdx.x; e < E; e += kGenThreads) {
            const int z = e / B2, yx = e - z * B2, y = yx / B, x = yx - y * B;
            const int pz = pos[z];
            const size_t o = (size_t)((pz >> 16) + y) * a.pitch + (pz & 0xFFFF) + x;
            if (a.kaiser) {
                const double wk = weight * a.kaiser[yx];
                fix_add_global(num + o, wk * X[e]);
                fix_add_global(den + o, wk);
            } else {
                fix_add_global(num + o, weight * X[e]);
                atomicAdd(den + o, wfix);
            }
        }
        __syncthreads();
    }
}

template <typename K>
hipError_t set_lds_g(K kernel, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

}  // namespace

size_t match_generic_lds_bytes(const MatchArgs &a)
{
    return align16((size_t)a.tileH * a.tilePitch * 8) + (size_t)kWaves * gen_match_wave_bytes(a.maxCand);
}

hipError_t launch_match_generic(const MatchArgs &a, bool guide64, hipStream_t s)
{
    const size_t lds = match_generic_lds_bytes(a);
    const dim3 grid((a.ngx + a.trx - 1) / a.trx, (a.ngy + a.try_ - 1) / a.try_);
    hipError_t e;
    if (guide64) {
        auto k = match_generic_kernel<double>;
        if ((e = set_lds_g(k, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a);
    } else {
        auto k = match_generic_kernel<float>;
        if ((e = set_lds_g(k, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a);
    }
    return hipGetLastError();
}

size_t filter_generic_lds_bytes(const GenFilterArgs &a, bool wiener)
{
    return (size_t)(wiener ? 2 : 1) * a.G * a.B * a.B * 8 + 64;
}

hipError_t launch_filter_generic(const GenFilterArgs &a, bool wiener, hipStream_t s)
{
    const size_t lds = filter_generic_lds_bytes(a, wiener);
    const unsigned refs = (unsigned)(a.ngx * a.ngy);
    hipError_t e;
    if (wiener) {
        auto k = filter_generic_kernel<true>;
        if ((e = set_lds_g(k, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3(refs), dim3(kGenThreads), lds, s, a);
    } else {
        auto k = filter_generic_kernel<false>;
        if ((e = set_lds_g(k, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3(refs), dim3(kGenThreads), lds, s, a);
    }
    return hipGetLastError();
}

}  // namespace ispbm3d
