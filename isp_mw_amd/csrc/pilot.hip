// Learned pilot of the BM3D Wiener step (gfx950): 3x3 convolutions on BF16
// NHWC activations with FP32 accumulation on the matrix cores, their
// gradients, Adam, and the input / output / loss stages of a residual CNN.
// No counterpart in ISP_MW: the pilot stands in for BM3D_Basic's estimate
// (source/BM3D.cpp:339-385) as the guide of BM3D_Final (source/BM3D.cpp:411-457).
//
// Conventions (all tensors device memory, row-major):
//   activations  [N][H][W][C] bf16, C a multiple of 16 (zero padded channels)
//   weights      FP32 master [cout][cin][9] (tap = ky*3 + kx, cross-correlation:
//                out[y][x][co] = sum w[co][ci][ky*3+kx] * in[y+ky-1][x+kx-1][ci],
//                zero outside the image)
//   packed       forward  wf[COUT][KP]  k = tap*CIN + ci    (bf16)
//                backward wb[CIN][KPb]  k = tap'*COUT + co, tap' = 8 - tap
//                (the input gradient is the same convolution of the output
//                gradient with the flipped, transposed kernel)
//   KP = 9*CIN rounded up to 32, the K step of v_mfma_f32_16x16x32_bf16.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/isp_pilot.h"

namespace pilot {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// output tile of a convolution workgroup: 8 x 32 pixels keeps the 48-channel
// kernel's LDS (77,184 bytes) within half of the 160 KB of a CU,
// so two workgroups overlap their staging with each other's products
constexpr int TH = 8, TW = 32;
constexpr int RPW = TH / 4;              // rows per wave
constexpr int MT = RPW * (TW / 16);      // 16-pixel M-tiles per wave
constexpr int SH = TH + 2, SW = TW + 2;  // staged input tile with halo
constexpr int THREADS = 256;             // 4 waves x RPW rows
// pixel tile of the weight-gradient kernel: 4 rows of 32 (one 32-pixel row per
// MFMA K chunk), staged as 6 rows with the halo
constexpr int WTH = 4, WTW = 32, WPIX = WTH * WTW, WSR = WTH + 2;
// channel pitches (elements) of the shifted input copies and of the transposed
// output gradient: odd multiples of 32 bytes (208 and 144 elements), so the
// 16-byte operand reads of half a wave (16 channels x 2 pixel groups) are
// conflict-free (see in_pitch)
constexpr int WCP = WSR * WTW + 16;
constexpr int WZP = WPIX + 16;
constexpr int WBLOCKS = 512;        // persistent grid of the weight gradient (two workgroups per CU)

__host__ __device__ constexpr int kpad(int cin) { return (9 * cin + 31) / 32 * 32; }
// LDS row pitches (elements) of the staged input tile (one row per pixel) and
// of the weights (one row per output channel).  A 16-byte LDS read is served
// per half-wave (32 lanes) from a 512-byte bank window, so the 16 pixels x 2
// K-groups of 8 channels that half a wave reads are conflict-free when a row
// is an odd multiple of 32 bytes: the input rows are unpadded (CIN = 16 or
// 48), the weight rows KP + 16 (KP is a multiple of 32).  The staging writes
// of the input tile are then contiguous.
__host__ __device__ constexpr int in_pitch(int cin) { return cin; }
__host__ __device__ constexpr int w_pitch(int cin) { return kpad(cin) + 16; }

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

// keep the 8 values of v whose mask value is > 0 (ReLU derivative of a
// post-ReLU activation), zero the others
__device__ __forceinline__ uint4 apply_mask(uint4 v, uint4 m)
{
    uint32_t vv[4] = {v.x, v.y, v.z, v.w}, mm[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t lo = mm[i] & 0xFFFFu, hi = mm[i] >> 16;
        const uint32_t keep_lo = (lo != 0 && !(lo & 0x8000u)) ? 0xFFFFu : 0u;
        const uint32_t keep_hi = (hi != 0 && !(hi & 0x8000u)) ? 0xFFFF0000u : 0u;
        vv[i] &= keep_lo | keep_hi;
    }
    return make_uint4(vv[0], vv[1], vv[2], vv[3]);
}

// One output tile of TH x TW pixels and all COUT channels per workgroup; wave
// w computes rows RPW*w .. RPW*w+RPW-1 (MT M-tiles of 16 pixels) against
// COUT/16 N-tiles.  LDS row pitches: in_pitch, w_pitch.
template <int CIN, int COUT>
__global__ __launch_bounds__(THREADS) void conv3x3_kernel(const uint16_t *__restrict__ in,
                                                          const uint16_t *__restrict__ mask,
                                                          const uint16_t *__restrict__ wts,
                                                          const float *__restrict__ bias, uint16_t *__restrict__ out,
                                                          float *__restrict__ out32, int H, int W, int relu, int N)
{
    constexpr int KP = kpad(CIN), KC = KP / 32, NT = COUT / 16;
    constexpr int SP = in_pitch(CIN), KS = w_pitch(CIN), CQ = CIN / 8;
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t *tin = reinterpret_cast<uint16_t *>(smem);  // [SH*SW][SP]
    uint16_t *tw = tin + SH * SW * SP;                    // [COUT][KS]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // persistent: the weights are staged once, then the workgroup walks the
    // tiles blockIdx.x, blockIdx.x + gridDim.x, ... (neighbouring workgroups
    // on neighbouring tiles)
    for (int k = tid; k < COUT * (KP / 8); k += THREADS) {
        const int row = k / (KP / 8), q = k - row * (KP / 8);
        *reinterpret_cast<uint4 *>(tw + (size_t)row * KS + 8 * q) =
            *reinterpret_cast<const uint4 *>(wts + (size_t)row * KP + 8 * q);
    }
    const int ntx = (W + TW - 1) / TW, nty = (H + TH - 1) / TH, per = ntx * nty;
    for (int t = blockIdx.x; t < per * N; t += gridDim.x) {
        const int n = t / per, tr = t - n * per, tyi = tr / ntx;
        const int x0 = (tr - tyi * ntx) * TW, y0 = tyi * TH;
        __syncthreads();  // the previous tile's reads of tin are done

        for (int k = tid; k < SH * SW * CQ; k += THREADS) {
            const int q = k % CQ, p = k / CQ, sx = p % SW, sy = p / SW;
            const int yy = y0 + sy - 1, xx = x0 + sx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const size_t o = (((size_t)n * H + yy) * W + xx) * CIN + 8 * q;
                v = *reinterpret_cast<const uint4 *>(in + o);
                if (mask) v = apply_mask(v, *reinterpret_cast<const uint4 *>(mask + o));
            }
            *reinterpret_cast<uint4 *>(tin + (size_t)p * SP + 8 * q) = v;
        }
        __syncthreads();

        const int m = lane & 15, g = lane >> 4;
        f32x4 acc[MT][NT];
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};

        for (int kc = 0; kc < KC; ++kc) {
            const int k = kc * 32 + 8 * g;
            int tap = k / CIN;
            const int ci0 = k - tap * CIN;  // in [0, CIN): the K padding reads tap 8, whose weights are 0
            if (tap > 8) tap = 8;
            const int ty = tap / 3, tx = tap - 3 * ty;
            bf16x8 b[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                b[nt] = *reinterpret_cast<const bf16x8 *>(tw + (size_t)(16 * nt + m) * KS + k);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int r = RPW * wave + (mt >> 1), xs = (mt & 1) * 16 + m;
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(tin + ((size_t)(r + ty) * SW + xs + tx) * SP + ci0);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[mt][nt], 0, 0, 0);
            }
        }
        // D layout: lane holds rows 4g..4g+3 (pixels) of column lane % 16 (channel)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int r = RPW * wave + (mt >> 1), yy = y0 + r;
            if (yy >= H) continue;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int co = 16 * nt + m;
                const float bv = bias ? bias[co] : 0.f;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int xx = x0 + (mt & 1) * 16 + 4 * g + i;
                    if (xx >= W) continue;
                    float v = acc[mt][nt][i] + bv;
                    if (relu) v = fmaxf(v, 0.f);
                    const size_t o = (((size_t)n * H + yy) * W + xx) * COUT + co;
                    out[o] = f2bf(v);
                    if (out32) out32[o] = v;
                }
            }
        }
    }
}

// Inference form of conv3x3_kernel (no ReLU mask, no FP32 copy of the
// outputs).  The products run as D[co][pixel] = W X, the weights being the A
// operand, so that a lane's four accumulator values are four consecutive
// output channels of one pixel: after bias, ReLU and the bf16 rounding (one
// v_cvt_pk_bf16_f32 per pair) they leave as one 8-byte store, and the three
// stores of a pixel's 48 channels fill its 96 bytes.  The operand fragments
// of K chunk kc + 1 are read from LDS while chunk kc's products run, and the
// next tile's input is loaded into registers during the whole tile.  Every
// thread's tile-relative offsets are computed once; tiles clear of the image
// border use them without bounds tests.  The shared area holds the staged
// input tile (with RESID, then the FP32 residual tile):
// max(SH*SW*SP, TH*TW*16 FP32 as bf16 elements).
template <int CIN, int COUT>
__host__ __device__ constexpr int fwd_area()
{
    return SH * SW * in_pitch(CIN) > TH * TW * COUT ? SH * SW * in_pitch(CIN) : TH * TW * COUT;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack_bf16(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));  // round to nearest even
}

// RESID: the network's last layer at inference.  Instead of the bf16 output
// tile, output channels c < creal (the residual of plane c) are staged in LDS
// as FP32 [c][TH*TW] and leave as est[n][c][y][x] = noisy[n][c][y][x] -
// (conv + bias), planar FP32, consecutive lanes on consecutive pixels.
template <int CIN, int COUT, bool RESID = false>
__global__ __launch_bounds__(THREADS, 2) void conv3x3_fwd_kernel(const uint16_t *__restrict__ in,
                                                                 const uint16_t *__restrict__ wts,
                                                                 const float *__restrict__ bias,
                                                                 uint16_t *__restrict__ out, int H, int W, int relu,
                                                                 int N, const float *__restrict__ noisy = nullptr,
                                                                 float *__restrict__ est = nullptr, int creal = 0)
{
    constexpr int KP = kpad(CIN), KC = KP / 32, NT = COUT / 16;
    constexpr int SP = in_pitch(CIN), KS = w_pitch(CIN), CQ = CIN / 8;
    constexpr int NIN = SH * SW * CQ, NPF = (NIN + THREADS - 1) / THREADS;
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t *tin = reinterpret_cast<uint16_t *>(smem);  // [SH*SW][SP]
    uint16_t *tw = tin + fwd_area<CIN, COUT>();          // [COUT][KS]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, g = lane >> 4;
    for (int k = tid; k < COUT * (KP / 8); k += THREADS) {
        const int row = k / (KP / 8), q = k - row * (KP / 8);
        *reinterpret_cast<uint4 *>(tw + (size_t)row * KS + 8 * q) =
            *reinterpret_cast<const uint4 *>(wts + (size_t)row * KP + 8 * q);
    }
    // this lane's output channels 16 nt + 4 g + i
    float bv[NT][4];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) bv[nt][i] = bias ? bias[16 * nt + 4 * g + i] : 0.f;
    // the 16-byte input pieces of this thread: k = tid + i THREADS is piece
    // k % CQ of halo pixel k / CQ at (sy, sx); ioff is its offset from the
    // tile's top-left halo pixel, ipos packs sy << 8 | sx of pieces 2j, 2j + 1
    // in its low and high halves
    int ioff[NPF], ipos[(NPF + 1) / 2];
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
        const int k = tid + i * THREADS, p = k / CQ, sy = p / SW, sx = p - sy * SW;
        ioff[i] = (sy * W + sx) * CIN + 8 * (k - p * CQ);
        if (i & 1) ipos[i >> 1] |= (sy << 8 | sx) << 16;
        else ipos[i >> 1] = sy << 8 | sx;
    }
    // the last round's pieces beyond the tile are neither loaded nor staged
    const bool last_in = NIN % THREADS == 0 || tid + (NPF - 1) * THREADS < NIN;
    // output pixel of M-tile mt: row RPW wave + mt / 2, column 16 (mt % 2) + m;
    // ooff is its offset (elements) from the tile's top-left pixel, channel 4 g
    int ooff[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) ooff[mt] = ((RPW * wave + (mt >> 1)) * W + (mt & 1) * 16 + m) * COUT + 4 * g;
    const int ntx = (W + TW - 1) / TW, nty = (H + TH - 1) / TH, per = ntx * nty, total = per * N;

    // the input tile (with halo, zeros outside the image) of tile t into registers
    uint4 pf[NPF];
    auto fetch = [&](int t) {
        const int n = t / per, tr = t - n * per, tyi = tr / ntx;
        const int x0 = (tr - tyi * ntx) * TW, y0 = tyi * TH;
        // element index of the top-left halo pixel (negative on the first row or column)
        const long long base = (((long long)n * H + y0 - 1) * W + x0 - 1) * CIN;
        const bool inner = y0 >= 1 && y0 + SH - 1 <= H && x0 >= 1 && x0 + SW - 1 <= W;
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
            bool ok = i < NPF - 1 || last_in;
            if (!inner) {
                const int pos = (ipos[i >> 1] >> (16 * (i & 1))) & 0xFFFF;
                const int yy = y0 + (pos >> 8) - 1, xx = x0 + (pos & 255) - 1;
                ok = ok && yy >= 0 && yy < H && xx >= 0 && xx < W;
            }
            pf[i] = ok ? *reinterpret_cast<const uint4 *>(in + (base + ioff[i])) : make_uint4(0, 0, 0, 0);
        }
    };
    if ((int)blockIdx.x < total) fetch(blockIdx.x);

    // operand fragments of K chunk kc: weights (A, rows 16 nt + m) and input
    // pixels (B, columns m of the M-tiles), K elements 8 g .. 8 g + 7
    auto frags = [&](int kc, bf16x8 *fw, bf16x8 *fx) {
        const int k = kc * 32 + 8 * g;
        int tap = k / CIN;
        const int ci0 = k - tap * CIN;  // in [0, CIN): the K padding reads tap 8, whose weights are 0
        if (tap > 8) tap = 8;
        const int ty = tap / 3, tx = tap - 3 * ty;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
            fw[nt] = *reinterpret_cast<const bf16x8 *>(tw + (size_t)(16 * nt + m) * KS + k);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int r = RPW * wave + (mt >> 1), xs = (mt & 1) * 16 + m;
            fx[mt] = *reinterpret_cast<const bf16x8 *>(tin + ((size_t)(r + ty) * SW + xs + tx) * SP + ci0);
        }
    };

    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int n = t / per, tr = t - n * per, tyi = tr / ntx;
        const int x0 = (tr - tyi * ntx) * TW, y0 = tyi * TH;
        __syncthreads();  // every wave is done with the previous tile's shared area
#pragma unroll
        for (int i = 0; i < NPF; ++i)
            if (i < NPF - 1 || last_in) {
                const int k = tid + i * THREADS;
                *reinterpret_cast<uint4 *>(tin + (size_t)(k / CQ) * SP + 8 * (k % CQ)) = pf[i];
            }
        __syncthreads();
        if (t + (int)gridDim.x < total) fetch(t + gridDim.x);  // in flight during the products

        f32x4 acc[MT][NT];
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
        bf16x8 fw[2][NT], fx[2][MT];
        frags(0, fw[0], fx[0]);
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
            if (kc + 1 < KC) frags(kc + 1, fw[(kc + 1) & 1], fx[(kc + 1) & 1]);
            // the reads stay ahead of this chunk's products (the scheduler
            // would otherwise sink them next to their use)
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw[kc & 1][nt], fx[kc & 1][mt], acc[mt][nt], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        // D layout: lane holds rows 4g..4g+3 (channels) of column lane % 16 (pixel)
        if constexpr (RESID) {
            __syncthreads();  // every wave is done reading the input tile
            float *tres = reinterpret_cast<float *>(tin);  // [creal][TH*TW]
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int pix = (RPW * wave + (mt >> 1)) * TW + (mt & 1) * 16 + m;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int co = 16 * nt + 4 * g + i;
                        if (co < creal) tres[co * (TH * TW) + pix] = acc[mt][nt][i] + bv[nt][i];
                    }
            }
            __syncthreads();
            for (int k = tid; k < creal * TH * TW; k += THREADS) {
                const int c = k / (TH * TW), pix = k - c * (TH * TW), r = pix / TW, px = pix - r * TW;
                const int yy = y0 + r, xx = x0 + px;
                if (yy < H && xx < W) {
                    const size_t o = (((size_t)n * creal + c) * H + yy) * W + xx;
                    est[o] = noisy[o] - tres[k];
                }
            }
            continue;
        }
        uint16_t *dst = out + (((size_t)n * H + y0) * W + x0) * COUT;
        const bool whole = y0 + TH <= H && x0 + TW <= W;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            if (!whole && (y0 + RPW * wave + (mt >> 1) >= H || x0 + (mt & 1) * 16 + m >= W)) continue;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    v[i] = acc[mt][nt][i] + bv[nt][i];
                    if (relu) v[i] = fmaxf(v[i], 0.f);
                }
                *reinterpret_cast<uint2 *>(dst + ooff[mt] + 16 * nt) =
                    make_uint2(pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]));
            }
        }
    }
}

// Weight and bias gradients on the matrix cores:
//   dw[co][tap][ci] = sum_p dz[p][co] * in[p + off(tap)][ci],  db[co] = sum_p dz[p][co],
// with dz = dout * (mask > 0) (mask may be null), as the product D = A B with
// A[co][p] = dz (M = output channels), B[p][(tap, ci)] = the shifted input
// (N = tap * CINP + ci, the packed forward weights' K order) and the pixels as
// the reduction axis, 32 per v_mfma_f32_16x16x32_bf16 (one tile row).  A
// workgroup stages a 4 x 32 pixel tile: dz transposed to [co][p], and the
// input (with its 1-pixel halo) transposed to [ci][row][x] in three copies
// shifted by kx = 0, 1, 2, so that every operand fragment (8 consecutive
// pixels of one channel) is one aligned 16-byte read.  The bias gradient is
// the same product with B = 1.  A persistent grid walks the tiles; workgroup
// b writes its sums to partial[b] ([cout][9][cin], then the cout bias sums),
// added over the workgroups in a fixed order by wgrad_sum_kernel.
template <int CINP, int COUTP>
__host__ __device__ constexpr size_t wgrad_lds_bytes()
{
    return ((size_t)3 * CINP * WCP + (size_t)COUTP * WZP) * 2;
}

template <int CINP, int COUTP>
__global__ __launch_bounds__(THREADS) void wgrad_mfma_kernel(const uint16_t *__restrict__ in,
                                                             const uint16_t *__restrict__ dout,
                                                             const uint16_t *__restrict__ mask,
                                                             float *__restrict__ partial, int N, int H, int W, int cin,
                                                             int cout)
{
    constexpr int MT = COUTP / 16;          // 16-channel tiles of the outputs
    constexpr int NTT = 9 * CINP / 16;      // 16-wide tiles of (tap, ci); one tap each (CINP % 16 == 0)
    constexpr int NTW = (NTT + 3) / 4;      // per wave (wave w takes tiles w, w + 4, ...)
    constexpr int CQI = CINP / 8, CQO = COUTP / 8;
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t *xin = reinterpret_cast<uint16_t *>(smem);  // [3][CINP][WCP]: copy kx, channel, row * WTW + x
    uint16_t *dzt = xin + 3 * CINP * WCP;                 // [COUTP][WZP]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    const int tx_n = (W + WTW - 1) / WTW, ty_n = (H + WTH - 1) / WTH;
    const long ntiles = (long)N * tx_n * ty_n;
    f32x4 acc[NTW][MT];
#pragma unroll
    for (int i = 0; i < NTW; ++i)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[i][m] = (f32x4){0.f, 0.f, 0.f, 0.f};
    f32x4 bacc = (f32x4){0.f, 0.f, 0.f, 0.f};
    bf16x8 ones;
#pragma unroll
    for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.f;

    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int n = (int)(t / ((long)tx_n * ty_n));
        const int rem = (int)(t - (long)n * tx_n * ty_n);
        const int y0 = (rem / tx_n) * WTH, x0 = (rem % tx_n) * WTW;
        __syncthreads();  // the previous tile's operand reads are done
        for (int k = tid; k < WPIX * CQO; k += THREADS) {
            const int q = k % CQO, p = k / CQO, yy = y0 + p / WTW, xx = x0 + p % WTW;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (yy < H && xx < W) {
                const size_t o = (((size_t)n * H + yy) * W + xx) * COUTP + 8 * q;
                v = *reinterpret_cast<const uint4 *>(dout + o);
                if (mask) v = apply_mask(v, *reinterpret_cast<const uint4 *>(mask + o));
            }
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
            uint16_t *d = dzt + (size_t)(8 * q) * WZP + p;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                d[(2 * j) * WZP] = (uint16_t)(w4[j] & 0xFFFFu);
                d[(2 * j + 1) * WZP] = (uint16_t)(w4[j] >> 16);
            }
        }
        // staged pixel (sy, sx) = input (y0 + sy - 1, x0 + sx - 1); copy kx holds x = sx - kx in [0, WTW)
        for (int k = tid; k < WSR * (WTW + 2) * CQI; k += THREADS) {
            const int q = k % CQI, sp = k / CQI, sx = sp % (WTW + 2), sy = sp / (WTW + 2);
            const int yy = y0 + sy - 1, xx = x0 + sx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *reinterpret_cast<const uint4 *>(in + (((size_t)n * H + yy) * W + xx) * CINP + 8 * q);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int x = sx - kx;
                if (x < 0 || x >= WTW) continue;
                uint16_t *d = xin + (size_t)(kx * CINP + 8 * q) * WCP + sy * WTW + x;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    d[(2 * j) * WCP] = (uint16_t)(w4[j] & 0xFFFFu);
                    d[(2 * j + 1) * WCP] = (uint16_t)(w4[j] >> 16);
                }
            }
        }
        __syncthreads();
        // A fragments: output channels 16m + l16, pixels 8g..8g+7 of tile row kc
        bf16x8 a[MT][WTH];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int kc = 0; kc < WTH; ++kc)
                a[m][kc] = *reinterpret_cast<const bf16x8 *>(dzt + (size_t)(16 * m + l16) * WZP + kc * WTW + 8 * g);
#pragma unroll
        for (int i = 0; i < NTW; ++i) {
            const int nt = wave + 4 * i;
            if (nt >= NTT) continue;  // wave-uniform
            const int tap = (16 * nt) / CINP, ci0 = 16 * nt - tap * CINP, ky = tap / 3, kx = tap - 3 * ky;
            const uint16_t *xb = xin + (size_t)(kx * CINP + ci0 + l16) * WCP + 8 * g;
#pragma unroll
            for (int kc = 0; kc < WTH; ++kc) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(xb + (kc + ky) * WTW);
#pragma unroll
                for (int m = 0; m < MT; ++m) acc[i][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m][kc], b, acc[i][m], 0, 0, 0);
            }
        }
        if (wave < MT) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
                if (m == wave) {
#pragma unroll
                    for (int kc = 0; kc < WTH; ++kc) bacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m][kc], ones, bacc, 0, 0, 0);
                }
        }
    }
    // D layout: lane holds rows 4g..4g+3 (output channels) of column l16
    float *dst = partial + (size_t)blockIdx.x * ((size_t)cout * 9 * cin + cout);
#pragma unroll
    for (int i = 0; i < NTW; ++i) {
        const int nt = wave + 4 * i;
        if (nt >= NTT) continue;
        const int kk = 16 * nt + l16, tap = kk / CINP, ci = kk - tap * CINP;
        if (ci >= cin) continue;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = 16 * m + 4 * g + r;
                if (co < cout) dst[((size_t)co * 9 + tap) * cin + ci] = acc[i][m][r];
            }
    }
    if (wave < MT && l16 == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = 16 * wave + 4 * g + r;
            if (co < cout) dst[(size_t)cout * 9 * cin + co] = bacc[r];
        }
    }
}

// dw[co][ci][tap] and db[co]: the partial sums of `blocks` workgroups, summed
// in a fixed order (wave w adds blocks w, w + 4, ...; the four wave sums are
// then added in wave order).
__global__ __launch_bounds__(256) void wgrad_sum_kernel(const float *__restrict__ partial, int blocks, int cin,
                                                        int cout, float *__restrict__ dw, float *__restrict__ db)
{
    __shared__ float part[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = cout * 9 * cin, stride = nw + cout;
    const int o = blockIdx.x * 64 + lane;
    float s = 0.f;
    if (o < stride)
        for (int b = wave; b < blocks; b += 4) s += partial[(size_t)b * stride + o];
    part[wave][lane] = s;
    __syncthreads();
    if (wave != 0 || o >= stride) return;
    s = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
    if (o >= nw) {
        if (db) db[o - nw] = s;
        return;
    }
    const int co = o / (9 * cin), r = o - co * 9 * cin, tap = r / cin, ci = r - tap * cin;
    dw[((size_t)co * cin + ci) * 9 + tap] = s;
}

__global__ void pack_kernel(const float *__restrict__ w, uint16_t *__restrict__ wf, uint16_t *__restrict__ wb, int cin,
                            int cout, int cinp, int coutp)
{
    const int kpf = kpad(cinp), kpb = kpad(coutp);
    const int nf = coutp * kpf, nb = cinp * kpb;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf + nb; i += gridDim.x * blockDim.x) {
        if (i < nf) {
            const int co = i / kpf, k = i - co * kpf, tap = k / cinp, ci = k - tap * cinp;
            const float v = (co < cout && tap < 9 && ci < cin) ? w[((size_t)co * cin + ci) * 9 + tap] : 0.f;
            wf[i] = f2bf(v);
        } else {
            const int j = i - nf, ci = j / kpb, k = j - ci * kpb, tp = k / coutp, co = k - tp * coutp;
            const float v = (ci < cin && tp < 9 && co < cout) ? w[((size_t)co * cin + ci) * 9 + (8 - tp)] : 0.f;
            wb[j] = f2bf(v);
        }
    }
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// network input [n][h][w][16]: channels 0..C-1 the noisy samples of the
// image's C planes plus each plane's offset, channel C sigma/255, the rest zero
__global__ void input_kernel(const float *__restrict__ noisy, int C, const float *__restrict__ sigma,
                             const float *__restrict__ offset, uint16_t *__restrict__ x, int N, int H, int W)
{
    const long hw = (long)H * W, total = (long)N * hw;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int n = (int)(i / hw);
        const long p = i - (long)n * hw;
        uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int c = 0; c < C; ++c) {
            const long q = (long)n * C + c;
            const uint32_t b = f2bf(noisy[q * hw + p] + (offset ? offset[q] : 0.f));
            v[c >> 1] |= b << (16 * (c & 1));
        }
        v[C >> 1] |= (uint32_t)f2bf(sigma[n] / 255.f) << (16 * (C & 1));
        uint4 *d = reinterpret_cast<uint4 *>(x + (size_t)i * 16);
        d[0] = make_uint4(v[0], v[1], v[2], v[3]);
        d[1] = make_uint4(v[4], v[5], v[6], v[7]);
    }
}

// out = noisy - residual per plane (residual channel c of res32 for plane c);
// with a target also the loss sum and d(loss)/d(residual) for
// loss = mean over the N * C planes of (out - target)^2.
__global__ void output_kernel(const float *__restrict__ noisy, const float *__restrict__ res32, int C,
                              float *__restrict__ out, const float *__restrict__ target, uint16_t *__restrict__ dres,
                              float *__restrict__ loss, int N, int H, int W)
{
    const long hw = (long)H * W, total = (long)N * hw;
    const float count = (float)(total * C);
    float part = 0.f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int n = (int)(i / hw);
        const long p = i - (long)n * hw;
        uint32_t g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int c = 0; c < C; ++c) {
            const long q = ((long)n * C + c) * hw + p;
            const float o = noisy[q] - res32[(size_t)i * 16 + c];
            if (out) out[q] = o;
            if (target) {
                const float d = o - target[q];
                part += d * d;
                g[c >> 1] |= (uint32_t)f2bf(-2.f * d / count) << (16 * (c & 1));
            }
        }
        if (target) {
            uint4 *q = reinterpret_cast<uint4 *>(dres + (size_t)i * 16);
            q[0] = make_uint4(g[0], g[1], g[2], g[3]);
            q[1] = make_uint4(g[4], g[5], g[6], g[7]);
        }
    }
    if (target && loss) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
        if ((threadIdx.x & 63) == 0) atomicAdd(loss, part / count);
    }
}

template <typename K>
hipError_t set_lds(K kernel, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

template <int CIN, int COUT>
hipError_t conv_t(const void *in, const void *mask, const void *w, const float *bias, void *out, float *out32, int n,
                  int h, int wd, int relu, hipStream_t s)
{
    const size_t lds = (size_t)SH * SW * in_pitch(CIN) * 2 + (size_t)COUT * w_pitch(CIN) * 2;
    // two persistent workgroups per CU (the LDS of one is within half a CU's)
    hipError_t e;
    int dev = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const long tiles = (long)((wd + TW - 1) / TW) * ((h + TH - 1) / TH) * n;
    const int grid = (int)std::min<long>(tiles, 2L * std::max(cus, 1));
    if (!mask && !out32) {
        auto k = conv3x3_fwd_kernel<CIN, COUT>;
        const size_t lf = (size_t)fwd_area<CIN, COUT>() * 2 + (size_t)COUT * w_pitch(CIN) * 2;
        if ((e = set_lds(k, lf)) != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3(grid), dim3(THREADS), lf, s, (const uint16_t *)in, (const uint16_t *)w, bias,
                           (uint16_t *)out, h, wd, relu, n, nullptr, nullptr, 0);
        return hipGetLastError();
    }
    auto k = conv3x3_kernel<CIN, COUT>;
    if ((e = set_lds(k, lds)) != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(THREADS), lds, s, (const uint16_t *)in, (const uint16_t *)mask,
                       (const uint16_t *)w, bias, (uint16_t *)out, out32, h, wd, relu, n);
    return hipGetLastError();
}

template <int CINP, int COUTP>
hipError_t wgrad_t(const void *in, const void *dout, const void *mask, float *partial, int blocks, int n, int h,
                   int wd, int cin, int cout, hipStream_t s)
{
    auto k = wgrad_mfma_kernel<CINP, COUTP>;
    constexpr size_t lds = wgrad_lds_bytes<CINP, COUTP>();
    hipError_t e = set_lds(k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(THREADS), lds, s, (const uint16_t *)in,
                       (const uint16_t *)dout, (const uint16_t *)mask, partial, n, h, wd, cin, cout);
    return hipGetLastError();
}

int wgrad_blocks(int n, int h, int w)
{
    const long tiles = (long)n * ((w + WTW - 1) / WTW) * ((h + WTH - 1) / WTH);
    return (int)(tiles < WBLOCKS ? tiles : WBLOCKS);
}

}  // namespace
}  // namespace pilot

using namespace pilot;

extern "C" {

int isp_pilot_conv3x3_residual(const void *in, const void *weights, const float *bias, const float *noisy,
                               float *est, int channels, int n, int h, int w, int cin_pad, int cout_pad, void *stream)
{
    if (!in || !weights || !noisy || !est || n <= 0 || h <= 0 || w <= 0 || channels < 1 || channels > 16)
        return ISP_PILOT_E_ARG;
    if (cin_pad != 48 || cout_pad != 16) return ISP_PILOT_E_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    int dev = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess || (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                                                            dev)) != hipSuccess)
        return ISP_PILOT_E_HIP;
    const long tiles = (long)((w + TW - 1) / TW) * ((h + TH - 1) / TH) * n;
    const int grid = (int)std::min<long>(tiles, 2L * std::max(cus, 1));
    auto k = conv3x3_fwd_kernel<48, 16, true>;
    const size_t lf = (size_t)fwd_area<48, 16>() * 2 + (size_t)16 * w_pitch(48) * 2;
    static_assert(fwd_area<48, 16>() * 2 >= 16 * TH * TW * 4, "the FP32 residual tile fits the shared area");
    if ((e = set_lds(k, lf)) != hipSuccess) return ISP_PILOT_E_HIP;
    hipLaunchKernelGGL(k, dim3(grid), dim3(THREADS), lf, s, (const uint16_t *)in, (const uint16_t *)weights, bias,
                       nullptr, h, w, 0, n, noisy, est, channels);
    return hipGetLastError() == hipSuccess ? ISP_PILOT_OK : ISP_PILOT_E_HIP;
}

int isp_pilot_conv3x3(const void *in, const void *mask, const void *weights, const float *bias, void *out,
                      float *out_f32, int n, int h, int w, int cin_pad, int cout_pad, int relu, void *stream)
{
    if (!in || !weights || !out || n <= 0 || h <= 0 || w <= 0) return ISP_PILOT_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (cin_pad == 16 && cout_pad == 48) e = conv_t<16, 48>(in, mask, weights, bias, out, out_f32, n, h, w, relu, s);
    else if (cin_pad == 48 && cout_pad == 48) e = conv_t<48, 48>(in, mask, weights, bias, out, out_f32, n, h, w, relu, s);
    else if (cin_pad == 48 && cout_pad == 16) e = conv_t<48, 16>(in, mask, weights, bias, out, out_f32, n, h, w, relu, s);
    else return ISP_PILOT_E_UNSUPPORTED;
    return e == hipSuccess ? ISP_PILOT_OK : ISP_PILOT_E_HIP;
}

size_t isp_pilot_wgrad_partial_floats(int n, int h, int w, int cin, int cout)
{
    if (n <= 0 || h <= 0 || w <= 0 || cin < 1 || cout < 1) return 0;
    return ((size_t)cout * 9 * cin + cout) * (size_t)wgrad_blocks(n, h, w);
}

int isp_pilot_conv3x3_wgrad(const void *in, const void *dout, const void *mask, float *partial, float *dw, float *db,
                            int n, int h, int w, int cin_pad, int cout_pad, int cin, int cout, void *stream)
{
    if (!in || !dout || !partial || !dw || n <= 0 || h <= 0 || w <= 0) return ISP_PILOT_E_ARG;
    if (cin < 1 || cin > cin_pad || cout < 1 || cout > cout_pad) return ISP_PILOT_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int blocks = wgrad_blocks(n, h, w);
    hipError_t e;
    if (cin_pad == 16 && cout_pad == 48) e = wgrad_t<16, 48>(in, dout, mask, partial, blocks, n, h, w, cin, cout, s);
    else if (cin_pad == 48 && cout_pad == 48) e = wgrad_t<48, 48>(in, dout, mask, partial, blocks, n, h, w, cin, cout, s);
    else if (cin_pad == 48 && cout_pad == 16) e = wgrad_t<48, 16>(in, dout, mask, partial, blocks, n, h, w, cin, cout, s);
    else return ISP_PILOT_E_UNSUPPORTED;
    if (e != hipSuccess) return ISP_PILOT_E_HIP;
    const int stride = cout * 9 * cin + cout;
    hipLaunchKernelGGL(wgrad_sum_kernel, dim3((stride + 63) / 64), dim3(256), 0, s, partial, blocks, cin, cout, dw, db);
    return hipGetLastError() == hipSuccess ? ISP_PILOT_OK : ISP_PILOT_E_HIP;
}

int isp_pilot_pack(const float *w, void *wf, void *wb, int cin, int cout, int cin_pad, int cout_pad, void *stream)
{
    if (!w || !wf || !wb || cin < 1 || cout < 1 || cin > cin_pad || cout > cout_pad || cin_pad % 16 ||
        cout_pad % 16)
        return ISP_PILOT_E_ARG;
    const int total = cout_pad * kpad(cin_pad) + cin_pad * kpad(cout_pad);
    hipLaunchKernelGGL(pack_kernel, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, w, (uint16_t *)wf,
                       (uint16_t *)wb, cin, cout, cin_pad, cout_pad);
    return hipGetLastError() == hipSuccess ? ISP_PILOT_OK : ISP_PILOT_E_HIP;
}

int isp_pilot_kpad(int cin_pad) { return kpad(cin_pad); }

int isp_pilot_adam(float *p, const float *g, float *m, float *v, long n, float lr, float beta1, float beta2, float eps,
                   int step, void *stream)
{
    if (!p || !g || !m || !v || n <= 0 || step < 1) return ISP_PILOT_E_ARG;
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    const long blocks = (n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048;
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, beta1,
                       beta2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? ISP_PILOT_OK : ISP_PILOT_E_HIP;
}

int isp_pilot_input_planes(const float *noisy, int channels, const float *sigma, const float *offset, void *x, int n,
                           int h, int w, void *stream)
{
    if (!noisy || !sigma || !x || n <= 0 || h <= 0 || w <= 0 || channels < 1 || channels > 15) return ISP_PILOT_E_ARG;
    const long total = (long)n * h * w;
    const long blocks = (total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096;
    hipLaunchKernelGGL(input_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, noisy, channels,
                       sigma, offset, (uint16_t *)x, n, h, w);
    return hipGetLastError() == hipSuccess ? ISP_PILOT_OK : ISP_PILOT_E_HIP;
}

int isp_pilot_input(const float *noisy, const float *sigma, const float *offset, void *x, int n, int h, int w,
                    void *stream)
{
    return isp_pilot_input_planes(noisy, 1, sigma, offset, x, n, h, w, stream);
}

int isp_pilot_output_planes(const float *noisy, const float *res_f32, float *out, const float *target, void *dres,
                            float *loss, int channels, int n, int h, int w, void *stream)
{
    if (!noisy || !res_f32 || n <= 0 || h <= 0 || w <= 0 || (target && !dres) || channels < 1 || channels > 16)
        return ISP_PILOT_E_ARG;
    const long total = (long)n * h * w;
    const long blocks = (total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096;
    hipLaunchKernelGGL(output_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, noisy, res_f32,
                       channels, out, target, (uint16_t *)dres, loss, n, h, w);
    return hipGetLastError() == hipSuccess ? ISP_PILOT_OK : ISP_PILOT_E_HIP;
}

int isp_pilot_output(const float *noisy, const float *res_f32, float *out, const float *target, void *dres, float *loss,
                     int n, int h, int w, void *stream)
{
    return isp_pilot_output_planes(noisy, res_f32, out, target, dres, loss, 1, n, h, w, stream);
}

}  // extern "C"
