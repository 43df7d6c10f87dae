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
// kernel's LDS (81,856 bytes with padding) within half of the 160 KB of a CU,
// so two workgroups overlap their staging with each other's products
constexpr int TH = 8, TW = 32;
constexpr int RPW = TH / 4;              // rows per wave
constexpr int MT = RPW * (TW / 16);      // 16-pixel M-tiles per wave
constexpr int SH = TH + 2, SW = TW + 2;  // staged input tile with halo
constexpr int THREADS = 256;             // 4 waves x RPW rows
constexpr int GTH = 8, GTW = 32;         // pixel tile of the weight-gradient kernel

__host__ __device__ constexpr int kpad(int cin) { return (9 * cin + 31) / 32 * 32; }

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

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
// COUT/16 N-tiles.
// LDS rows are padded by 8 elements so that the 16-byte operand reads of 8
// neighbouring lanes fall in distinct banks.
template <int CIN, int COUT>
__global__ __launch_bounds__(THREADS) void conv3x3_kernel(const uint16_t *__restrict__ in,
                                                          const uint16_t *__restrict__ mask,
                                                          const uint16_t *__restrict__ wts,
                                                          const float *__restrict__ bias, uint16_t *__restrict__ out,
                                                          float *__restrict__ out32, int H, int W, int relu, int N)
{
    constexpr int KP = kpad(CIN), KC = KP / 32, NT = COUT / 16;
    constexpr int SP = CIN + 8, KS = KP + 8, CQ = CIN / 8;
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
// outputs): the next tile's input is loaded into registers while the current
// tile's products run, and the outputs pass through LDS (bias, ReLU, bf16) so
// that whole pixel rows go out as 16-byte stores.  The shared area holds the
// staged input tile, then the output tile: max(SH*SW*SP, TH*TW*COUT) elements.
template <int CIN, int COUT>
__host__ __device__ constexpr int fwd_area()
{
    return SH * SW * (CIN + 8) > TH * TW * COUT ? SH * SW * (CIN + 8) : TH * TW * COUT;
}

template <int CIN, int COUT>
__global__ __launch_bounds__(THREADS) void conv3x3_fwd_kernel(const uint16_t *__restrict__ in,
                                                              const uint16_t *__restrict__ wts,
                                                              const float *__restrict__ bias,
                                                              uint16_t *__restrict__ out, int H, int W, int relu,
                                                              int N)
{
    constexpr int KP = kpad(CIN), KC = KP / 32, NT = COUT / 16;
    constexpr int SP = CIN + 8, KS = KP + 8, CQ = CIN / 8, CO8 = COUT / 8;
    constexpr int NIN = SH * SW * CQ, NPF = (NIN + THREADS - 1) / THREADS;
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t *tin = reinterpret_cast<uint16_t *>(smem);  // [SH*SW][SP], then the outputs [TH*TW][COUT]
    uint16_t *tw = tin + fwd_area<CIN, COUT>();          // [COUT][KS]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < COUT * (KP / 8); k += THREADS) {
        const int row = k / (KP / 8), q = k - row * (KP / 8);
        *reinterpret_cast<uint4 *>(tw + (size_t)row * KS + 8 * q) =
            *reinterpret_cast<const uint4 *>(wts + (size_t)row * KP + 8 * q);
    }
    const int ntx = (W + TW - 1) / TW, nty = (H + TH - 1) / TH, per = ntx * nty, total = per * N;

    // the input tile (with halo, zeros outside the image) of tile t into registers
    uint4 pf[NPF];
    auto fetch = [&](int t) {
        const int n = t / per, tr = t - n * per, tyi = tr / ntx;
        const int x0 = (tr - tyi * ntx) * TW, y0 = tyi * TH;
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
            const int k = tid + i * THREADS;
            const int q = k % CQ, p = k / CQ, sx = p % SW, sy = p / SW;
            const int yy = y0 + sy - 1, xx = x0 + sx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (k < NIN && yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *reinterpret_cast<const uint4 *>(in + (((size_t)n * H + yy) * W + xx) * CIN + 8 * q);
            pf[i] = v;
        }
    };
    if ((int)blockIdx.x < total) fetch(blockIdx.x);

    const int m = lane & 15, g = lane >> 4;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int n = t / per, tr = t - n * per, tyi = tr / ntx;
        const int x0 = (tr - tyi * ntx) * TW, y0 = tyi * TH;
        __syncthreads();  // the previous tile's output rows are out of the shared area
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
            const int k = tid + i * THREADS;
            if (k < NIN) *reinterpret_cast<uint4 *>(tin + (size_t)(k / CQ) * SP + 8 * (k % CQ)) = pf[i];
        }
        __syncthreads();
        if (t + (int)gridDim.x < total) fetch(t + gridDim.x);  // in flight during the products

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
        __syncthreads();  // every wave is done reading the input tile
        // D layout: lane holds rows 4g..4g+3 (pixels) of column lane % 16 (channel)
        uint16_t *tout = tin;  // [TH*TW][COUT]
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int r = RPW * wave + (mt >> 1);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int co = 16 * nt + m;
                const float bv = bias ? bias[co] : 0.f;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int px = (mt & 1) * 16 + 4 * g + i;
                    float v = acc[mt][nt][i] + bv;
                    if (relu) v = fmaxf(v, 0.f);
                    tout[(size_t)(r * TW + px) * COUT + co] = f2bf(v);
                }
            }
        }
        __syncthreads();
        for (int k = tid; k < TH * TW * CO8; k += THREADS) {
            const int pix = k / CO8, q = k - pix * CO8, r = pix / TW, px = pix - r * TW;
            const int yy = y0 + r, xx = x0 + px;
            if (yy < H && xx < W)
                *reinterpret_cast<uint4 *>(out + (((size_t)n * H + yy) * W + xx) * COUT + 8 * q) =
                    *reinterpret_cast<const uint4 *>(tout + (size_t)pix * COUT + 8 * q);
        }
    }
}

// Weight and bias gradients: dw[co][ci][tap] = sum_p dz[p][co] * in[p + off(tap)][ci],
// db[co] = sum_p dz[p][co], with dz = dout * (mask > 0) (mask may be null).
// A persistent grid walks the pixel tiles; workgroup b writes its sums to
// partial[b][*] (work item it = ((cb * cin) + ci) * 9 + tap holds CB outputs
// co = cb*CB + j, then cout bias sums), reduced in a fixed order afterwards.
template <int CIN, int COUT, int CB, int IPT>
__global__ __launch_bounds__(THREADS) void wgrad_kernel(const uint16_t *__restrict__ in,
                                                        const uint16_t *__restrict__ dout,
                                                        const uint16_t *__restrict__ mask, float *__restrict__ partial,
                                                        int N, int H, int W, int cin, int cout)
{
    constexpr int ASH = GTH + 2, ASW = GTW + 2;
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t *ta = reinterpret_cast<uint16_t *>(smem);  // [ASH*ASW][CIN]
    uint16_t *tz = ta + ASH * ASW * CIN;                 // [GTH*GTW][COUT]
    const int tid = threadIdx.x;
    const int ncb = (cout + CB - 1) / CB;
    const int items = ncb * cin * 9;
    const int tx_n = (W + GTW - 1) / GTW, ty_n = (H + GTH - 1) / GTH;
    const long ntiles = (long)N * tx_n * ty_n;
    float acc[IPT][CB];
#pragma unroll
    for (int j = 0; j < IPT; ++j)
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[j][c] = 0.f;
    float bacc = 0.f;

    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int n = (int)(t / ((long)tx_n * ty_n));
        const int rem = (int)(t - (long)n * tx_n * ty_n);
        const int y0 = (rem / tx_n) * GTH, x0 = (rem % tx_n) * GTW;
        __syncthreads();
        for (int k = tid; k < ASH * ASW * (CIN / 8); k += THREADS) {
            const int q = k % (CIN / 8), p = k / (CIN / 8), sx = p % ASW, sy = p / ASW;
            const int yy = y0 + sy - 1, xx = x0 + sx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *reinterpret_cast<const uint4 *>(in + (((size_t)n * H + yy) * W + xx) * CIN + 8 * q);
            *reinterpret_cast<uint4 *>(ta + (size_t)p * CIN + 8 * q) = v;
        }
        for (int k = tid; k < GTH * GTW * (COUT / 8); k += THREADS) {
            const int q = k % (COUT / 8), p = k / (COUT / 8), px = p % GTW, py = p / GTW;
            const int yy = y0 + py, xx = x0 + px;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (yy < H && xx < W) {
                const size_t o = (((size_t)n * H + yy) * W + xx) * COUT + 8 * q;
                v = *reinterpret_cast<const uint4 *>(dout + o);
                if (mask) v = apply_mask(v, *reinterpret_cast<const uint4 *>(mask + o));
            }
            *reinterpret_cast<uint4 *>(tz + (size_t)p * COUT + 8 * q) = v;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int it = tid + j * THREADS;
            if (it >= items) continue;
            const int tap = it % 9, ci = (it / 9) % cin, cb = it / (9 * cin);
            const int ky = tap / 3, kx = tap - 3 * ky;
            for (int p = 0; p < GTH * GTW; ++p) {
                const int py = p / GTW, px = p - py * GTW;
                const float a = bf2f(ta[((py + ky) * ASW + px + kx) * CIN + ci]);
                const uint16_t *z = tz + (size_t)p * COUT + cb * CB;
#pragma unroll
                for (int c = 0; c < CB; ++c) acc[j][c] = fmaf(a, bf2f(z[c]), acc[j][c]);
            }
        }
        if (tid < cout)
            for (int p = 0; p < GTH * GTW; ++p) bacc += bf2f(tz[(size_t)p * COUT + tid]);
    }
    const int stride = items * CB + cout;
    float *dst = partial + (size_t)blockIdx.x * stride;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const int it = tid + j * THREADS;
        if (it >= items) continue;
#pragma unroll
        for (int c = 0; c < CB; ++c) dst[it * CB + c] = acc[j][c];
    }
    if (tid < cout) dst[items * CB + tid] = bacc;
}

// dw[co][ci][tap] and db[co] from the partial sums of `blocks` workgroups.
__global__ void wgrad_reduce_kernel(const float *__restrict__ partial, int blocks, int cin, int cout, int CB,
                                    float *__restrict__ dw, float *__restrict__ db)
{
    const int ncb = (cout + CB - 1) / CB, items = ncb * cin * 9, stride = items * CB + cout;
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= stride) return;
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * stride + o];
    if (o >= items * CB) {
        if (db) db[o - items * CB] = s;
        return;
    }
    const int it = o / CB, c = o - it * CB;
    const int tap = it % 9, ci = (it / 9) % cin, co = (it / (9 * cin)) * CB + c;
    if (co < cout) dw[((size_t)co * cin + ci) * 9 + tap] = s;
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
    constexpr int KP = kpad(CIN);
    const size_t lds = (size_t)SH * SW * (CIN + 8) * 2 + (size_t)COUT * (KP + 8) * 2;
    // two persistent workgroups per CU (the LDS of one is within half a CU's)
    hipError_t e;
    int dev = 0, cus = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const long tiles = (long)((wd + TW - 1) / TW) * ((h + TH - 1) / TH) * n;
    const int grid = (int)std::min<long>(tiles, 2L * std::max(cus, 1));
    if (!mask && !out32) {
        auto k = conv3x3_fwd_kernel<CIN, COUT>;
        const size_t lf = (size_t)fwd_area<CIN, COUT>() * 2 + (size_t)COUT * (KP + 8) * 2;
        if ((e = set_lds(k, lf)) != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3(grid), dim3(THREADS), lf, s, (const uint16_t *)in, (const uint16_t *)w, bias,
                           (uint16_t *)out, h, wd, relu, n);
        return hipGetLastError();
    }
    auto k = conv3x3_kernel<CIN, COUT>;
    if ((e = set_lds(k, lds)) != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(THREADS), lds, s, (const uint16_t *)in, (const uint16_t *)mask,
                       (const uint16_t *)w, bias, (uint16_t *)out, out32, h, wd, relu, n);
    return hipGetLastError();
}

template <int CIN, int COUT, int CB, int IPT>
hipError_t wgrad_t(const void *in, const void *dout, const void *mask, float *partial, int blocks, int n, int h,
                   int wd, int cin, int cout, hipStream_t s)
{
    const size_t lds = ((size_t)(GTH + 2) * (GTW + 2) * CIN + (size_t)GTH * GTW * COUT) * 2;
    auto k = wgrad_kernel<CIN, COUT, CB, IPT>;
    hipError_t e = set_lds(k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(THREADS), lds, s, (const uint16_t *)in, (const uint16_t *)dout,
                       (const uint16_t *)mask, partial, n, h, wd, cin, cout);
    return hipGetLastError();
}

// output channels per work item of the weight gradient: 12 of a hidden
// layer's 48; the last layer's 1 (grey) or 3 (OPP planes) in one item
int cb_of(int cout) { return cout > 12 ? 12 : cout; }

int wgrad_blocks(int n, int h, int w)
{
    const long tiles = (long)n * ((w + GTW - 1) / GTW) * ((h + GTH - 1) / GTH);
    return (int)(tiles < 1024 ? tiles : 1024);
}

}  // namespace
}  // namespace pilot

using namespace pilot;

extern "C" {

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
    const int cb = cb_of(cout);
    const size_t stride = (size_t)((cout + cb - 1) / cb) * cin * 9 * cb + cout;
    return stride * (size_t)wgrad_blocks(n, h, w);
}

int isp_pilot_conv3x3_wgrad(const void *in, const void *dout, const void *mask, float *partial, float *dw, float *db,
                            int n, int h, int w, int cin_pad, int cout_pad, int cin, int cout, void *stream)
{
    if (!in || !dout || !partial || !dw || n <= 0 || h <= 0 || w <= 0) return ISP_PILOT_E_ARG;
    if (cin < 1 || cin > cin_pad || cout < 1 || cout > cout_pad) return ISP_PILOT_E_ARG;
    const int cb = cb_of(cout);
    // CB 12 for the 48 outputs of a hidden layer, 1 or 3 for the last layer
    if ((cout_pad == 48) != (cb == 12) || (cout_pad == 16 && cout != 1 && cout != 3)) return ISP_PILOT_E_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    const int blocks = wgrad_blocks(n, h, w);
    // work items per thread the instantiations hold (items = ceil(cout/CB) * cin * 9)
    const int ipt = cin_pad == 16 ? 1 : (cout_pad == 48 ? 7 : 2);
    if (((cout + cb - 1) / cb) * cin * 9 > ipt * THREADS) return ISP_PILOT_E_UNSUPPORTED;
    hipError_t e;
    if (cin_pad == 16 && cout_pad == 48) e = wgrad_t<16, 48, 12, 1>(in, dout, mask, partial, blocks, n, h, w, cin, cout, s);
    else if (cin_pad == 48 && cout_pad == 48) e = wgrad_t<48, 48, 12, 7>(in, dout, mask, partial, blocks, n, h, w, cin, cout, s);
    else if (cin_pad == 48 && cout_pad == 16 && cout == 1)
        e = wgrad_t<48, 16, 1, 2>(in, dout, mask, partial, blocks, n, h, w, cin, cout, s);
    else if (cin_pad == 48 && cout_pad == 16 && cout == 3)
        e = wgrad_t<48, 16, 3, 2>(in, dout, mask, partial, blocks, n, h, w, cin, cout, s);
    else return ISP_PILOT_E_UNSUPPORTED;
    if (e != hipSuccess) return ISP_PILOT_E_HIP;
    const int stride = ((cout + cb - 1) / cb) * cin * 9 * cb + cout;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((stride + 255) / 256), dim3(256), 0, s, partial, blocks, cin, cout, cb,
                       dw, db);
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
