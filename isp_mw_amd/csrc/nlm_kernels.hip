// HIP kernels of the NL-means denoiser (gfx950, wave64).
//
// NLMeans (source/NLMeans.cpp:150-375) forms the same groups as BM3D's
// matching (match_kernel, with the unscaled thMSE) and replaces the 3D
// transform by a weighted average of each group:
//   nlm_kernel          - per reference block, weights exp(-key / strength^2)
//                         and the weighted mean (WeightedAverage, :6-61) or the
//                         mean corrected by a stationarity test
//                         (WeightedAverage_Correction, :66-147), summed into an
//                         LDS tile of 64-bit fixed-point accumulators and flushed
//                         to HBM with integer atomics (AddTo, :223-224).
//   nlm_finalize_kernel - num / den with den the number of grid blocks that
//                         cover the pixel (CountTo, :224, :367), then the output
//                         conversion of the plane or frame form (:230, :254,
//                         :355-372).
// Every sum is formed in the reference's order in FP64 with separate multiply
// and add roundings, so outputs equal the reference's up to the last bits of
// exp() and of the aggregation (2^-52 relative to the fixed-point scale).
#include "bm3d_internal.h"
#include "tile_order.h"

namespace ispbm3d {
namespace {

typedef unsigned long long u64;

// 8 waves per workgroup: one wave per group, and a tile's LDS accumulators
// allow few workgroups per CU, so more waves per workgroup keep more groups
// (and their sample loads) in flight
constexpr int kNlmWaves = 8;
constexpr int kNlmThreads = 64 * kNlmWaves;

__device__ __forceinline__ int grid_pos(int k, int s, int last) { return min(k * s, last); }

template <typename ST>
__device__ __forceinline__ double load(const void *p, size_t o)
{
    return (double)static_cast<const ST *>(p)[o];
}

// floor(v * fix) added to a 64-bit fixed-point LDS accumulator.  |v| * fix
// stays below 2^57 by the choice of fix, so the product, floor and conversion
// are exact.
__device__ __forceinline__ void acc_add(u64 *acc, double v, double fix)
{
    atomicAdd(acc, (u64)(long long)floor(v * fix));
}

// lane l's value of v, l wave-uniform
__device__ __forceinline__ double readlane_d(double v, int l)
{
    const u64 b = __double_as_longlong(v);
    const u64 lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const u64 hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(lo | (hi << 32)));
}

// Members' samples are loaded this many at a time, then summed in member order
constexpr int kNlmBatch = 8;

template <typename ST>
__device__ void nlm_tile(const NlmArgs &a, u64 *acc, int tx, int ty)
{
    // separate mul and add roundings, as the reference's double code
#pragma clang fp contract(off)
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int trx = a.trx, tnum = a.trx * a.try_;
    const int tx0 = tx * trx, ty0 = ty * a.try_;
    const int lastx = a.W - kBlock, lasty = a.H - kBlock;
    const int ox = grid_pos(tx0, a.bstep, lastx), oy = grid_pos(ty0, a.bstep, lasty);
    const int FW = a.FW, FH = a.FH, np = a.nplanes;
    const int u = lane >> 3, v = lane & 7;

    for (int k = tid; k < np * FW * FH; k += kNlmThreads) acc[k] = 0;
    __syncthreads();

    for (int r = wave; r < tnum; r += kNlmWaves) {
        const int gxi = tx0 + (r % trx), gyi = ty0 + (r / trx);
        if (gxi >= a.ngx || gyi >= a.ngy) continue;
        const int gx = grid_pos(gxi, a.bstep, lastx), gy = grid_pos(gyi, a.bstep, lasty);
        const size_t refIdx = (size_t)gyi * a.ngx + gxi;
        const int *pos = a.pos + refIdx * a.G;
        const int N = a.cnt[refIdx];
        const size_t self = (size_t)(gy + u) * a.spitch + gx + v;
        const int ai = (gy - oy + u) * FW + gx - ox + v;
        if (N < 2) {
            // a group of one: the source block itself (:15-19)
            for (int c = 0; c < np; ++c) acc_add(&acc[c * FW * FH + ai], load<ST>(a.src[c], self), a.fix);
            continue;
        }
        // weights of the group members, lane z holding member z's; their sum
        // in member order
        const double wz = lane < N ? exp(a.keys[refIdx * a.G + lane] * a.exp_mul) : 0.0;
        const int pl = lane < N ? pos[lane] : 0;
        double wsum = 0.0;
        for (int z = 0; z < N; ++z) wsum += readlane_d(wz, z);
        const double rec = 1.0 / wsum;
        for (int c = 0; c < np; ++c) {
            const void *src = a.src[c];
            double s1 = 0.0, s2 = 0.0;
            for (int z0 = 0; z0 < N; z0 += kNlmBatch) {
                double xs[kNlmBatch];
#pragma unroll
                for (int j = 0; j < kNlmBatch; ++j) {
                    const int pz = __builtin_amdgcn_readlane(pl, min(z0 + j, N - 1));
                    xs[j] = load<ST>(src, (size_t)((pz >> 16) + u) * a.spitch + (pz & 0xFFFF) + v);
                }
#pragma unroll
                for (int j = 0; j < kNlmBatch; ++j) {
                    if (z0 + j >= N) break;
                    const double w = readlane_d(wz, z0 + j), x = xs[j];
                    if (a.correction) {
                        const double t = x * w;
                        s1 += t;
                        s2 += x * t;
                    } else {
                        s1 += x * w;
                    }
                }
            }
            double out;
            if (a.correction) {
                const double X = load<ST>(src, self);
                const double EX = s1 * rec;
                const double VarX = s2 * rec - EX * EX;
                out = VarX > a.varN ? X - (a.varN / VarX) * (X - EX) : EX;
            } else {
                out = s1 * rec;
            }
            acc_add(&acc[c * FW * FH + ai], out, a.fix);
        }
    }
    __syncthreads();
    // the tile's footprint into num; neighbouring tiles overlap by 8 - BlockStep
    for (int k = tid; k < np * FW * FH; k += kNlmThreads) {
        const u64 s = acc[k];
        if (s == 0) continue;
        const int c = k / (FW * FH), q = k - c * FW * FH;
        const int py = q / FW, px = q - py * FW;
        const int yy = oy + py, xx = ox + px;
        if (yy >= a.H || xx >= a.W) continue;
        atomicAdd(&a.num[c][(size_t)yy * a.pitch + xx], s);
    }
}

template <typename ST>
__global__ __launch_bounds__(kNlmThreads) void nlm_kernel(NlmArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int tx, ty;
    xcd_tile(tx, ty);
    nlm_tile<ST>(a, reinterpret_cast<u64 *>(smem), tx, ty);
}

// Number of grid positions min(k*s, last) in [x - 7, x]: the multiples of s
// below last, and last itself.
__device__ __forceinline__ int cover(int x, int s, int last)
{
    int c = (x >= last && x <= last + kBlock - 1) ? 1 : 0;
    const int lo = max(0, x - (kBlock - 1)), hi = min(x, last - 1);
    if (hi >= lo) c += max(0, hi / s - (lo + s - 1) / s + 1);
    return c;
}

__global__ void nlm_finalize_kernel(NlmFinalizeArgs a)
{
    // separate mul and add roundings, as the reference's double code
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= a.W) return;
    const size_t o = (size_t)y * a.pitch + x;
    const size_t oo = (size_t)y * a.out_pitch + x;
    const double den = (double)(cover(x, a.bstep, a.W - kBlock) * cover(y, a.bstep, a.H - kBlock));
    for (int c = 0; c < a.nplanes; ++c) {
        const double num = (double)(long long)a.num[c][o] * a.inv_fix;
        const double val = num / den;
        if (a.out_int) {
            double t = val * a.gain + a.offset;
            if (a.clip) t = t < a.lo ? a.lo : (t > a.hi ? a.hi : t);
            // uncovered pixels (BlockStep > BlockSize) are 0/0 there; x86's
            // conversion of NaN gives INT_MIN
            static_cast<int *>(a.out[c])[oo] = t == t ? (int)t : (int)0x80000000u;
        } else {
            static_cast<float *>(a.out[c])[oo] = (float)val;
        }
    }
}

template <typename K>
hipError_t set_lds(K kernel, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename ST>
hipError_t launch_nlm_t(const NlmArgs &a, hipStream_t s)
{
    const size_t lds = nlm_lds_bytes(a);
    hipError_t e;
    if ((e = set_lds(nlm_kernel<ST>, lds)) != hipSuccess) return e;
    const dim3 grid((a.ngx + a.trx - 1) / a.trx, (a.ngy + a.try_ - 1) / a.try_);
    hipLaunchKernelGGL(nlm_kernel<ST>, grid, dim3(kNlmThreads), lds, s, a);
    return hipGetLastError();
}

}  // namespace

size_t nlm_lds_bytes(const NlmArgs &a) { return (size_t)a.nplanes * a.FW * a.FH * sizeof(u64); }

hipError_t launch_nlm(const NlmArgs &a, hipStream_t s)
{
    switch (a.src_type) {
    case 0: return launch_nlm_t<float>(a, s);
    case 1: return launch_nlm_t<double>(a, s);
    case 2: return launch_nlm_t<int>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_nlm_finalize(const NlmFinalizeArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(nlm_finalize_kernel, dim3((a.W + 255) / 256, a.H), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace ispbm3d
