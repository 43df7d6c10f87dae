// Multi-frame grouping (gfx950): the merge of per-frame candidate lists into
// the groups of one reference frame.
//
// The reference carries Pos3 group codes (frame, y, x) and gathers /
// aggregates over a vector of planes (BlockGroup, include/Block.h:843-846,
// 927-936, 1141-1208) without forming such groups itself.  Here a group of
// reference block (t, j, i) is its BlockMatchingMulti extended to frames
// |z - t| <= radius: the same window, test and (key, scan order) sort, frames
// scanned in the order t, t-1, t+1, t-2, ...  match_kernel produces each
// frame's list (the own frame's with the block itself first, the others
// without it), each holding that frame's G-1 best; the group's G-1 best lie
// in their union, and a stable merge by (key, frame order, position in list)
// keeps the scan order among equal keys.
#include "bm3d_internal.h"

namespace ispbm3d {
namespace {

// One wave per reference block.  The lists' entries (the own list without
// its first entry, the block itself) are laid out in frame order, each list
// in its own order, and each entry's place in the merge is its rank: the
// number of entries with a smaller key, plus those with an equal key earlier
// in that layout.  Every list is sorted by (key, scan order), so this is the
// stable merge by (key, frame order, position in list).  All loads of a block
// are in flight together, instead of one dependent key load per output.
constexpr int kMergeWaves = 4;
constexpr int kMergeMax = kMaxStackLists * kMaxGroup;
constexpr int kMergePerLane = (kMergeMax + 63) / 64;

__global__ __launch_bounds__(64 * kMergeWaves) void stack_merge_kernel(StackMergeArgs a)
{
    __shared__ double skey[kMergeWaves][kMergeMax];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int r = (int)blockIdx.x * kMergeWaves + wave;
    if (r >= a.refs) return;
    const size_t base = (size_t)r * a.G;
    // entries of list l in the layout: len[l], after those of lists 0..l-1
    int len[kMaxStackLists];
    int E = 0;
#pragma unroll
    for (int l = 0; l < kMaxStackLists; ++l) {
        len[l] = 0;
        if (l < a.nlists) {
            const int c = a.cnt[l][r];
            len[l] = l == 0 ? max(c - 1, 0) : c;
        }
        E += len[l];
    }
    double *key = skey[wave];
    int mypos[kMergePerLane], myslot[kMergePerLane];
#pragma unroll
    for (int q = 0; q < kMergePerLane; ++q) {
        const int e = lane + 64 * q;
        if (e < E) {
            // this entry's list m and index in it
            int st = 0, m0 = 0, st0 = 0;
#pragma unroll
            for (int m = 0; m < kMaxStackLists; ++m) {
                if (e >= st && e < st + len[m]) {
                    m0 = m;
                    st0 = st;
                }
                st += len[m];
            }
            const int *pl = a.pos[0];
            const double *kl = a.key[0];
            int sl = a.slot[0];
#pragma unroll
            for (int m = 1; m < kMaxStackLists; ++m)
                if (m0 == m) {
                    pl = a.pos[m];
                    kl = a.key[m];
                    sl = a.slot[m];
                }
            const size_t o = base + (size_t)(e - st0 + (m0 == 0 ? 1 : 0));
            key[e] = kl[o];
            mypos[q] = pl[o];
            myslot[q] = sl;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int q = 0; q < kMergePerLane; ++q) {
        const int e = lane + 64 * q;
        if (e >= E) continue;
        const double k = key[e];
        int rank = 0;
        for (int f = 0; f < E; ++f) {
            const double kf = key[f];
            rank += (kf < k || (kf == k && f < e)) ? 1 : 0;
        }
        if (rank < a.G - 1) {
            a.pos_out[base + 1 + rank] = mypos[q];
            a.slot_out[base + 1 + rank] = myslot[q];
        }
    }
    if (lane == 0) {
        a.pos_out[base] = a.pos[0][base];
        a.slot_out[base] = a.slot[0];
        a.cnt_out[r] = 1 + min(E, a.G - 1);
    }
}

}  // namespace

hipError_t launch_stack_merge(const StackMergeArgs &a, hipStream_t s)
{
    if (a.nlists < 1 || a.nlists > kMaxStackLists) return hipErrorInvalidValue;
    hipLaunchKernelGGL(stack_merge_kernel, dim3((a.refs + kMergeWaves - 1) / kMergeWaves), dim3(64 * kMergeWaves), 0, s, a);
    return hipGetLastError();
}

}  // namespace ispbm3d
