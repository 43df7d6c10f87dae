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

__global__ void stack_merge_kernel(StackMergeArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.refs) return;
    const size_t base = (size_t)r * a.G;
    int cur[kMaxStackLists], end[kMaxStackLists];
    for (int l = 0; l < a.nlists; ++l) {
        cur[l] = l == 0 ? 1 : 0;  // the own list's first entry is the block itself
        end[l] = a.cnt[l][r];
    }
    a.pos_out[base] = a.pos[0][base];
    a.slot_out[base] = a.slot[0];
    int n = 1;
    while (n < a.G) {
        int best = -1;
        double bk = 0.0;
        for (int l = 0; l < a.nlists; ++l) {
            if (cur[l] >= end[l]) continue;
            const double k = a.key[l][base + cur[l]];
            if (best < 0 || k < bk) {  // strict: equal keys keep the earlier frame
                best = l;
                bk = k;
            }
        }
        if (best < 0) break;
        a.pos_out[base + n] = a.pos[best][base + cur[best]];
        a.slot_out[base + n] = a.slot[best];
        ++cur[best];
        ++n;
    }
    a.cnt_out[r] = n;
}

}  // namespace

hipError_t launch_stack_merge(const StackMergeArgs &a, hipStream_t s)
{
    if (a.nlists < 1 || a.nlists > kMaxStackLists) return hipErrorInvalidValue;
    hipLaunchKernelGGL(stack_merge_kernel, dim3((a.refs + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace ispbm3d
