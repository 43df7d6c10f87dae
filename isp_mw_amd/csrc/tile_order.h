// Tile order of a 2D launch on the MI355X's 8 XCDs.
#pragma once

#include <hip/hip_runtime.h>

namespace ispbm3d {

// Consecutive tiles (row-major) that one XCD processes together.
constexpr int kXcdRun = 8;

// Workgroups are dispatched to the 8 XCDs in turn (linear id w to XCD w % 8),
// and each XCD has its own L2.  Here each block of 8 * kXcdRun tiles in
// row-major order is split into 8 runs of kXcdRun neighbouring tiles, run k
// going to XCD k: tiles whose footprints overlap are read through one L2,
// while every XCD still takes work from the whole image (the tiles' cost
// varies with the image content).  The last partial block keeps the plain
// order.  A bijection of the grid's tiles: every tile is processed once.
__device__ __forceinline__ void xcd_tile(int &tx, int &ty)
{
    constexpr int kBlockTiles = 8 * kXcdRun;
    const int gx = (int)gridDim.x, n = gx * (int)gridDim.y;
    const int w = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    int t = w;
    if (w < n - n % kBlockTiles) {
        const int b = w / kBlockTiles, wi = w % kBlockTiles;
        t = b * kBlockTiles + (wi & 7) * kXcdRun + (wi >> 3);
    }
    tx = t % gx;
    ty = t / gx;
}

}  // namespace ispbm3d
