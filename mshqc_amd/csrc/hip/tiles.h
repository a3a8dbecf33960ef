// Output-tile scheduling shared by the GEMM kernels.
#pragma once
#include <hip/hip_runtime.h>

// Maps blockIdx.x of a 1-D grid (a multiple of 8 workgroups) to an output
// tile (mt, nt) of an nt_m x nt_n tile grid.  The dispatcher deals
// workgroups to the 8 XCDs round-robin; the remap gives every XCD a
// contiguous range of tiles, and within it tiles are swept in groups of GM
// row-tiles, column by column, so the workgroups resident on one XCD at a
// time share GM row panels and ~8 column panels in that XCD's L2 instead of
// each streaming its own from memory.  GM is at most 8, chosen so that the
// groups are as even as possible (19 row-tiles: 7, 7, 5 rather than 8, 8,
// 3, whose last narrow group would fill the last XCDs' whole ranges).
// Returns false for padding workgroups.
__device__ __forceinline__ bool grouped_tile(int nt_m, int nt_n, int& mt, int& nt) {
    const int ng = (nt_m + 7) / 8;
    const int GM = (nt_m + ng - 1) / ng;
    const long total = (long)nt_m * nt_n;
    const long per = gridDim.x >> 3;
    const long w = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (w >= total) return false;
    const long group_size = (long)GM * nt_n;
    const int first_m = (int)(w / group_size) * GM;
    const int gm = min(GM, nt_m - first_m);
    const long r = w % group_size;
    mt = first_m + (int)(r % gm);
    nt = (int)(r / gm);
    return true;
}

inline unsigned grouped_grid(int nt_m, int nt_n) {
    const long total = (long)nt_m * nt_n;
    return (unsigned)(((total + 7) / 8) * 8);
}
