// tile.h -- cell tiles shared by the Verlet-list build and the pair pass.
//
// A tile is one home cell c plus its 27-cell periodic stencil.  The stencil's
// atoms (cell-sorted, so each stencil cell is one contiguous range) are
// concatenated in stencil order st = 9(a+1) + 3(b+1) + (g+1), a,b,g in {-1,0,1};
// a candidate's position in that concatenation is its tile-local index u.
// Verlet rows store u as uint16, and the whole concatenation is staged in LDS
// once per workgroup, so the pair pass reads neighbour positions from LDS.
// Each stencil cell carries the periodic image shift s (k L, k in {-1,0,1})
// such that x_j - s is the image of x_j next to the home cell.
#pragma once
#include "engine.h"

namespace rbmd {

constexpr int kTileThreads = 512;
constexpr int kTileWaves = kTileThreads / 64;
constexpr int kHomeSeg = 13;  // stencil slot of the home cell (a = b = g = 0)

struct TileHdr {
    int off[28];     // segment offsets in the concatenation, off[27] = total
    int gstart[27];  // first cell-sorted index of each segment
    int cnt[27];
    double sx[27], sy[27], sz[27];
};

#ifdef __HIPCC__
// Fill the tile header of home cell c (all threads of the block call it).
__device__ inline void tile_setup(int c, int3 dims, const Box& box,
                                  const int32_t* __restrict__ cell_start, TileHdr& h)
{
    const int t = threadIdx.x;
    if (t < 64) {
        int cnt = 0;
        if (t < 27) {
            const int cz = c / (dims.x * dims.y);
            const int cy = (c / dims.x) % dims.y;
            const int cx = c % dims.x;
            int ax = cx + t / 9 - 1, ay = cy + (t / 3) % 3 - 1, az = cz + t % 3 - 1;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            if (ax < 0) { ax += dims.x; sx = box.L[0]; }
            else if (ax >= dims.x) { ax -= dims.x; sx = -box.L[0]; }
            if (ay < 0) { ay += dims.y; sy = box.L[1]; }
            else if (ay >= dims.y) { ay -= dims.y; sy = -box.L[1]; }
            if (az < 0) { az += dims.z; sz = box.L[2]; }
            else if (az >= dims.z) { az -= dims.z; sz = -box.L[2]; }
            const int cc = ax + ay * dims.x + az * dims.x * dims.y;
            const int g0 = cell_start[cc];
            cnt = cell_start[cc + 1] - g0;
            h.gstart[t] = g0;
            h.cnt[t] = cnt;
            h.sx[t] = sx;
            h.sy[t] = sy;
            h.sz[t] = sz;
        }
        // wave-level inclusive scan of the 27 segment sizes
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (t >= o) incl += v;
        }
        if (t < 27) h.off[t] = incl - cnt;
        if (t == 26) h.off[27] = incl;
    }
    __syncthreads();
}

// Segment holding tile-local index u (largest st with off[st] <= u; empty
// segments before it have smaller st, empty ones after it start above u).
__device__ inline int tile_seg(const TileHdr& h, int u)
{
    int lo = 0, hi = 26;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (h.off[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
#endif

}  // namespace rbmd
