// half.h -- half-shell cell tiles for the Newton's-third-law pair pass (FP32).
//
// A half tile is the home cell c (segment 0) followed by the 13 "forward" cells of
// its 27-cell stencil, the offsets (a, b, g) with g > 0, or g = 0 and b > 0, or
// g = b = 0 and a > 0.  With at least three cells per axis every unordered pair
// of neighbouring cells is (home, forward) of exactly one tile, so each pair of
// atoms within the list radius sits in exactly one tile: in the row of the home
// atom i for a forward atom j, or, both in the home cell, in the row of the one
// with the smaller tile-local index.  The pair pass evaluates each pair once and
// applies +F to i and -F to j (the reference evaluates every pair from both
// sides, ComputeNearForceVerletWorklet, rbmd/framework/run/worklet/RunWorklet.cpp:170-237).
//
// The forces on the tile's atoms accumulate in LDS with FP64 atomic adds of values
// rounded to a 2^-30 grid: sums of such values below 2^22 are exact in FP64, so the
// result does not depend on the order the atomics land in, and repeated runs give
// the same bits.  Each tile writes the sums of its 14 segments to 14 force planes
// (plane k holds, for every atom of a forward-k cell, what the tile behind it added;
// plane 0 the home atoms' own sums), every plane entry written by exactly one tile;
// a per-atom pass then adds the 14 planes in a fixed order (PlaneSum, pair.hip).
#pragma once
#include "tile.h"

namespace rbmd {

constexpr int kHalfSegs = 14;

struct HalfHdr {
    int off[kHalfSegs + 1];  // segment offsets in the concatenation, off[14] = total
    int gstart[kHalfSegs];   // first cell-sorted index of each segment
    int cnt[kHalfSegs];
    double sx[kHalfSegs], sy[kHalfSegs], sz[kHalfSegs];
    double4 dummy;           // the row sentinel's position: beyond every cutoff, charge 0
};

// Grid of the order-independent LDS sums: x rounded to a multiple of 2^-30
// ((x + 1.5 2^22) - 1.5 2^22, exact for |x| < 2^21).
constexpr double kHalfGridMagic = 6291456.0;  // 1.5 * 2^22

#ifdef __HIPCC__
__device__ inline double half_grid(double x)
{
#pragma clang fp contract(off)
    return (x + kHalfGridMagic) - kHalfGridMagic;
}

// Fill the half-tile header of home cell c (all threads of the block call it).
// Stencil slot t = 9(a+1) + 3(b+1) + (g+1) is forward when (g, b, a) > (0, 0, 0)
// lexicographically; forward slots take segments 1..13 in slot order.
__device__ inline void half_setup(int c, int3 dims, const Box& box,
                                  const int32_t* __restrict__ cell_start, HalfHdr& h)
{
    const int t = threadIdx.x;
    if (t < 64) {
        const int a = t / 9 - 1, b = (t / 3) % 3 - 1, g = t % 3 - 1;
        const bool fwd = t < 27 && (g > 0 || (g == 0 && (b > 0 || (b == 0 && a > 0))));
        const uint64_t mf = __ballot(fwd);
        const int seg = t == kHomeSeg ? 0 : (fwd ? 1 + (int)mbcnt64(mf) : -1);
        int cnt = 0;
        if (seg >= 0) {
            const int cz = c / (dims.x * dims.y);
            const int cy = (c / dims.x) % dims.y;
            const int cx = c % dims.x;
            int ax = cx + a, ay = cy + b, az = cz + g;
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
            h.gstart[seg] = g0;
            h.cnt[seg] = cnt;
            h.sx[seg] = sx;
            h.sy[seg] = sy;
            h.sz[seg] = sz;
        }
        // segment sizes in segment order: lane j < 14 fetches segment j's count
        __syncwarp();
        const int sc = t < kHalfSegs ? h.cnt[t] : 0;
        int incl = sc;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (t >= o) incl += v;
        }
        if (t < kHalfSegs) h.off[t] = incl - sc;
        if (t == kHalfSegs - 1) h.off[kHalfSegs] = incl;
        if (t == 0)
            h.dummy = make_double4(box.lo[0] + 8.0 * box.L[0] + 64.0,
                                   box.lo[1] + 8.0 * box.L[1] + 64.0,
                                   box.lo[2] + 8.0 * box.L[2] + 64.0, 0.0);
    }
    __syncthreads();
}

// Segment holding tile-local index u of a half tile.
__device__ inline int half_seg(const HalfHdr& h, int u)
{
    int lo = 0, hi = kHalfSegs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (h.off[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ inline double4 half_ref(const HalfHdr& h, const double4* __restrict__ xs_ref, int u)
{
#pragma clang fp contract(off)
    const int st = half_seg(h, u);
    const double4 p = xs_ref[h.gstart[st] + (u - h.off[st])];
    return make_double4(p.x - h.sx[st], p.y - h.sy[st], p.z - h.sz[st], p.w);
}

// Half rows of the home atoms of a half tile: for the home atom at tile index ui,
// the candidates u in (ui, total) -- the later home atoms and every forward atom --
// with the criterion and distance convention of list_wave (tile.h); rows in the same
// layout (LJ-range entries, the rest, the sentinel `total`).  Staged: the reference
// stage holds the tile and its dummy up to index total.
template <bool STAGED>
__device__ inline void list_wave_half(const HalfHdr& h, const RefStage& rs,
                                      const double4* __restrict__ xs_ref, int first, int a1,
                                      int h0, int lane, double rl2, double rin2, int cap,
                                      uint16_t* hb, uint16_t* __restrict__ nlist,
                                      int2* __restrict__ ncount,
                                      unsigned long long* __restrict__ stats)
{
#pragma clang fp contract(off)
    const int total = h.off[kHalfSegs];
    for (int s = first; s < a1; s += kTileWaves) {
        const int ui = s - h0;  // home segment first
        if (total >= kMaxTile) {
            if (lane == 0) {
                ncount[s] = make_int2(cap + 1, 0);
                atomicAdd(&stats[2], 1ull);
                atomicMax(&stats[4], (unsigned long long)(cap + 1));
            }
            continue;
        }
        double3 pi;
        if (STAGED) {
            pi = rs.get(ui);
        } else {
            const double4 p = xs_ref[s];
            pi = make_double3(p.x, p.y, p.z);
        }
        int n = 0;
        for (int ub = (ui + 1) & ~63; ub < total; ub += kSweepLen) {
            double3 pj[kSweep];
#pragma unroll
            for (int k = 0; k < kSweep; ++k) {
                const int u = ub + 64 * k + lane;
                const int uu = (u > ui && u < total) ? u : total;
                if (STAGED) {
                    pj[k] = rs.get(uu);
                } else {
                    const double4 p = uu < total ? half_ref(h, xs_ref, uu) : h.dummy;
                    pj[k] = make_double3(p.x, p.y, p.z);
                }
            }
#pragma unroll
            for (int k = 0; k < kSweep; ++k) {
                const int u = ub + 64 * k + lane;
                const double dx = pj[k].x - pi.x;
                const double dy = pj[k].y - pi.y;
                const double dz = pj[k].z - pi.z;
                const double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
                const bool live = u > ui && u < total;
                const bool hit = live && ((rl2 - d2) > 1e-4) && (d2 > 1e-4);
                const uint64_t mh = __builtin_amdgcn_ballot_w64(live) &
                                    __builtin_amdgcn_ballot_w64((rl2 - d2) > 1e-4) &
                                    __builtin_amdgcn_ballot_w64(d2 > 1e-4);
                const int slot = n + (int)mbcnt64(mh);
                if (hit && slot < cap) hb[slot] = (uint16_t)(d2 < rin2 ? (u | kInnerFlag) : u);
                n += __popcll(mh);
            }
        }
        uint16_t* __restrict__ row = nlist + (size_t)s * cap;
        int n_in = n;
        if (n < cap) {
            __syncwarp();
            n_in = 0;
            for (int k0 = 0; k0 < n; k0 += 64) {
                const int k = k0 + lane;
                n_in += __popcll(__ballot(k < n && (hb[k] & kInnerFlag) != 0));
            }
            int ci = 0, co = n_in;
            for (int k0 = 0; k0 < n; k0 += 64) {
                const int k = k0 + lane;
                const int e = k < n ? (int)hb[k] : 0;
                const bool in = k < n && (e & kInnerFlag) != 0;
                const bool out = k < n && (e & kInnerFlag) == 0;
                const uint64_t mi = __ballot(in), mo = __ballot(out);
                if (in) row[ci + (int)mbcnt64(mi)] = (uint16_t)(e & kMaxTile);
                if (out) row[co + (int)mbcnt64(mo)] = (uint16_t)e;
                ci += __popcll(mi);
                co += __popcll(mo);
            }
            if (lane == 0) row[n] = (uint16_t)total;
            __syncwarp();
        }
        if (lane == 0) {
            ncount[s] = n < cap ? make_int2(n_in, n - n_in) : make_int2(n, 0);
            if (n >= cap) {
                atomicAdd(&stats[2], 1ull);
                atomicMax(&stats[4], (unsigned long long)(n + 1));
            }
        }
    }
}
#endif

}  // namespace rbmd
