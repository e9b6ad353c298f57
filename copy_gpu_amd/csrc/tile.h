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
    double4 dummy;   // position of the row sentinel `total`: beyond every cutoff, charge 0
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
        if (t == 0)
            h.dummy = make_double4(box.lo[0] + 8.0 * box.L[0] + 64.0,
                                   box.lo[1] + 8.0 * box.L[1] + 64.0,
                                   box.lo[2] + 8.0 * box.L[2] + 64.0, 0.0);
    }
    __syncthreads();
}

// ---------------------------------------------------------------- binning
// Wrap into the box (x - L floor((x-lo)/L), no contraction: the CPU oracle
// computes the same bits) and the cell trunc((x-lo)/cell), clamped.
__device__ inline int wrap_cell(const double p[3], const Box& box, int3 dims, double3 cell,
                                double w[3])
{
#pragma clang fp contract(off)
    const int dm[3] = {dims.x, dims.y, dims.z};
    const double cs[3] = {cell.x, cell.y, cell.z};
    int c[3];
    for (int d = 0; d < 3; ++d) {
        double f = floor((p[d] - box.lo[d]) / box.L[d]);
        double fl = f * box.L[d];
        w[d] = p[d] - fl;
        int v = (int)((w[d] - box.lo[d]) / cs[d]);
        v = v < 0 ? 0 : v;
        v = v > dm[d] - 1 ? dm[d] - 1 : v;
        c[d] = v;
    }
    return c[0] + c[1] * dims.x + c[2] * dims.x * dims.y;
}

// Wave-aggregated binning for lanes that mostly share one cell (the home atoms
// of a tile): one atomic per distinct cell of the wave, ranks by ballot order.
// Atom v sits in slot s.  Every lane of the wave calls it; inactive lanes bin
// nothing.
__device__ inline void md_bin_wave(const MdBins& b, int buf, bool active, int s, int v,
                                   const double p[3], const Box& box, int3 dims, double3 cell)
{
    int c = -1;
    if (active) {
        double w[3];
        c = wrap_cell(p, box, dims, cell, w);
        b.cell[s] = c;
        b.vid[s] = v;
        b.xw[s] = w[0];
        b.yw[s] = w[1];
        b.zw[s] = w[2];
    }
    const int lane = (int)__lane_id();
    uint64_t pending = __ballot(active);
    while (pending != 0ull) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int cl = __shfl(c, leader, 64);
        const uint64_t grp = __ballot(c == cl) & pending;
        int base = 0;
        if (lane == leader) {
            base = atomicAdd(&b.count[buf * b.ncell + cl], __popcll(grp));
            if (b.nchunk > 0) atomicAdd(&b.chunk[buf * b.nchunk + cl / kBinChunk], __popcll(grp));
        }
        base = __shfl(base, leader, 64);
        if ((grp >> lane) & 1ull) {
            const int rank = base + (int)mbcnt64(grp);
            if (rank < b.cap) b.bucket[((size_t)buf * b.ncell + cl) * b.cap + rank] = s;
        }
        pending &= ~grp;
    }
}

__device__ inline void md_bin_atom(const MdBins& b, int buf, int s, int v, const double p[3],
                                   const Box& box, int3 dims, double3 cell)
{
    double w[3];
    const int c = wrap_cell(p, box, dims, cell, w);
    b.cell[s] = c;
    b.vid[s] = v;
    b.xw[s] = w[0];
    b.yw[s] = w[1];
    b.zw[s] = w[2];
    const int bc = buf * b.ncell + c;
    const int rank = atomicAdd(&b.count[bc], 1);
    if (rank < b.cap) b.bucket[(size_t)bc * b.cap + rank] = s;
}

// ---------------------------------------------------------------- tile order
// Order of the tile workgroups, for load balance only (a tile's results do not
// depend on which workgroup computes it).  Used when the tiles fit about one
// dispatch round (ncell * split <= 3 ncu): then the kernel ends when its most
// loaded CU does, and the dispatcher gives workgroup t to CU t mod ncu.  The
// work of cell c is estimated as w_c = n_c x (atoms of its 27-cell stencil);
// cells are counting-sorted into kOrderBuckets classes of w, heaviest first
// (order within a class is arbitrary), and the first round takes the heaviest
// tiles in reverse order, so a CU that starts with a lighter tile receives
// heavier ones in the later rounds.  With more tiles the dispatcher balances by
// itself and the cell order (spatial, L2-friendly) is kept.  One workgroup of
// at least 256 threads runs it, ncell <= kOrderMaxCells; lds holds kOrderLds ints.
constexpr int kOrderBuckets = 256;
constexpr int kOrderMaxCells = 1024;
constexpr int kOrderLds = 2 * kOrderBuckets + 8 + 2 * kOrderMaxCells;

__device__ inline void tile_order_build(const int32_t* __restrict__ count, int ncell, int3 dims,
                                        int split, int ncu, int32_t* __restrict__ order,
                                        int32_t* lds)
{
    const int nt = (int)blockDim.x, t = (int)threadIdx.x;
    int32_t* hist = lds;
    int32_t* cur = lds + kOrderBuckets;
    int32_t* misc = lds + 2 * kOrderBuckets;  // [0] max weight, [1..4] wave sums
    int32_t* cnt = misc + 8;
    int32_t* w = cnt + kOrderMaxCells;
    for (int k = t; k < kOrderBuckets; k += nt) hist[k] = 0;
    for (int c = t; c < ncell; c += nt) cnt[c] = count[c];
    if (t == 0) misc[0] = 0;
    __syncthreads();
    int mx = 0;
    for (int c = t; c < ncell; c += nt) {
        const int cz = c / (dims.x * dims.y), cy = (c / dims.x) % dims.y, cx = c % dims.x;
        int sum = 0;
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
                for (int g = -1; g <= 1; ++g) {
                    const int x = (cx + a + dims.x) % dims.x, y = (cy + b + dims.y) % dims.y,
                              z = (cz + g + dims.z) % dims.z;
                    sum += cnt[x + y * dims.x + z * dims.x * dims.y];
                }
        const int wc = cnt[c] * sum;
        w[c] = wc;
        mx = wc > mx ? wc : mx;
    }
    atomicMax(&misc[0], mx);
    __syncthreads();
    const long long scale = (long long)misc[0] + 1;
    auto bucket = [&](int wc) {
        return kOrderBuckets - 1 - (int)(((long long)wc * kOrderBuckets) / scale);
    };
    for (int c = t; c < ncell; c += nt) atomicAdd(&hist[bucket(w[c])], 1);
    __syncthreads();
    // exclusive prefix over the buckets: threads [0, 256) one bucket each
    // (4 waves: wave scans, then the waves' totals)
    int v = 0, incl = 0;
    if (t < kOrderBuckets) {
        v = hist[t];
        incl = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if ((t & 63) >= o) incl += u;
        }
        if ((t & 63) == 63) misc[1 + (t >> 6)] = incl;
    }
    __syncthreads();
    if (t < kOrderBuckets) {
        int off = 0;
        for (int k = 0; k < (t >> 6); ++k) off += misc[1 + k];
        cur[t] = off + incl - v;
    }
    __syncthreads();
    const int ntile = ncell * split;
    const int r0 = ncu < ntile ? ncu : ntile;
    for (int c = t; c < ncell; c += nt) {
        const int pos = atomicAdd(&cur[bucket(w[c])], 1);
        for (int p = 0; p < split; ++p) {
            const int slot = pos * split + p;
            order[slot < r0 ? r0 - 1 - slot : slot] = c * split + p;
        }
    }
}

// Whether a system's tiles get the load-balanced order (ncu passed to the
// builder) or keep the cell order (0).
inline int tile_order_ncu(int ncell, int split, int num_cu)
{
    return (ncell <= kOrderMaxCells && ncell * split <= 3 * num_cu) ? num_cu : 0;
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
// ---------------------------------------------------------------- Verlet list
// One workgroup per (home cell, part): the tile's reference positions are
// staged in LDS already moved to the home cell's image, x_j - s (s the image
// shift of j's stencil cell), then one wave per home atom i sweeps the whole
// tile, kSweep chunks of 64 candidates per iteration.  A candidate is kept iff
// (r_list^2 - d2 > 1e-4) && (d2 > 1e-4) with dx = (x_j - s) - x_i and
// d2 = fma(dz,dz, fma(dy,dy, dx*dx)) -- the reference's operation order
// (RunWorklet.cpp:77-83) and the exact bits the CPU oracle uses.  The stage is
// two arrays, (x, y) pairs then z, so the sweep's consecutive reads are
// bank-conflict-free.  Tiles larger than the LDS stage read candidates from
// global memory; tiles beyond the 15-bit index range mark their rows overflowed.
//
// Row layout (uint16 tile-local indices, cap slots per home atom): the LJ-range
// entries (d2 < r_in^2) first, then the others, then one sentinel slot holding
// the tile's dummy index `total` (staged far from every atom), so a reader may
// clamp its slot to the row length instead of testing it.  A row of n entries
// needs n + 1 <= cap slots; ncount = (LJ-range, other) entries, or (n, 0) with
// n >= cap for a row that did not fit.
//
// The sweep compacts hits with one ballot + mbcnt into this wave's LDS hit
// buffer (cap entries, LJ-range hits tagged with kInnerFlag), then one pass
// over the buffer writes the row in its final order.
constexpr int kInnerFlag = 0x8000;
constexpr int kMaxTile = 0x7fff;  // tile-local indices and the sentinel fit in 15 bits

__device__ inline double4 stage_ref(const TileHdr& h, const double4* __restrict__ xs_ref, int u)
{
#pragma clang fp contract(off)
    const int st = tile_seg(h, u);
    const double4 p = xs_ref[h.gstart[st] + (u - h.off[st])];
    return make_double4(p.x - h.sx[st], p.y - h.sy[st], p.z - h.sz[st], p.w);
}

// Reference-position stage of list_wave: (x, y) of candidate u at xy[u], z at
// zz[u], for u below the stage size rounded up to the sweep's kSweepLen candidates
// (entries past the tile hold its dummy position, so the sweep needs no bounds
// test).  The per-wave hit buffers follow (hit_buffer).
constexpr int kSweep = 4;  // 64-candidate chunks per sweep iteration (independent chains)
constexpr int kSweepLen = 64 * kSweep;
__host__ __device__ inline int ref_stage_len(int stage_cap)
{
    return (stage_cap + kSweepLen - 1) / kSweepLen * kSweepLen;
}

struct RefStage {
    double2* xy;
    double* zz;
    __device__ RefStage(double4* lds, int stage_cap)
        : xy((double2*)lds), zz((double*)lds + 2 * (size_t)ref_stage_len(stage_cap))
    {
    }
    __device__ void put(int u, const double4& p) const
    {
        xy[u] = make_double2(p.x, p.y);
        zz[u] = p.z;
    }
    __device__ double3 get(int u) const
    {
        const double2 a = xy[u];
        return make_double3(a.x, a.y, zz[u]);
    }
};

// LDS bytes of the list build: the reference stage plus one hit buffer of cap
// uint16 per wave (and, for a random batch list, the candidates' atom ids).
inline size_t list_lds_bytes(int stage_cap, int cap, bool rbl = false)
{
    return 24 * (size_t)ref_stage_len(stage_cap) +
           (size_t)kTileWaves * (size_t)cap * sizeof(uint16_t) +
           (rbl ? 4 * (size_t)ref_stage_len(stage_cap) : 0);
}

__device__ inline uint16_t* hit_buffer(double4* lds, int stage_cap, int cap, int wv)
{
    return (uint16_t*)((double*)lds + 3 * (size_t)ref_stage_len(stage_cap)) + (size_t)wv * cap;
}

// Atom ids of the staged candidates (random batch list builds), after the hit buffers.
__device__ inline int32_t* stage_ids(double4* lds, int stage_cap, int cap)
{
    return (int32_t*)(hit_buffer(lds, stage_cap, cap, kTileWaves));
}

// Random batch list of a build: rows keep every pair closer than r_near = r_core +
// skin (it may come inside r_core before the next build) and the shell pairs that
// rbl_select keeps, tagged with kRblSel so the pair pass weights them.
constexpr int kRblSel = 0x8000;
struct RblBuild {
    double near2;      // (r_core + skin)^2
    uint64_t seed;
    uint32_t epoch;    // list build number
    uint32_t pice;
    const int32_t* perm;  // sorted slot -> atom id
    const int32_t* ids;   // LDS: atom id of each staged candidate (null: from perm)
};

// Atom id of tile-local candidate u (random batch list builds and their fallback).
__device__ inline int32_t tile_atom_id(const TileHdr& h, const int32_t* __restrict__ perm, int u)
{
    const int st = tile_seg(h, u);
    return perm[h.gstart[st] + (u - h.off[st])];
}

template <bool STAGED, bool RBL = false>
__device__ inline void list_wave(const TileHdr& h, const RefStage& rs,
                                 const double4* __restrict__ xs_ref, int first, int a1, int h0,
                                 int lane, double rl2, double rin2, int cap, uint16_t* hb,
                                 uint16_t* __restrict__ nlist, int2* __restrict__ ncount,
                                 unsigned long long* __restrict__ stats,
                                 const RblBuild& rb = RblBuild{})
{
#pragma clang fp contract(off)
    const int total = h.off[27];
    for (int s = first; s < a1; s += kTileWaves) {
        if (total >= kMaxTile) {
            if (lane == 0) {
                ncount[s] = make_int2(cap + 1, 0);
                atomicAdd(&stats[2], 1ull);
                atomicMax(&stats[4], (unsigned long long)(cap + 1));
            }
            continue;
        }
        const int ui = h.off[kHomeSeg] + (s - h0);
        double3 pi;
        if (STAGED) {
            pi = rs.get(ui);
        } else {
            const double4 p = xs_ref[s];
            pi = make_double3(p.x, p.y, p.z);
        }
        int n = 0;
        for (int ub = 0; ub < total; ub += kSweepLen) {
            double3 pj[kSweep];
#pragma unroll
            for (int k = 0; k < kSweep; ++k) {
                const int u = ub + 64 * k + lane;
                const int uu = STAGED || u < total ? u : 0;
                if (STAGED) {
                    pj[k] = rs.get(uu);
                } else {
                    const double4 p = stage_ref(h, xs_ref, uu);
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
                // staged: padding entries past the tile are far, so they miss
                const bool live = STAGED || u < total;
                const bool hit = live && ((rl2 - d2) > 1e-4) && (d2 > 1e-4);
                // the same mask as ballot(hit), formed from the compare masks
                // directly (no per-lane boolean round trip)
                const uint64_t mh = __builtin_amdgcn_ballot_w64(live) &
                                    __builtin_amdgcn_ballot_w64((rl2 - d2) > 1e-4) &
                                    __builtin_amdgcn_ballot_w64(d2 > 1e-4);
                const int slot = n + (int)mbcnt64(mh);
                // tag: LJ range (full list) or the always-kept near band (random batch)
                const double tag2 = RBL ? rb.near2 : rin2;
                if (hit && slot < cap) hb[slot] = (uint16_t)(d2 < tag2 ? (u | kInnerFlag) : u);
                n += __popcll(mh);
            }
        }
        uint16_t* __restrict__ row = nlist + (size_t)s * cap;
        int n_in = n;
        if (RBL && n < cap) {
            // random batch list: the near band, and the shell pairs this list build
            // selects (tagged kRblSel), in sweep order; the pair pass evaluates LJ on
            // every entry
            __syncwarp();
            const int32_t id_i = rb.perm[s];
            int nk = 0;
            for (int k0 = 0; k0 < n; k0 += 64) {
                const int k = k0 + lane;
                const int e = k < n ? (int)hb[k] : 0;
                const int u = e & kMaxTile;
                // every hit draws its selection: a near-band entry kept for being near
                // counts only if selected should it lie in the shell when evaluated
                bool sel = false;
                if (k < n) {
                    const int32_t id_j = rb.ids ? rb.ids[u] : tile_atom_id(h, rb.perm, u);
                    sel = rbl_select(rb.seed, rb.epoch, id_i, id_j, rb.pice);
                }
                const bool keep = k < n && ((e & kInnerFlag) != 0 || sel);
                const uint64_t mk = __ballot(keep);
                if (keep) row[nk + (int)mbcnt64(mk)] = (uint16_t)(sel ? (u | kRblSel) : u);
                nk += __popcll(mk);
            }
            if (lane == 0) row[nk] = (uint16_t)total;
            __syncwarp();
            n_in = n = nk;
        } else if (n < cap) {
            __syncwarp();  // the hit buffer, written by every lane of this wave
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
            __syncwarp();  // the buffer is reused for the wave's next atom
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
