// nlist.hip -- cell list (counting sort) of the force pass, per-step displacement
// check and position gather, the MD-run binning prologue and list statistics.
// (The Verlet rows are built inside the tile pass, pair.hip; the MD-step kernels
// are in md.h / md.hip.)
//
// Reference behaviour followed (paths relative to the reference checkout):
//   cell grid / binning / sort    rbmd/framework/locator/ContPointLocator.{h,cxx} (h:44-53, cxx:38-92)
//   27-cell stencil + periodic    rbmd/framework/locator/ExecPointLocator.h:303-319
//   list criterion                rbmd/framework/run/worklet/RunWorklet.cpp:40-96 (:83)
// The list radius adds a Verlet skin; the list is rebuilt on the device when an
// atom has moved more than half the skin since the last build.
#include "state.h"
#include "tile.h"

namespace rbmd {

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------- per-step check
// Max-displacement test against the positions of the last build; zeroes the
// cell histogram so that a rebuild can follow.
struct Check {
    static constexpr int kThreads = kBlock;
    struct Args {
        const double *x, *y, *z, *xr, *yr, *zr;
        int n;
        Box box;
        double thr2;
        uint32_t* flags;
        int32_t* cell_count;
        int ncell;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        const int i = blk * kBlock + threadIdx.x;
        if (i < a.ncell) a.cell_count[i] = 0;
        if (i < a.n) {
            double dx = min_image(a.x[i] - a.xr[i], a.box.L[0], a.box.invL[0]);
            double dy = min_image(a.y[i] - a.yr[i], a.box.L[1], a.box.invL[1]);
            double dz = min_image(a.z[i] - a.zr[i], a.box.L[2], a.box.invL[2]);
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > a.thr2) a.flags[0] = 1u;
            if (d2 != d2) a.flags[1] = 1u;
        }
    }
};

__global__ void k_set_flag(uint32_t* flags) { flags[0] = 1u; }

// ---------------------------------------------------------------- binning
// Rank inside the cell from an atomic histogram.
struct Bin {
    static constexpr int kThreads = kBlock;
    struct Args {
        const double *x, *y, *z;
        double *xr, *yr, *zr;
        int n;
        Box box;
        int3 dims;
        double3 cell;
        int32_t *cell_count, *atom_cell, *atom_rank;
        const uint32_t* flags;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        if (a.flags[0] == 0u) return;
        const int i = blk * kBlock + threadIdx.x;
        if (i >= a.n) return;
        const double p[3] = {a.x[i], a.y[i], a.z[i]};
        double w[3];
        const int cid = wrap_cell(p, a.box, a.dims, a.cell, w);
        a.xr[i] = w[0];
        a.yr[i] = w[1];
        a.zr[i] = w[2];
        a.atom_cell[i] = cid;
        a.atom_rank[i] = atomicAdd(&a.cell_count[cid], 1);
    }
};

// ---------------------------------------------------------------- scan
// Exclusive scan of the cell histogram by one workgroup (wave prefix sums
// combined through LDS, chunks of 1024 with a running carry).
struct Scan {
    static constexpr int kThreads = 1024;
    struct Args {
        const int32_t* count;
        int32_t* start;
        int ncell;
        uint32_t* flags;
        unsigned long long* stats;
        int3 dims;
        int split, ncu;  // ncu 0: keep the cell order
        int32_t* tile_order;
    };
    __device__ static void run(const Args& a, int, int)
    {
        // consume the rebuild request: flags[2] tells the rest of this step's
        // build kernels and the tile pass whether they rebuild
        __shared__ int go;
        if (threadIdx.x == 0) {
            go = a.flags[0] != 0u;
            a.flags[2] = (uint32_t)go;
            if (go) a.flags[0] = 0u;
        }
        __syncthreads();
        if (!go) return;
        __shared__ int32_t wsum[16];
        __shared__ int32_t carry;
        const int t = threadIdx.x, lane = t & 63, w = t >> 6;
        if (t == 0) {
            carry = 0;
            a.stats[0] += 1ull;
            a.stats[1] = 0ull;
            a.stats[2] = 0ull;
            a.stats[3] = 0ull;
            a.stats[4] = 0ull;
            a.stats[6] = 0ull;
        }
        __syncthreads();
        for (int base = 0; base < a.ncell; base += 1024) {
            const int i = base + t;
            const int v = i < a.ncell ? a.count[i] : 0;
            int incl = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int u = __shfl_up(incl, o, 64);
                if (lane >= o) incl += u;
            }
            if (lane == 63) wsum[w] = incl;
            __syncthreads();
            int woff = 0;
            for (int k = 0; k < w; ++k) woff += wsum[k];
            const int c0 = carry;
            if (i < a.ncell) a.start[i] = c0 + woff + incl - v;
            __syncthreads();
            if (t == 1023) carry = c0 + woff + incl;
            __syncthreads();
        }
        if (t == 0) a.start[a.ncell] = carry;
        if (a.ncu > 0) {
            __shared__ int32_t olds[kOrderLds];
            tile_order_build(a.count, a.ncell, a.dims, a.split, a.ncu, a.tile_order, olds);
        }
    }
};

struct Scatter {
    static constexpr int kThreads = kBlock;
    struct Args {
        const int32_t *atom_cell, *atom_rank, *start;
        int32_t* perm;
        int n;
        const uint32_t* flags;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        if (a.flags[2] == 0u) return;
        const int i = blk * kBlock + threadIdx.x;
        if (i >= a.n) return;
        a.perm[a.start[a.atom_cell[i]] + a.atom_rank[i]] = i;
    }
};

// ---------------------------------------------------------------- per-cell sort
// One wave per cell: order the cell's atoms by original id (rank sort in LDS)
// so the layout is deterministic, then write the sorted reference positions
// (also the current sorted positions: a build happens at the current
// positions), charges, types, cells and the inverse permutation.
constexpr int kCellSortMax = 2048;

struct CellSort {
    static constexpr int kThreads = 64;
    struct Args {
        int32_t *perm, *inv_perm;
        const int32_t* start;
        const double *xr, *yr, *zr, *q;
        const int32_t* type;
        double4 *xs_ref, *xs;
        int32_t *scell, *ts;
        const uint32_t* flags;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        if (a.flags[2] == 0u) return;
        __shared__ int32_t buf[kCellSortMax];
        const int c = blk;
        const int lane = threadIdx.x;
        const int s0 = a.start[c], cnt = a.start[c + 1] - s0;
        const bool sortable = cnt <= kCellSortMax;
        if (sortable) {
            for (int k = lane; k < cnt; k += 64) buf[k] = a.perm[s0 + k];
            __syncthreads();
        }
        for (int k = lane; k < cnt; k += 64) {
            int v, r;
            if (sortable) {
                v = buf[k];
                r = 0;
                for (int t = 0; t < cnt; ++t) r += (buf[t] < v) ? 1 : 0;
            } else {
                v = a.perm[s0 + k];
                r = k;
            }
            const int s = s0 + r;
            if (sortable) a.perm[s] = v;
            a.inv_perm[v] = s;
            const double4 p = make_double4(a.xr[v], a.yr[v], a.zr[v], a.q[v]);
            a.xs_ref[s] = p;
            a.xs[s] = p;  // current positions = the reference just built
            a.scell[s] = c;
            a.ts[s] = a.type ? a.type[v] : 0;
        }
    }
};

// MD-step binning before the first step of a run: bins the current positions
// into buffer (epoch & 1) of the MD bins (which the caller cleared), as the tile
// pass of a previous step would have.
// It also opens the run's device state: the long-step flag (flags[4]) cleared and
// the graph's step counter set to the run's first step (no fill or copy launches).
struct MdBin {
    static constexpr int kThreads = kBlock;
    struct Args {
        const double *x, *y, *z;
        int n;
        Box box;
        int3 dims;
        double3 cell;
        MdBins bins;
        uint32_t* flags;
        unsigned long long* step;  // device step counter of graph replays
        unsigned long long step0;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        const int i = blk * kBlock + threadIdx.x;
        if (i == 0) {
            a.flags[4] = 0u;
            *a.step = a.step0;
        }
        if (i >= a.n) return;
        const double p[3] = {a.x[i], a.y[i], a.z[i]};
        // any one-to-one slot numbering serves the cell sort: atom i bins in slot i
        md_bin_atom(a.bins, (int)(a.flags[3] & 1u), i, i, p, a.box, a.dims, a.cell);
    }
};

// Zeroes n ints (a replica's MD bin counts before MdBin, one grid for the batch).
struct ClearInts {
    static constexpr int kThreads = kBlock;
    struct Args {
        int32_t* p;
        int n;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        const int i = blk * kBlock + threadIdx.x;
        if (i < a.n) a.p[i] = 0;
    }
};

// One replica's build statistics and flags into its kStatsOutWords words of the
// batch's readback block, so a batched run ends with one copy to the host.
struct StatsOut {
    static constexpr int kThreads = 64;
    struct Args {
        const unsigned long long* stats;
        const uint32_t* flags;
        unsigned long long* out;
    };
    __device__ static void run(const Args& a, int, int)
    {
        const int t = threadIdx.x;
        if (t < 8)
            a.out[t] = a.stats[t];
        else if (t < 13)
            a.out[t] = a.flags[t - 8];
    }
};

// Chunk counts of MdBin's bins: block k sums the counts of cells [256 k, 256 k + 256)
// of the buffer MdBin filled and writes chunk k (no atomics: atoms in id order land
// in every chunk, so per-atom or per-wave chunk atomics serialise on few counters).
struct ChunkSum {
    static constexpr int kThreads = kBinChunk;
    struct Args {
        MdBins bins;
        const uint32_t* flags;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        __shared__ int red[kThreads / 64];
        const int b = (int)(a.flags[3] & 1u);
        const int c = blk * kBinChunk + (int)threadIdx.x;
        int v = c < a.bins.ncell ? a.bins.count[b * a.bins.ncell + c] : 0;
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int w = 0; w < kThreads / 64; ++w) s += red[w];
            a.bins.chunk[b * a.bins.nchunk + blk] = s;
        }
    }
};

// ---------------------------------------------------------------- Verlet list
// (row builder list_wave and stage_ref live in tile.h, shared with the fused
// list + pair pass)
// List statistics on demand: entries, LJ-range entries, longest row.
__global__ __launch_bounds__(256) void k_list_stats(const int2* __restrict__ ncount, int n,
                                                    unsigned long long* __restrict__ out)
{
    __shared__ unsigned long long se[256], si[256], sm[256];
    unsigned long long e = 0, in = 0, mx = 0;
    for (int s = threadIdx.x; s < n; s += 256) {
        const int2 c = ncount[s];
        e += (unsigned long long)(c.x + c.y);
        in += (unsigned long long)c.x;
        mx = mx > (unsigned long long)(c.x + c.y) ? mx : (unsigned long long)(c.x + c.y);
    }
    se[threadIdx.x] = e;
    si[threadIdx.x] = in;
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            se[threadIdx.x] += se[threadIdx.x + o];
            si[threadIdx.x] += si[threadIdx.x + o];
            sm[threadIdx.x] = sm[threadIdx.x] > sm[threadIdx.x + o] ? sm[threadIdx.x]
                                                                    : sm[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = se[0];
        out[1] = si[0];
        out[2] = sm[0];
    }
}

// ---------------------------------------------------------------- gather
// Current positions into the sorted packed array, unwrapped against the
// reference positions (x_ref + min_image(x - x_ref)), so that the image of
// every list pair stays the one seen at the build.
struct Gather {
    static constexpr int kThreads = kBlock;
    struct Args {
        const double *x, *y, *z, *xr, *yr, *zr, *q;
        const int32_t* inv_perm;
        int n;
        Box box;
        double4* xs;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        const int i = blk * kBlock + threadIdx.x;
        if (i >= a.n) return;
        const double p = a.xr[i], b = a.yr[i], c = a.zr[i];
        const double dx = min_image(a.x[i] - p, a.box.L[0], a.box.invL[0]);
        const double dy = min_image(a.y[i] - b, a.box.L[1], a.box.invL[1]);
        const double dz = min_image(a.z[i] - c, a.box.L[2], a.box.invL[2]);
        a.xs[a.inv_perm[i]] = make_double4(p + dx, b + dy, c + dz, a.q[i]);
    }
};

__global__ __launch_bounds__(kBlock) void k_copy(uint4* __restrict__ dst,
                                                 const uint4* __restrict__ src, int64_t n16)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

inline int nblk(int64_t n, int b = kBlock) { return (int)((n + b - 1) / b); }

// ---------------------------------------------------------------- argument packing
Check::Args check_args(Engine& e, const double* x, const double* y, const double* z)
{
    const double h = 0.5 * e.skin;
    return Check::Args{x, y, z, e.xr, e.yr, e.zr, (int)e.n, e.box, h * h, e.flags,
                       e.cell_count, e.ncell};
}

Bin::Args bin_args(Engine& e, const double* x, const double* y, const double* z)
{
    return Bin::Args{x, y, z, e.xr, e.yr, e.zr, (int)e.n, e.box,
                     make_int3(e.dims[0], e.dims[1], e.dims[2]),
                     make_double3(e.cell[0], e.cell[1], e.cell[2]), e.cell_count, e.atom_cell,
                     e.atom_rank, e.flags};
}

Scan::Args scan_args(Engine& e)
{
    return Scan::Args{e.cell_count,
                      e.cell_start,
                      e.ncell,
                      e.flags,
                      e.stats,
                      make_int3(e.dims[0], e.dims[1], e.dims[2]),
                      e.split,
                      tile_order_ncu(e.ncell, e.split, e.num_cu),
                      e.tile_order};
}

Scatter::Args scatter_args(Engine& e)
{
    return Scatter::Args{e.atom_cell, e.atom_rank, e.cell_start, e.perm, (int)e.n, e.flags};
}

CellSort::Args cellsort_args(Engine& e)
{
    return CellSort::Args{e.perm, e.atom_rank /* reused as inverse perm */, e.cell_start, e.xr,
                          e.yr,   e.zr, e.q, e.type, e.xs_ref, e.xs, e.scell, e.ts, e.flags};
}

Gather::Args gather_args(Engine& e, const double* x, const double* y, const double* z)
{
    return Gather::Args{x, y, z, e.xr, e.yr, e.zr, e.q, e.atom_rank, (int)e.n, e.box, e.xs};
}

MdBin::Args mdbin_args(Engine& e, const double* x, const double* y, const double* z,
                       unsigned long long step0)
{
    return MdBin::Args{x,         y,           z,          (int)e.n, e.box,
                       make_int3(e.dims[0], e.dims[1], e.dims[2]),
                       make_double3(e.cell[0], e.cell[1], e.cell[2]),
                       md_bins(e), e.flags,     e.d_step,   step0};
}

}  // namespace

// ================================================================= launchers

void launch_check(Engine& e, const double* x, const double* y, const double* z, hipStream_t s)
{
    ScopedTimer t(e, P_CHECK, s);
    int64_t cover = e.n > e.ncell ? e.n : e.ncell;
    k_one<Check><<<nblk(cover), kBlock, 0, s>>>(check_args(e, x, y, z));
}

void launch_force_rebuild(Engine& e, hipStream_t s) { k_set_flag<<<1, 1, 0, s>>>(e.flags); }

void launch_build(Engine& e, const double* x, const double* y, const double* z, hipStream_t s)
{
    const int n = (int)e.n;
    {
        ScopedTimer t(e, P_BIN, s);
        k_one<Bin><<<nblk(n), kBlock, 0, s>>>(bin_args(e, x, y, z));
    }
    {
        ScopedTimer t(e, P_SCAN, s);
        k_one<Scan><<<1, 1024, 0, s>>>(scan_args(e));
    }
    {
        ScopedTimer t(e, P_SCATTER, s);
        k_one<Scatter><<<nblk(n), kBlock, 0, s>>>(scatter_args(e));
    }
    {
        ScopedTimer t(e, P_CELLSORT, s);
        k_one<CellSort><<<e.ncell, 64, 0, s>>>(cellsort_args(e));
    }
    // the Verlet rows are built inside the tile pass (pair.hip)
}

void launch_md_bin(Engine& e, const double* x, const double* y, const double* z,
                   unsigned long long step0, hipStream_t s)
{
    // the counts cleared by a kernel of the library (a runtime fill costs the host
    // about twice the launch time)
    const int nc = (int)(md_count_bytes(e.ncell) / sizeof(int32_t));
    k_one<ClearInts><<<nblk(nc), kBlock, 0, s>>>(ClearInts::Args{e.md_count, nc});
    k_one<MdBin><<<nblk(e.n > 0 ? e.n : 1), kBlock, 0, s>>>(mdbin_args(e, x, y, z, step0));
    if (md_nchunk(e.ncell) > 0)
        k_one<ChunkSum><<<md_nchunk(e.ncell), ChunkSum::kThreads, 0, s>>>(
            ChunkSum::Args{md_bins(e), e.flags});
}

void launch_list_stats(Engine& e, unsigned long long* out, hipStream_t s)
{
    k_list_stats<<<1, 256, 0, s>>>(e.ncount, (int)e.n, out);
}

void launch_gather(Engine& e, const double* x, const double* y, const double* z, hipStream_t s)
{
    ScopedTimer t(e, P_GATHER, s);
    const int n = (int)e.n;
    k_one<Gather><<<nblk(n > 0 ? n : 1), kBlock, 0, s>>>(gather_args(e, x, y, z));
}

void launch_copy(void* dst, const void* src, int64_t nbytes, hipStream_t s)
{
    const int64_t n16 = nbytes / 16;
    int blocks = (int)((n16 + kBlock - 1) / kBlock);
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    k_copy<<<blocks, kBlock, 0, s>>>((uint4*)dst, (const uint4*)src, n16);
}

// ================================================================= batched phases

hipError_t nlist_batch_prepare(Batch& b)
{
    const size_t R = b.es.size();
    std::vector<Bin::Args> bi(R);
    std::vector<Scan::Args> sc(R);
    std::vector<Scatter::Args> sp(R);
    std::vector<CellSort::Args> cs(R);
    std::vector<MdBin::Args> mb(R);
    std::vector<ClearInts::Args> mc(R);
    std::vector<StatsOut::Args> so(R);
    std::vector<int> nmc(R);
    std::vector<ChunkSum::Args> cks(R);
    std::vector<Gather::Args> ga(R);
    std::vector<int> nat(R), none(R, 1), ncell(R), nch(R);
    for (size_t r = 0; r < R; ++r) {
        Engine& e = *b.es[r];
        const RepState& st = b.st[r];
        bi[r] = bin_args(e, st.x, st.y, st.z);
        sc[r] = scan_args(e);
        sp[r] = scatter_args(e);
        cs[r] = cellsort_args(e);
        mb[r] = mdbin_args(e, st.x, st.y, st.z, b.step0);
        const int ncount = (int)(md_count_bytes(e.ncell) / sizeof(int32_t));
        mc[r] = ClearInts::Args{e.md_count, ncount};
        nmc[r] = (ncount + kBlock - 1) / kBlock;
        so[r] = StatsOut::Args{e.stats, e.flags, b.d_out ? b.d_out + r * kStatsOutWords : nullptr};
        cks[r] = ChunkSum::Args{md_bins(e), e.flags};
        ga[r] = gather_args(e, st.x, st.y, st.z);
        nat[r] = nblk(e.n > 0 ? e.n : 1);
        ncell[r] = e.ncell;
        nch[r] = md_nchunk(e.ncell);
    }
    hipError_t err;
    if ((err = phase_upload(b.ph[BP_BIN], bi, nat, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_SCAN], sc, none, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_SCATTER], sp, nat, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_CELLSORT], cs, ncell, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_MDCLEAR], mc, nmc, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_MDBIN], mb, nat, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_CHUNKSUM], cks, nch, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_GATHER], ga, nat, 0)) != hipSuccess) return err;
    if (b.d_out && (err = phase_upload(b.ph[BP_STATS_OUT], so, none, 0)) != hipSuccess)
        return err;
    return hipSuccess;
}

void nlist_batch_launch(Batch& b, int phase, hipStream_t s)
{
    switch (phase) {
    case BP_BIN: phase_launch<Bin>(b.ph[phase], s); break;
    case BP_SCAN: phase_launch<Scan>(b.ph[phase], s); break;
    case BP_SCATTER: phase_launch<Scatter>(b.ph[phase], s); break;
    case BP_CELLSORT: phase_launch<CellSort>(b.ph[phase], s); break;
    case BP_MDCLEAR: phase_launch<ClearInts>(b.ph[phase], s); break;
    case BP_MDBIN: phase_launch<MdBin>(b.ph[phase], s); break;
    case BP_STATS_OUT: phase_launch<StatsOut>(b.ph[phase], s); break;
    case BP_CHUNKSUM: phase_launch<ChunkSum>(b.ph[phase], s); break;
    case BP_GATHER: phase_launch<Gather>(b.ph[phase], s); break;
    default: break;
    }
}

}  // namespace rbmd
