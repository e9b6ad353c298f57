// md.h -- kernel bodies of the velocity-Verlet MD step.
//
// An MD run issues, per step s, three kernels (plus one opening kernel per run
// segment):
//   KickDrift      (segment start) v += dt/2 F/m, x += v dt, wrap + image flags,
//                  displacement check, gather into the cell-sorted array, and the
//                  k-vector draw of step s
//   SortRho        rho(k) partials of step s, and, when the check asked for a list
//                  rebuild, the cell sort of the bins made one step earlier
//   TileRho        (pair.hip) rho(k) final reduction + the tile pass (Verlet rows
//                  on rebuild steps, pair forces, binning for the next rebuild)
//   KForceDrift    k-space force, closing half kick of step s, and the opening
//                  kick/drift/check/gather + k-vector draw of step s+1
//   or KForce<T, true> at the end of a segment (closing half kick only).
// Every atom's closing kick and next opening kick/drift run in the same lane with
// the same operations as the split kernels, so a run gives the same bits however
// it is cut into segments.
//
// Reference behaviour followed (paths relative to the reference checkout):
//   kick / drift + image flags    rbmd/framework/run/worklet/RunWorklet.cpp:1764-1782, 1854-1877
//   cell sort                     rbmd/framework/locator/ContPointLocator.cxx:38-92
#pragma once
#include "kspace.h"
#include "state.h"
#include "tile.h"

namespace rbmd {

constexpr int kMdBlock = 256;      // KickDrift workgroup (one atom per thread)
constexpr int kSamplerLds = 2048;  // CDF entries staged in LDS by the fused samplers
constexpr int kSortWaveBuf = 1024;  // bucket ids ranked in LDS per cell-sort wave

// ---------------------------------------------------------------- k-vector draw
// The first ceil(P / kThreads) workgroups of the kick/drift kernels only draw
// k-vectors (l = blk * blockDim.x + thread), so the draws run beside the atom
// work instead of ahead of it; the CDF tables are staged in LDS (cdf,
// kSamplerLds entries) when they fit.  Call from every thread of the group.
__device__ inline int sampler_blocks(const SampleArgs& sa, int threads)
{
    return (sa.P + threads - 1) / threads;
}

__device__ inline void sample_block(const SampleArgs& sa, int blk, double* cdf)
{
    const int len = 2 * (sa.M.x + sa.M.y + sa.M.z) + 3;
    const bool in_lds = len <= kSamplerLds;
    if (in_lds)
        for (int k = threadIdx.x; k < len; k += blockDim.x) cdf[k] = sa.cdf[k];
    __syncthreads();
    const int l = blk * (int)blockDim.x + (int)threadIdx.x;
    if (l < sa.P) {
        const uint64_t step = sa.step_dev ? (uint64_t)(*sa.step_dev) : sa.step;
        const int Ms[3] = {sa.M.x, sa.M.y, sa.M.z};
        int m[3];
        const double w = draw_kvector(sa.seed, step, (uint32_t)l, in_lds ? cdf : sa.cdf, Ms,
                                      sa.gp, sa.P, sa.w, sa.w_out, m);
        sa.km[3 * l] = m[0];
        sa.km[3 * l + 1] = m[1];
        sa.km[3 * l + 2] = m[2];
        sa.kw[l] = w;
    }
}

// ---------------------------------------------------------------- kick / drift
template <typename T>
struct KickDriftArgs {
    double *x, *y, *z, *vx, *vy, *vz;
    const T *fx, *fy, *fz;
    int32_t* img;
    const double* mass;
    double dt;
    int n;
    Box box;
    const double *xr, *yr, *zr;
    double thr2;
    uint32_t* flags;
    const double* q;
    const int32_t* inv_perm;
    double4* xs;
    SampleArgs sa;
    const double* vscale = nullptr;  // thermostat velocity scale (v = c v before the kick)
    ThermoArgs th;  // set: reduce the step before's temperature here (no ThermoReduce)
};

// Per-atom inputs of the kick/drift besides velocity and force.
struct KdIn {
    double x[3], xr[3], m, q;
    int32_t slot, img[3];
};

template <typename T>
__device__ inline KdIn kd_load(const KickDriftArgs<T>& a, int i)
{
    KdIn in;
    in.x[0] = a.x[i];
    in.x[1] = a.y[i];
    in.x[2] = a.z[i];
    in.xr[0] = a.xr[i];
    in.xr[1] = a.yr[i];
    in.xr[2] = a.zr[i];
    in.m = a.mass ? a.mass[i] : 1.0;
    in.q = a.q[i];
    in.slot = a.inv_perm[i];
    in.img[0] = in.img[1] = in.img[2] = 0;
    if (a.img) {
        in.img[0] = a.img[3 * i];
        in.img[1] = a.img[3 * i + 1];
        in.img[2] = a.img[3 * i + 2];
    }
    return in;
}

// Opening half kick v += dt/2 F/m from velocity v0 and force f, drift
// x += v dt, wrap into the box with image-flag update, displacement check
// against the reference positions of the last build, and the gather of the new
// position into the cell-sorted array, unwrapped against the reference
// (x_ref + min_image(x - x_ref)) so that every list pair keeps its image.
template <typename T>
__device__ inline void kick_drift_atom(const KickDriftArgs<T>& a, int i, const KdIn& in, double v0,
                                       double v1, double v2, T f0, T f1, T f2)
{
#pragma clang fp contract(off)
    const Box& box = a.box;
    const double dt = a.dt;
    const double h = 0.5 * dt / in.m;
    double v[3] = {v0 + h * (double)f0, v1 + h * (double)f1, v2 + h * (double)f2};
    double p[3] = {in.x[0] + v[0] * dt, in.x[1] + v[1] * dt, in.x[2] + v[2] * dt};
    int fl[3];
    for (int d = 0; d < 3; ++d) {
        const double f = floor((p[d] - box.lo[d]) * box.invL[d]);
        p[d] = p[d] - f * box.L[d];
        fl[d] = (int)f;
    }
    a.vx[i] = v[0];
    a.vy[i] = v[1];
    a.vz[i] = v[2];
    a.x[i] = p[0];
    a.y[i] = p[1];
    a.z[i] = p[2];
    if (a.img) {
        a.img[3 * i] = in.img[0] + fl[0];
        a.img[3 * i + 1] = in.img[1] + fl[1];
        a.img[3 * i + 2] = in.img[2] + fl[2];
    }
    const double dx = min_image(p[0] - in.xr[0], box.L[0], box.invL[0]);
    const double dy = min_image(p[1] - in.xr[1], box.L[1], box.invL[1]);
    const double dz = min_image(p[2] - in.xr[2], box.L[2], box.invL[2]);
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > a.thr2) a.flags[0] = 1u;
    if (d2 != d2) a.flags[1] = 1u;
    // A rebuild takes its reference positions from the binning one step earlier, so
    // the list stays exact only while no atom moves more than half the skin in one
    // step (flags[4] reports a step that did; the run then returns an error).
    const double s0 = v[0] * dt, s1 = v[1] * dt, s2 = v[2] * dt;
    if (s0 * s0 + s1 * s1 + s2 * s2 > a.thr2) a.flags[4] = 1u;
    a.xs[in.slot] = make_double4(in.xr[0] + dx, in.xr[1] + dy, in.xr[2] + dz, in.q);
}

// Opening kernel of a segment: the step's k-vector draw (first workgroups), then
// the kick/drift of every atom (one per thread).
template <typename T>
struct KickDrift {
    static constexpr int kThreads = kMdBlock;
    using Args = KickDriftArgs<T>;
    __device__ static void run(const Args& a, int blk, int)
    {
        const int ns = sampler_blocks(a.sa, kMdBlock);
        if (blk < ns) {
            __shared__ double cdf[kSamplerLds];
            sample_block(a.sa, blk, cdf);
            return;
        }
        double c = 1.0;
        if (a.th.type != kThermoNone) {
            // the step before's temperature, summed by every workgroup in the same
            // order as ThermoReduce; the first atom workgroup also logs it and
            // updates the thermostat state (which no other workgroup reads here)
            const double s = thermo_sum<kMdBlock>(a.th);
            if (blk == ns && threadIdx.x == 0) thermo_finish(a.th, s, a.dt);
            c = thermo_scale(a.th, s);
        } else if (a.vscale) {
            c = *a.vscale;
        }
        const int i = (blk - ns) * kMdBlock + threadIdx.x;
        if (i >= a.n) return;
        if (a.vscale || a.th.type != kThermoNone) {
            // Berendsen / rescale coefficient of the step before
            // (UpdateVelocityRescaleWorklet, rbmd/framework/run/worklet/RunWorklet.cpp:1818-1830)
            kick_drift_atom(a, i, kd_load(a, i), c * a.vx[i], c * a.vy[i], c * a.vz[i], a.fx[i],
                            a.fy[i], a.fz[i]);
            return;
        }
        kick_drift_atom(a, i, kd_load(a, i), a.vx[i], a.vy[i], a.vz[i], a.fx[i], a.fy[i],
                        a.fz[i]);
    }
};

// Start of a thermostatted run: scale 1 for the first opening kick, temperature
// log from position 0.  One thread.
struct ThermoBegin {
    static constexpr int kThreads = 64;
    struct Args {
        double* st;
        uint32_t* cnt;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        if (blk != 0 || threadIdx.x != 0) return;
        a.st[0] = 1.0;
        a.cnt[1] = 0u;
    }
};

// End of a thermostatted run: v = c v with the scale the last step's closing kick
// formed (the opening kick of a following step would otherwise apply it).
struct VScale {
    static constexpr int kThreads = kMdBlock;
    struct Args {
        double *vx, *vy, *vz;
        const double* scale;
        int n;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        const int i = blk * kMdBlock + threadIdx.x;
        if (i >= a.n) return;
        const double c = *a.scale;
        a.vx[i] = c * a.vx[i];
        a.vy[i] = c * a.vy[i];
        a.vz[i] = c * a.vz[i];
    }
};

// Closing half kick of step s fused with the opening kick/drift of step s+1
// (and its k-vector draw): the lane that completes atom i's force carries its
// velocity through both half kicks; the atom's other inputs are loaded before
// the k-space sum.
template <typename T>
struct KForceDrift {
    static constexpr int kThreads = kKfThreads;
    struct Args {
        KForceArgs<T> kf;
        KickDriftArgs<T> kd;
    };
    struct In {
        double v[3];
        KdIn kd;
    };
    __device__ static void run(const Args& A, int blk, int)
    {
#pragma clang fp contract(off)
        __shared__ union {
            KForceLds<T> kf;
            double cdf[kSamplerLds];
        } lds;
        if (blk == 0 && threadIdx.x == 0) A.kf.flags[3] += 1u;  // MD step epoch
        const int ns = sampler_blocks(A.kd.sa, kKfThreads);
        if (blk < ns) {
            sample_block(A.kd.sa, blk, lds.cdf);
            return;
        }
        const KForceArgs<T>& kf = A.kf;
        kforce_block(
            kf, blk - ns, lds.kf,
            [&](int i) {
                In in;
                in.v[0] = kf.vx[i];
                in.v[1] = kf.vy[i];
                in.v[2] = kf.vz[i];
                in.kd = kd_load(A.kd, i);
                return in;
            },
            [&](int i, T f0, T f1, T f2, const In& in) {
#pragma clang fp contract(off)
                kf.fx[i] = f0;
                kf.fy[i] = f1;
                kf.fz[i] = f2;
                const double h = 0.5 * kf.dt / in.kd.m;
                const double v0 = in.v[0] + h * (double)f0;
                const double v1 = in.v[1] + h * (double)f1;
                const double v2 = in.v[2] + h * (double)f2;
                kick_drift_atom(A.kd, i, in.kd, v0, v1, v2, f0, f1, f2);
            });
    }
};

// ---------------------------------------------------------------- MD cell sort
// Rebuild-step cell sort from the bins the tile pass made one step earlier
// (buffer flags[3] & 1): one wave per cell.  The cell's start is the sum of the
// counts before it; its atoms come from the bucket (arrival order) ranked by id,
// or, for a bucket that overflowed, from a scan of every binning slot's cell.  The
// sorted slot s of atom v receives perm, inv_perm, the binned position as the
// new reference, the packed reference and current (unwrapped) positions, its
// cell and type.  The wave of cell 0 publishes the rebuild decision (flags[2])
// and resets the build statistics.
struct CellSortMDArgs {
    const double *x, *y, *z, *q;
    const int32_t* type;
    int n, ncell;
    Box box;
    MdBins bins;
    int32_t *cell_start, *perm, *inv_perm, *scell, *ts;
    double *xr, *yr, *zr;
    double4 *xs_ref, *xs;
    uint32_t* flags;
    unsigned long long* stats;
};

// Atom v, binned in slot so, takes sorted slot s.
__device__ inline void cellsort_md_emit(const CellSortMDArgs& a, int c, int v, int so, int s)
{
    a.perm[s] = v;
    a.inv_perm[v] = s;
    const double w0 = a.bins.xw[so], w1 = a.bins.yw[so], w2 = a.bins.zw[so];
    a.xr[v] = w0;
    a.yr[v] = w1;
    a.zr[v] = w2;
    const double qv = a.q[v];
    a.xs_ref[s] = make_double4(w0, w1, w2, qv);
    a.xs[s] = make_double4(w0 + min_image(a.x[v] - w0, a.box.L[0], a.box.invL[0]),
                           w1 + min_image(a.y[v] - w1, a.box.L[1], a.box.invL[1]),
                           w2 + min_image(a.z[v] - w2, a.box.L[2], a.box.invL[2]), qv);
    a.scell[s] = c;
    a.ts[s] = a.type ? a.type[v] : 0;
}

// One wave sorts cell c (c >= ncell: nothing); buf holds kSortWaveBuf ids of
// this wave's own LDS.  The bucket lists binning slots; atoms are ranked by id.
__device__ inline void cellsort_md_cell(const CellSortMDArgs& a, int c, int lane, int32_t* buf)
{
    const uint32_t req = a.flags[0];
    if (c == 0 && lane == 0) {
        a.flags[2] = req;
        if (req) {
            a.stats[0] += 1ull;
            a.stats[1] = 0ull;
            a.stats[2] = 0ull;
            a.stats[3] = 0ull;
            a.stats[4] = 0ull;
            a.stats[6] = 0ull;
        }
    }
    if (req == 0u || c >= a.ncell) return;
    const int b = (int)(a.flags[3] & 1u);
    const int32_t* count = a.bins.count + (size_t)b * a.ncell;
    const int32_t* chunk = a.bins.chunk + (size_t)b * a.bins.nchunk;
    // start = the counts of the whole chunks before c, then of the cells of c's own
    // chunk before c: at most ncell / kBinChunk + kBinChunk loads per cell (without
    // chunk counts, every earlier cell)
    const int c0 = a.bins.nchunk > 0 ? (c / kBinChunk) * kBinChunk : 0;
    int part = 0;
    if (a.bins.nchunk > 0)
        for (int k = lane; k < c / kBinChunk; k += 64) part += chunk[k];
    for (int k = c0 + lane; k < c; k += 64) part += count[k];
    const int s0 = wave_sum(part);
    const int cnt = count[c];
    if (lane == 0) {
        a.cell_start[c] = s0;
        if (c == a.ncell - 1) a.cell_start[a.ncell] = s0 + cnt;
    }
    if (cnt <= a.bins.cap) {
        const int32_t* bk = a.bins.bucket + ((size_t)b * a.ncell + c) * a.bins.cap;
        if (cnt <= kSortWaveBuf) {
            // ids staged in LDS, each lane ranks its own entries
            for (int k = lane; k < cnt; k += 64) buf[k] = a.bins.vid[bk[k]];
            __syncwarp();
            for (int k = lane; k < cnt; k += 64) {
                const int v = buf[k];
                int r = 0;
                for (int t = 0; t < cnt; ++t) r += (buf[t] < v) ? 1 : 0;
                cellsort_md_emit(a, c, v, bk[k], s0 + r);
            }
        } else {
            for (int k = lane; k < cnt; k += 64) {
                const int v = a.bins.vid[bk[k]];
                int r = 0;
                for (int t = 0; t < cnt; ++t) r += (a.bins.vid[bk[t]] < v) ? 1 : 0;
                cellsort_md_emit(a, c, v, bk[k], s0 + r);
            }
        }
    } else {
        // the bucket overflowed: scan every binning slot (order: by slot)
        int found = 0;
        for (int base = 0; base < a.n; base += 64) {
            const int i = base + lane;
            const bool hit = i < a.n && a.bins.cell[i] == c;
            const uint64_t m = __ballot(hit);
            if (hit) cellsort_md_emit(a, c, a.bins.vid[i], i, s0 + found + (int)mbcnt64(m));
            found += __popcll(m);
        }
    }
}

// rho(k) partials (workgroups [0, nrho)) and the MD cell sort (the rest, one
// cell per wave) in one grid: both read only the step's drifted positions.
template <typename T>
struct SortRho {
    static constexpr int kThreads = kRhoBlock;
    static constexpr int kWaves = kThreads / 64;
    static constexpr size_t kSortBytes = kWaves * kSortWaveBuf * sizeof(int32_t);
    static constexpr size_t kLdsBytes =
        RhoPartial<T>::kLdsBytes > kSortBytes ? RhoPartial<T>::kLdsBytes : kSortBytes;
    static_assert(kLdsBytes >= kOrderLds * sizeof(int32_t),
                  "the tile-order workgroup shares the stage");
    struct Args {
        typename RhoPartial<T>::Args rho;
        int nrho;
        CellSortMDArgs cs;
        int nsort;  // cell-sort workgroups; the one after them orders the tiles
        int3 dims;
        int split, ncu;  // ncu 0: keep the cell order
        int32_t* tile_order;
    };
    __device__ static void run(const Args& A, int blk, int)
    {
        __shared__ double4 lds[kLdsBytes / sizeof(double4)];
        if (blk < A.nrho) {
            RhoPartial<T>::run_lds(A.rho, blk, lds);
            return;
        }
        blk -= A.nrho;
        if (blk == A.nsort) {
            // rebuild steps: order the tile workgroups by the new cells' work
            if (A.ncu > 0 && A.cs.flags[0] != 0u) {
                const int32_t* count =
                    A.cs.bins.count + (size_t)(A.cs.flags[3] & 1u) * A.cs.ncell;
                tile_order_build(count, A.cs.ncell, A.dims, A.split, A.ncu, A.tile_order,
                                 (int32_t*)lds);
            }
            return;
        }
        const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
        cellsort_md_cell(A.cs, blk * kWaves + wv, lane, (int32_t*)lds + wv * kSortWaveBuf);
    }
};

inline int sortrho_cell_blocks(int ncell) { return (ncell + kRhoBlock / 64 - 1) / (kRhoBlock / 64); }

}  // namespace rbmd
