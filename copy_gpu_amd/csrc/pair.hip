// pair.hip -- LJ + erfc real-space Coulomb pair pass and the FP64 energy/virial reduction.
//
// Physics (paths relative to the reference checkout):
//   LJ       rbmd/framework/forceFunction/ExecForceFunction.h:74-96 (energy :296-313)
//   Coulomb  ExecForceFunction.h:125-140 with Gnear :217-224 (energy :315-329)
//   per-atom full-list loop  rbmd/framework/run/worklet/RunWorklet.cpp:170-237
//   virial convention        rbmd-gpu src/force/src/ljforce_op/rocm/ljforce_op.hip.cu:555-575
//
// Mapping: one wave64 per cell-sorted atom; the lanes walk the Verlet row 64
// entries at a time (the row's LJ-range entries come first, so only the first
// iterations evaluate LJ), then a wave reduction.  Pair displacements are formed
// in FP64 (minimum image) and the force maths runs in T (float or double).
// Cutoff decisions use the FP64 squared distance.
#include "state.h"

namespace rbmd {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

__device__ inline double min_image(double t, double L, double invL)
{
    return t - L * rint(t * invL);
}

// erfc(z) for z >= 0 in FP32: Chebyshev-fitted form of Numerical Recipes
// (Press et al., 2nd ed., section 6.2, "erfcc"), fractional error < 1.2e-7:
//   t = 1/(1 + z/2),  erfc(z) = t exp(-z^2 + p(t)).
__device__ inline float erfc_nr(float z, float mz2)
{
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.5f * z);
    float p = 0.17087277f;
    p = fmaf(p, t, -0.82215223f);
    p = fmaf(p, t, 1.48851587f);
    p = fmaf(p, t, -1.13520398f);
    p = fmaf(p, t, 0.27886807f);
    p = fmaf(p, t, -0.18628806f);
    p = fmaf(p, t, 0.09678418f);
    p = fmaf(p, t, 0.37409196f);
    p = fmaf(p, t, 1.00002368f);
    p = fmaf(p, t, -1.26551223f);
    return t * __expf(mz2 + p);
}

struct PairScalars {
    double rcl2, rcc2;
    double alpha, sqrt_alpha, tsp, qqr2e;
    double eps24, sig2;  // type 0 - type 0 (ntypes == 1)
    int ntypes;
};

// Scalar force factor fs (force on i = fs * r_ij, r_ij = r_j - r_i) and, with
// EV, the two energies, for one pair at squared distance d2 (FP64 decisions).
template <typename T>
struct PairMath;

template <>
struct PairMath<float> {
    __device__ static inline float eval(double d2, bool lj_range, double qq_d, float eps24,
                                        float sig2, const PairScalars& ps, float* e_lj,
                                        float* e_re, bool ev)
    {
        const float r2 = (float)d2;
        const float rinv = __builtin_amdgcn_rsqf(r2);
        const float r2inv = rinv * rinv;
        float fs = 0.0f;
        if (lj_range && d2 < ps.rcl2) {
            const float sr2 = sig2 * r2inv;
            const float sr6 = sr2 * sr2 * sr2;
            fs = -eps24 * (2.0f * sr6 - 1.0f) * sr6 * r2inv;
            if (ev) *e_lj += 0.5f * (eps24 * (1.0f / 6.0f)) * (sr6 * sr6 - sr6);
        }
        if (d2 < ps.rcc2) {
            const float qq = (float)qq_d;
            const float r = r2 * rinv;
            const float alpha = (float)ps.alpha;
            const float z = (float)ps.sqrt_alpha * r;
            const float mz2 = -alpha * r2;
            const float erfcv = erfc_nr(z, mz2);
            const float g = erfcv * r2inv + (float)ps.tsp * __expf(mz2) * rinv;
            fs -= qq * g * rinv;
            if (ev) *e_re += 0.5f * qq * erfcv * rinv;
        }
        return fs;
    }
};

template <>
struct PairMath<double> {
    __device__ static inline double eval(double d2, bool lj_range, double qq, double eps24,
                                         double sig2, const PairScalars& ps, double* e_lj,
                                         double* e_re, bool ev)
    {
        double fs = 0.0;
        if (lj_range && d2 < ps.rcl2) {
            const double sr2 = sig2 / d2;
            const double sr6 = sr2 * sr2 * sr2;
            fs = -eps24 * (2.0 * sr6 - 1.0) * sr6 / d2;
            if (ev) *e_lj += 0.5 * (eps24 / 6.0) * (sr6 * sr6 - sr6);
        }
        if (d2 < ps.rcc2) {
            const double r = sqrt(d2);
            const double erfcv = erfc(ps.sqrt_alpha * r);
            const double g = erfcv / d2 + ps.tsp * exp(-ps.alpha * d2) / r;
            fs -= qq * g / r;
            if (ev) *e_re += 0.5 * qq * erfcv / r;
        }
        return fs;
    }
};

template <typename T, bool EV>
__device__ inline void accumulate_pair(const double4& pi, const double4& pj, int ti, int tj,
                                       bool lj_range, const Box& box, const PairScalars& ps,
                                       const PairParams* __restrict__ pp, T& fx, T& fy, T& fz,
                                       double* acc)
{
    const double dx = min_image(pj.x - pi.x, box.L[0], box.invL[0]);
    const double dy = min_image(pj.y - pi.y, box.L[1], box.invL[1]);
    const double dz = min_image(pj.z - pi.z, box.L[2], box.invL[2]);
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 > 1e-4)) return;
    T eps24, sig2;
    if (ps.ntypes == 1) {
        eps24 = (T)ps.eps24;
        sig2 = (T)ps.sig2;
    } else {
        eps24 = (T)pp->eps24[ti * kMaxTypes + tj];
        sig2 = (T)pp->sig2[ti * kMaxTypes + tj];
    }
    T e_lj = 0, e_re = 0;
    const double qq = pi.w * pj.w * ps.qqr2e;
    const T fs = PairMath<T>::eval(d2, lj_range, qq, eps24, sig2, ps, &e_lj, &e_re, EV);
    fx += fs * (T)dx;
    fy += fs * (T)dy;
    fz += fs * (T)dz;
    if (EV) {
        const double h = -0.5 * (double)fs;
        acc[0] += (double)e_lj;
        acc[1] += (double)e_re;
        acc[2] += h * dx * dx;
        acc[3] += h * dy * dy;
        acc[4] += h * dz * dz;
        acc[5] += h * dx * dy;
        acc[6] += h * dx * dz;
        acc[7] += h * dy * dz;
    }
}

template <typename T, bool EV>
__global__ __launch_bounds__(kBlock) void k_pair(
    const double4* __restrict__ xs, const int32_t* __restrict__ ts,
    const int32_t* __restrict__ perm, const int32_t* __restrict__ nlist,
    const int2* __restrict__ ncount, const int32_t* __restrict__ scell,
    const int32_t* __restrict__ cell_start, int n, int cap, int3 dims, Box box, PairScalars ps,
    const PairParams* __restrict__ pp, T* __restrict__ fx, T* __restrict__ fy,
    T* __restrict__ fz, double* __restrict__ ev_part)
{
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int s = uniform(blockIdx.x * kWavesPerBlock + wv);
    T ax = 0, ay = 0, az = 0;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (s < n) {
        const double4 pi = xs[s];
        const int ti = ps.ntypes == 1 ? 0 : ts[s];
        const int2 nc = ncount[s];
        const int n_in = uniform(nc.x);
        const int tot = uniform(nc.x + nc.y);
        if (tot <= cap) {
            const int32_t* __restrict__ row = nlist + (size_t)s * cap;
            for (int base = 0; base < tot; base += 64) {
                const int v = base + lane;
                if (v < tot) {
                    const int idx = v < n_in ? v : cap - 1 - (v - n_in);
                    const int j = row[idx];
                    const double4 pj = xs[j];
                    const int tj = ps.ntypes == 1 ? 0 : ts[j];
                    accumulate_pair<T, EV>(pi, pj, ti, tj, base < n_in, box, ps, pp, ax, ay, az,
                                           acc);
                }
            }
        } else {
            // Row overflowed its capacity: scan the 27-cell stencil directly.
            const int c = uniform(scell[s]);
            const int cz = c / (dims.x * dims.y);
            const int cy = (c / dims.x) % dims.y;
            const int cx = c % dims.x;
            for (int st = 0; st < 27; ++st) {
                const int bx = (cx + st / 9 - 1 + dims.x) % dims.x;
                const int by = (cy + (st / 3) % 3 - 1 + dims.y) % dims.y;
                const int bz = (cz + st % 3 - 1 + dims.z) % dims.z;
                const int cc = bx + by * dims.x + bz * dims.x * dims.y;
                const int j0 = cell_start[cc], j1 = cell_start[cc + 1];
                for (int j = j0 + lane; j < j1; j += 64) {
                    const double4 pj = xs[j];
                    const int tj = ps.ntypes == 1 ? 0 : ts[j];
                    accumulate_pair<T, EV>(pi, pj, ti, tj, true, box, ps, pp, ax, ay, az, acc);
                }
            }
        }
        ax = wave_sum(ax);
        ay = wave_sum(ay);
        az = wave_sum(az);
        if (lane == 0) {
            const int i = perm[s];
            fx[i] = ax;
            fy[i] = ay;
            fz[i] = az;
        }
    }
    if (EV) {
        __shared__ double red[kWavesPerBlock][8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double v = wave_sum(acc[k]);
            if (lane == 0) red[wv][k] = v;
        }
        __syncthreads();
        if (threadIdx.x < 8) {
            double v = 0.0;
            for (int w = 0; w < kWavesPerBlock; ++w) v += red[w][threadIdx.x];
            ev_part[(size_t)blockIdx.x * 8 + threadIdx.x] = v;
        }
    }
}

// Deterministic final reduction: pair partials [nb][8] + per-k terms [P][7]
// + self energy -> ev[10] = {e_lj, e_real, e_self, e_k, W_xx..W_yz}.
__global__ __launch_bounds__(256) void k_ev_final(const double* __restrict__ ev_part, int nb,
                                                  const double* __restrict__ kev, int P,
                                                  double e_self, double* __restrict__ ev)
{
    __shared__ double red[256];
    const int t = threadIdx.x;
    double out[10];
    for (int c = 0; c < 15; ++c) {
        double v = 0.0;
        if (c < 8) {
            for (int b = t; b < nb; b += 256) v += ev_part[(size_t)b * 8 + c];
        } else {
            for (int k = t; k < P; k += 256) v += kev[(size_t)k * 7 + (c - 8)];
        }
        red[t] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) red[t] += red[t + o];
            __syncthreads();
        }
        const double r = red[0];
        __syncthreads();
        // c: 0 e_lj, 1 e_real, 2..7 pair virial, 8 e_k, 9..14 k virial
        if (c == 0) out[0] = r;
        else if (c == 1) out[1] = r;
        else if (c < 8) out[4 + (c - 2)] = r;
        else if (c == 8) out[3] = r;
        else out[4 + (c - 9)] += r;
    }
    if (t == 0) {
        out[2] = e_self;
        for (int k = 0; k < 10; ++k) ev[k] = out[k];
    }
}

}  // namespace

void launch_pair(Engine& e, void* fx, void* fy, void* fz, bool ev, hipStream_t s)
{
    const int n = (int)e.n;
    if (n == 0) return;
    PairScalars ps;
    ps.rcl2 = e.pp.rcl2;
    ps.rcc2 = e.pp.rcc2;
    ps.alpha = e.pp.alpha;
    ps.sqrt_alpha = e.pp.sqrt_alpha;
    ps.tsp = e.pp.two_sqrt_alpha_over_pi;
    ps.qqr2e = e.pp.qqr2e;
    ps.eps24 = e.pp.eps24[0];
    ps.sig2 = e.pp.sig2[0];
    ps.ntypes = e.pp.ntypes;
    const int3 dims = make_int3(e.dims[0], e.dims[1], e.dims[2]);
    const int blocks = (n + kWavesPerBlock - 1) / kWavesPerBlock;
    ScopedTimer t(e, P_PAIR, s);
#define RBMD_PAIR_ARGS                                                                        \
    e.xs, e.ts, e.perm, e.nlist, e.ncount, e.scell, e.cell_start, n, e.cap, dims, e.box, ps,  \
        e.d_pp
    if (e.precision == 32) {
        if (ev)
            k_pair<float, true><<<blocks, kBlock, 0, s>>>(RBMD_PAIR_ARGS, (float*)fx, (float*)fy,
                                                           (float*)fz, e.ev_part);
        else
            k_pair<float, false><<<blocks, kBlock, 0, s>>>(RBMD_PAIR_ARGS, (float*)fx,
                                                            (float*)fy, (float*)fz, e.ev_part);
    } else {
        if (ev)
            k_pair<double, true><<<blocks, kBlock, 0, s>>>(RBMD_PAIR_ARGS, (double*)fx,
                                                            (double*)fy, (double*)fz, e.ev_part);
        else
            k_pair<double, false><<<blocks, kBlock, 0, s>>>(RBMD_PAIR_ARGS, (double*)fx,
                                                             (double*)fy, (double*)fz, e.ev_part);
    }
#undef RBMD_PAIR_ARGS
}

void launch_ev_final(Engine& e, bool with_pair, hipStream_t s)
{
    ScopedTimer t(e, P_EV_FINAL, s);
    const int nb = with_pair ? (int)((e.n + kWavesPerBlock - 1) / kWavesPerBlock) : 0;
    const int P = (e.kspace != 0) ? e.P : 0;
    k_ev_final<<<1, 256, 0, s>>>(e.ev_part, nb, e.kev, P, e.e_self, e.ev);
}

}  // namespace rbmd
