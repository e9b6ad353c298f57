// kspace.hip -- Random Batch Ewald (and full Ewald) k-space: device k-sampler,
// structure factor rho(k) = sum_j q_j exp(i k.r_j), per-k energy/virial terms,
// and the per-atom k-space force.
//
// Reference (paths relative to the reference checkout):
//   sampler target / normaliser S   rbmd/framework/model/RBEPSample.h:38-61, 87-93, 143-175
//   rho(k)                          rbmd/framework/run/worklet/RunWorklet.cpp:1616-1677,
//                                   rbmd/framework/run/ExecutionMD.cpp:412-448
//   RBE force                       RunWorklet.cpp:1508-1548, ExecForceFunction.h:247-269
//   Ewald force / energy            ExecForceFunction.h:226-245, output/ThermoOutput.cpp:209-255
//
// Phases are formed from FP64 fractional coordinates, phi = frac(m . s) with
// s = (x - lo)/L, so FP32 sines keep full accuracy at any |k.r|; sin/cos of the
// turn fraction come from the hardware sine/cosine in FP32 and sincospi in FP64.  FP32 mode accumulates rho with Kahan compensation per lane
// and reduces the per-block partials in FP64, deterministically.
#include "state.h"

namespace rbmd {

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- sampler
__global__ __launch_bounds__(256) void k_sample(SampleArgs sa)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= sa.P) return;
    const uint64_t step = sa.step_dev ? (uint64_t)(*sa.step_dev) : sa.step;
    const int Ms[3] = {sa.M.x, sa.M.y, sa.M.z};
    int m[3];
    const double w =
        draw_kvector(sa.seed, step, (uint32_t)l, sa.cdf, Ms, sa.gp, sa.P, sa.w, sa.w_out, m);
    sa.km[3 * l] = m[0];
    sa.km[3 * l + 1] = m[1];
    sa.km[3 * l + 2] = m[2];
    sa.kw[l] = w;
}

__global__ void k_fill(double* __restrict__ a, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------- trig helpers
template <typename T>
__device__ inline void sincos2pi(double phi, T* s, T* c);

// FP32: the hardware sine/cosine take their argument in turns (sin(2 pi x));
// on [-0.5, 0.5] turns their error is 1.3e-7 absolute (measured against FP64),
// the size of the FP32 rounding of the turn fraction itself.
template <>
__device__ inline void sincos2pi<float>(double phi, float* s, float* c)
{
    const float f = (float)(phi - rint(phi));  // [-0.5, 0.5] turns
    *s = __builtin_amdgcn_sinf(f);
    *c = __builtin_amdgcn_cosf(f);
}

template <>
__device__ inline void sincos2pi<double>(double phi, double* s, double* c)
{
    const double f = phi - rint(phi);
    sincospi(2.0 * f, s, c);
}

// Kahan accumulator in T (used for float; a no-op compensation in double).
template <typename T>
struct Kahan {
    T s = 0, c = 0;
    __device__ inline void add(T v)
    {
        const T y = v - c;
        const T t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ inline double value() const { return (double)s - (double)c; }
};

template <>
struct Kahan<double> {
    double s = 0;
    __device__ inline void add(double v) { s += v; }
    __device__ inline double value() const { return s; }
};

// ---------------------------------------------------------------- rho partials
// grid = (atom tiles, k tiles of 64).  Lane <-> k within the tile, the 4 waves
// split the tile's atoms; atoms are staged through LDS 256 at a time.
constexpr int kRhoBlock = 256;
constexpr int kRhoStage = 256;

template <typename T>
__global__ __launch_bounds__(kRhoBlock) void k_rho_partial(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ q, int n, int atoms_per_block, Box box,
    const int32_t* __restrict__ km, int P, double* __restrict__ kpart)
{
#pragma clang fp contract(off)
    __shared__ double4 st[kRhoStage];
    __shared__ double red[kRhoBlock / 64][64][2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kidx = blockIdx.y * 64 + lane;
    const bool kok = kidx < P;
    double mx = 0, my = 0, mz = 0;
    if (kok) {
        mx = (double)km[3 * kidx];
        my = (double)km[3 * kidx + 1];
        mz = (double)km[3 * kidx + 2];
    }
    Kahan<T> re, im;
    const int a0 = blockIdx.x * atoms_per_block;
    int a1 = a0 + atoms_per_block;
    a1 = a1 < n ? a1 : n;
    for (int base = a0; base < a1; base += kRhoStage) {
        __syncthreads();
        const int i = base + threadIdx.x;
        if (threadIdx.x < kRhoStage) {
            if (i < a1)
                st[threadIdx.x] = make_double4((x[i] - box.lo[0]) * box.invL[0],
                                               (y[i] - box.lo[1]) * box.invL[1],
                                               (z[i] - box.lo[2]) * box.invL[2], q[i]);
            else
                st[threadIdx.x] = make_double4(0, 0, 0, 0);
        }
        __syncthreads();
        const int cnt = (a1 - base) < kRhoStage ? (a1 - base) : kRhoStage;
        if (kok) {
            for (int a = wv; a < cnt; a += kRhoBlock / 64) {
                const double4 p = st[a];
                const double phi = fma(mz, p.z, fma(my, p.y, mx * p.x));
                T sn, cs;
                sincos2pi<T>(phi, &sn, &cs);
                const T qa = (T)p.w;
                re.add(qa * cs);
                im.add(qa * sn);
            }
        }
    }
    red[wv][lane][0] = re.value();
    red[wv][lane][1] = im.value();
    __syncthreads();
    if (wv == 0 && kok) {
        double r = 0.0, m = 0.0;
        for (int w = 0; w < kRhoBlock / 64; ++w) {
            r += red[w][lane][0];
            m += red[w][lane][1];
        }
        const size_t o = ((size_t)blockIdx.x * P + kidx) * 2;
        kpart[o] = r;
        kpart[o + 1] = m;
    }
}

// ---------------------------------------------------------------- rho final
// One block per k: FP64 tree over the atom-tile partials, then the force
// coefficients and the energy / virial terms of this k.
template <typename T>
__global__ __launch_bounds__(256) void k_rho_final(
    const double* __restrict__ kpart, int nbx, int P, const int32_t* __restrict__ km,
    const double* __restrict__ kw, Box box, double alpha, double qqr2e,
    double* __restrict__ rho, KCoef<T>* __restrict__ coef, double* __restrict__ kev,
    unsigned long long* __restrict__ step_dev)
{
    __shared__ double sr[256], si[256];
    const int k = blockIdx.x, t = threadIdx.x;
    double r = 0.0, m = 0.0;
    for (int b = t; b < nbx; b += 256) {
        r += kpart[((size_t)b * P + k) * 2];
        m += kpart[((size_t)b * P + k) * 2 + 1];
    }
    sr[t] = r;
    si[t] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            sr[t] += sr[t + o];
            si[t] += si[t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double re = sr[0], im = si[0];
        rho[2 * k] = re;
        rho[2 * k + 1] = im;
        const double mm[3] = {(double)km[3 * k], (double)km[3 * k + 1], (double)km[3 * k + 2]};
        double kv[3];
        for (int d = 0; d < 3; ++d) kv[d] = 2.0 * kPi * mm[d] / box.L[d];
        const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        const double V = box.L[0] * box.L[1] * box.L[2];
        const double w = kw[k];
        const double c = w * (-4.0 * kPi) / (V * k2) * qqr2e;
        KCoef<T> kc;
        for (int d = 0; d < 3; ++d) {
            kc.m[d] = mm[d];
            kc.k[d] = (T)kv[d];
        }
        kc.A = (T)(c * im);
        kc.B = (T)(c * re);
        coef[k] = kc;
        const double e = w * (2.0 * kPi / V) * (re * re + im * im) / k2 * qqr2e;
        const double b = 2.0 * (1.0 / k2 + 1.0 / (4.0 * alpha));
        double* o = kev + (size_t)k * 7;
        o[0] = e;
        o[1] = e * (1.0 - b * kv[0] * kv[0]);
        o[2] = e * (1.0 - b * kv[1] * kv[1]);
        o[3] = e * (1.0 - b * kv[2] * kv[2]);
        o[4] = -e * b * kv[0] * kv[1];
        o[5] = -e * b * kv[0] * kv[2];
        o[6] = -e * b * kv[1] * kv[2];
        if (k == 0 && step_dev) *step_dev += 1ull;
    }
}

// ---------------------------------------------------------------- k-space force
// Block = 64 atoms (lanes, original order) x 8 k-slices (waves): wave w takes
// k = w, w+8, ... of the coefficients staged in LDS (every lane reads the same
// k: broadcast), then the 8 partial forces of each atom are summed in a fixed
// order.  F_i += q_i sum_k k (cos(phi) A_k - sin(phi) B_k).  With KICK the
// final half kick v += dt/2 F/m of the velocity-Verlet step is fused in.
constexpr int kKfAtoms = 64;
constexpr int kKfSlices = 8;
constexpr int kKfThreads = kKfAtoms * kKfSlices;
constexpr int kKfStage = 512;

template <typename T, bool KICK>
__global__ __launch_bounds__(kKfThreads) void k_kforce(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ q, int n, Box box, const KCoef<T>* __restrict__ coef, int P,
    T* __restrict__ fx, T* __restrict__ fy, T* __restrict__ fz, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz, const double* __restrict__ mass, double dt)
{
#pragma clang fp contract(off)
    __shared__ KCoef<T> st[kKfStage];
    __shared__ T red[kKfSlices][3][kKfAtoms];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * kKfAtoms + lane;
    double sx = 0, sy = 0, sz = 0;
    if (i < n) {
        sx = (x[i] - box.lo[0]) * box.invL[0];
        sy = (y[i] - box.lo[1]) * box.invL[1];
        sz = (z[i] - box.lo[2]) * box.invL[2];
    }
    T ax = 0, ay = 0, az = 0;
    for (int base = 0; base < P; base += kKfStage) {
        __syncthreads();
        for (int k = threadIdx.x; k < kKfStage && base + k < P; k += kKfThreads)
            st[k] = coef[base + k];
        __syncthreads();
        const int cnt = (P - base) < kKfStage ? (P - base) : kKfStage;
        for (int k = wv; k < cnt; k += kKfSlices) {
            const KCoef<T>& c = st[k];
            const double phi = fma(c.m[2], sz, fma(c.m[1], sy, c.m[0] * sx));
            T sn, cs;
            sincos2pi<T>(phi, &sn, &cs);
            const T g = cs * c.A - sn * c.B;
            ax += g * c.k[0];
            ay += g * c.k[1];
            az += g * c.k[2];
        }
    }
    red[wv][0][lane] = ax;
    red[wv][1][lane] = ay;
    red[wv][2][lane] = az;
    __syncthreads();
    if (wv == 0 && i < n) {
        T s0 = 0, s1 = 0, s2 = 0;
        for (int w = 0; w < kKfSlices; ++w) {
            s0 += red[w][0][lane];
            s1 += red[w][1][lane];
            s2 += red[w][2][lane];
        }
        const T qi = (T)q[i];
        const T f0 = fx[i] + qi * s0, f1 = fy[i] + qi * s1, f2 = fz[i] + qi * s2;
        fx[i] = f0;
        fy[i] = f1;
        fz[i] = f2;
        if (KICK) {
            const double h = 0.5 * dt / (mass ? mass[i] : 1.0);
            vx[i] += h * (double)f0;
            vy[i] += h * (double)f1;
            vz[i] += h * (double)f2;
        }
    }
}

}  // namespace

// ================================================================= launchers

SampleArgs sample_args(const Engine& e, int P, uint64_t seed, const uint64_t* step_host)
{
    SampleArgs sa;
    sa.P = P;
    sa.seed = seed;
    sa.step = step_host ? *step_host : 0ull;
    sa.step_dev = step_host ? nullptr : e.d_step;
    sa.cdf = e.kcdf;
    sa.M = make_int3(e.kM[0], e.kM[1], e.kM[2]);
    sa.w = e.S / (double)P;
    sa.gp = e.gp;
    sa.w_out = e.gp.cdf ? e.S / ((1.0 - e.gp.lambda) * (double)P) : 0.0;
    sa.km = e.km;
    sa.kw = e.kw;
    return sa;
}

void launch_sample(Engine& e, int P, uint64_t seed, const uint64_t* step_host, hipStream_t s)
{
    ScopedTimer t(e, P_SAMPLE, s);
    k_sample<<<(P + 255) / 256, 256, 0, s>>>(sample_args(e, P, seed, step_host));
}

void launch_kweights_default(Engine& e, hipStream_t s)
{
    k_fill<<<(e.P + 255) / 256, 256, 0, s>>>(e.kw, e.P, e.S / (double)e.P);
}

static void rho_geometry(const Engine& e, int* atoms_per_block, int* nbx, int* nkt)
{
    const int n = (int)e.n;
    *nkt = (e.P + 63) / 64;
    int want = 2048 / *nkt;
    if (want < 1) want = 1;
    int apb = (n + want - 1) / want;
    apb = ((apb + 63) / 64) * 64;
    if (apb < 64) apb = 64;
    *atoms_per_block = apb;
    *nbx = (n + apb - 1) / apb;
    if (*nbx < 1) *nbx = 1;
}

void launch_rho(Engine& e, const double* x, const double* y, const double* z, hipStream_t s)
{
    int apb, nbx, nkt;
    rho_geometry(e, &apb, &nbx, &nkt);
    {
        ScopedTimer t(e, P_RHO, s);
        dim3 grid(nbx, nkt);
        if (e.precision == 32)
            k_rho_partial<float><<<grid, kRhoBlock, 0, s>>>(x, y, z, e.q, (int)e.n, apb, e.box,
                                                             e.km, e.P, e.kpart);
        else
            k_rho_partial<double><<<grid, kRhoBlock, 0, s>>>(x, y, z, e.q, (int)e.n, apb, e.box,
                                                              e.km, e.P, e.kpart);
    }
    {
        ScopedTimer t(e, P_RHO_FINAL, s);
        if (e.precision == 32)
            k_rho_final<float><<<e.P, 256, 0, s>>>(e.kpart, nbx, e.P, e.km, e.kw, e.box, e.alpha,
                                                    e.qqr2e, e.rho, (KCoef<float>*)e.kcoef, e.kev,
                                                    e.capturing ? e.d_step : nullptr);
        else
            k_rho_final<double><<<e.P, 256, 0, s>>>(e.kpart, nbx, e.P, e.km, e.kw, e.box, e.alpha,
                                                     e.qqr2e, e.rho, (KCoef<double>*)e.kcoef,
                                                     e.kev, e.capturing ? e.d_step : nullptr);
    }
}

void launch_kforce(Engine& e, const double* x, const double* y, const double* z, void* fx,
                   void* fy, void* fz, double* vx, double* vy, double* vz, const double* mass,
                   double dt, hipStream_t s)
{
    ScopedTimer t(e, P_KFORCE, s);
    const int n = (int)e.n;
    const int blocks = (n + kKfAtoms - 1) / kKfAtoms;
    const bool kick = vx != nullptr;
#define RBMD_KF(T, K)                                                                          \
    k_kforce<T, K><<<blocks, kKfThreads, 0, s>>>(x, y, z, e.q, n, e.box,                       \
                                                 (const KCoef<T>*)e.kcoef, e.P, (T*)fx, (T*)fy, \
                                                 (T*)fz, vx, vy, vz, mass, dt)
    if (e.precision == 32) {
        if (kick) RBMD_KF(float, true);
        else RBMD_KF(float, false);
    } else {
        if (kick) RBMD_KF(double, true);
        else RBMD_KF(double, false);
    }
#undef RBMD_KF
}

}  // namespace rbmd
