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

void rho_geometry(const Engine& e, int* atoms_per_block, int* nbx, int* nkt)
{
    const int n = (int)e.n;
    *nkt = (e.P + 63) / 64;
    int apb;
    if (e.rho_apb > 0) {
        apb = e.rho_apb;
    } else {
        int want = 2048 / *nkt;
        if (want < 1) want = 1;
        apb = (n + want - 1) / want;
    }
    // the partials [nbx][P][2] must fit the kpart buffer
    const int64_t max_nbx = e.P > 0 ? e.kpart_cap / (2 * (int64_t)e.P) : 1;
    if (max_nbx >= 1 && (int64_t)apb * max_nbx < (int64_t)n)
        apb = (int)(((int64_t)n + max_nbx - 1) / max_nbx);
    apb = ((apb + 63) / 64) * 64;
    if (apb < 64) apb = 64;
    *atoms_per_block = apb;
    *nbx = (n + apb - 1) / apb;
    if (*nbx < 1) *nbx = 1;
}

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- sampler
struct Sample {
    static constexpr int kThreads = 256;
    using Args = SampleArgs;
    __device__ static void run(const SampleArgs& sa, int blk, int)
    {
        const int l = blk * 256 + threadIdx.x;
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
};

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
// Workgroups = (atom tiles) x (k tiles of 64), atom tile fastest.  Lane <-> k
// within the tile, the 4 waves split the tile's atoms; atoms are staged
// through LDS 256 at a time.
constexpr int kRhoBlock = 256;
constexpr int kRhoStage = 256;

template <typename T>
struct RhoPartial {
    static constexpr int kThreads = kRhoBlock;
    struct Args {
        const double *x, *y, *z, *q;
        int n, atoms_per_block, nbx;
        Box box;
        const int32_t* km;
        int P;
        double* kpart;
    };
    __device__ static void run(const Args& A, int blk, int)
    {
#pragma clang fp contract(off)
        __shared__ double4 st[kRhoStage];
        __shared__ double red[kRhoBlock / 64][64][2];
        const int bx = blk % A.nbx, ky = blk / A.nbx;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int kidx = ky * 64 + lane;
        const int P = A.P;
        const bool kok = kidx < P;
        double mx = 0, my = 0, mz = 0;
        if (kok) {
            mx = (double)A.km[3 * kidx];
            my = (double)A.km[3 * kidx + 1];
            mz = (double)A.km[3 * kidx + 2];
        }
        Kahan<T> re, im;
        const int a0 = bx * A.atoms_per_block;
        int a1 = a0 + A.atoms_per_block;
        a1 = a1 < A.n ? a1 : A.n;
        const Box& box = A.box;
        for (int base = a0; base < a1; base += kRhoStage) {
            __syncthreads();
            const int i = base + threadIdx.x;
            if (threadIdx.x < kRhoStage) {
                if (i < a1)
                    st[threadIdx.x] = make_double4((A.x[i] - box.lo[0]) * box.invL[0],
                                                   (A.y[i] - box.lo[1]) * box.invL[1],
                                                   (A.z[i] - box.lo[2]) * box.invL[2], A.q[i]);
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
            const size_t o = ((size_t)bx * P + kidx) * 2;
            A.kpart[o] = r;
            A.kpart[o + 1] = m;
        }
    }
};

// ---------------------------------------------------------------- rho final
// One workgroup per k: FP64 tree over the atom-tile partials, then the force
// coefficients and the energy / virial terms of this k.
template <typename T>
struct RhoFinal {
    static constexpr int kThreads = 256;
    struct Args {
        const double* kpart;
        int nbx, P;
        const int32_t* km;
        const double* kw;
        Box box;
        double alpha, qqr2e;
        double* rho;
        KCoef<T>* coef;
        double* kev;
        unsigned long long* step_dev;
    };
    __device__ static void run(const Args& A, int blk, int)
    {
        __shared__ double sr[256], si[256];
        const int k = blk, t = threadIdx.x;
        const int P = A.P;
        double r = 0.0, m = 0.0;
        for (int b = t; b < A.nbx; b += 256) {
            r += A.kpart[((size_t)b * P + k) * 2];
            m += A.kpart[((size_t)b * P + k) * 2 + 1];
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
            const Box& box = A.box;
            const double re = sr[0], im = si[0];
            A.rho[2 * k] = re;
            A.rho[2 * k + 1] = im;
            const double mm[3] = {(double)A.km[3 * k], (double)A.km[3 * k + 1],
                                  (double)A.km[3 * k + 2]};
            double kv[3];
            for (int d = 0; d < 3; ++d) kv[d] = 2.0 * kPi * mm[d] / box.L[d];
            const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
            const double V = box.L[0] * box.L[1] * box.L[2];
            const double w = A.kw[k];
            const double c = w * (-4.0 * kPi) / (V * k2) * A.qqr2e;
            KCoef<T> kc;
            for (int d = 0; d < 3; ++d) {
                kc.m[d] = mm[d];
                kc.k[d] = (T)kv[d];
            }
            kc.A = (T)(c * im);
            kc.B = (T)(c * re);
            A.coef[k] = kc;
            const double e = w * (2.0 * kPi / V) * (re * re + im * im) / k2 * A.qqr2e;
            const double b = 2.0 * (1.0 / k2 + 1.0 / (4.0 * A.alpha));
            double* o = A.kev + (size_t)k * 7;
            o[0] = e;
            o[1] = e * (1.0 - b * kv[0] * kv[0]);
            o[2] = e * (1.0 - b * kv[1] * kv[1]);
            o[3] = e * (1.0 - b * kv[2] * kv[2]);
            o[4] = -e * b * kv[0] * kv[1];
            o[5] = -e * b * kv[0] * kv[2];
            o[6] = -e * b * kv[1] * kv[2];
            if (k == 0 && A.step_dev) *A.step_dev += 1ull;
        }
    }
};

// ---------------------------------------------------------------- k-space force
// Workgroup = 64 atoms (lanes, original order) x 8 k-slices (waves): wave w takes
// k = w, w+8, ... of the coefficients staged in LDS (every lane reads the same
// k: broadcast), then the 8 partial forces of each atom are summed in a fixed
// order.  F_i += q_i sum_k k (cos(phi) A_k - sin(phi) B_k).  With KICK the
// final half kick v += dt/2 F/m of the velocity-Verlet step is fused in.
constexpr int kKfAtoms = 64;
constexpr int kKfSlices = 8;
constexpr int kKfThreads = kKfAtoms * kKfSlices;
constexpr int kKfStage = 512;

template <typename T, bool KICK>
struct KForce {
    static constexpr int kThreads = kKfThreads;
    struct Args {
        const double *x, *y, *z, *q;
        int n;
        Box box;
        const KCoef<T>* coef;
        int P;
        T *fx, *fy, *fz;
        double *vx, *vy, *vz;
        const double* mass;
        double dt;
        uint32_t* flags;
    };
    __device__ static void run(const Args& A, int blk, int)
    {
#pragma clang fp contract(off)
        if (KICK && blk == 0 && threadIdx.x == 0) A.flags[3] += 1u;  // MD step epoch
        __shared__ KCoef<T> st[kKfStage];
        __shared__ T red[kKfSlices][3][kKfAtoms];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int i = blk * kKfAtoms + lane;
        const int n = A.n, P = A.P;
        const Box& box = A.box;
        double sx = 0, sy = 0, sz = 0;
        if (i < n) {
            sx = (A.x[i] - box.lo[0]) * box.invL[0];
            sy = (A.y[i] - box.lo[1]) * box.invL[1];
            sz = (A.z[i] - box.lo[2]) * box.invL[2];
        }
        T ax = 0, ay = 0, az = 0;
        for (int base = 0; base < P; base += kKfStage) {
            __syncthreads();
            for (int k = threadIdx.x; k < kKfStage && base + k < P; k += kKfThreads)
                st[k] = A.coef[base + k];
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
            const T qi = (T)A.q[i];
            const T f0 = A.fx[i] + qi * s0, f1 = A.fy[i] + qi * s1, f2 = A.fz[i] + qi * s2;
            A.fx[i] = f0;
            A.fy[i] = f1;
            A.fz[i] = f2;
            if (KICK) {
                const double h = 0.5 * A.dt / (A.mass ? A.mass[i] : 1.0);
                A.vx[i] += h * (double)f0;
                A.vy[i] += h * (double)f1;
                A.vz[i] += h * (double)f2;
            }
        }
    }
};


template <typename T>
typename RhoPartial<T>::Args rho_args(Engine& e, const double* x, const double* y,
                                      const double* z, int* blocks)
{
    int apb, nbx, nkt;
    rho_geometry(e, &apb, &nbx, &nkt);
    *blocks = nbx * nkt;
    return typename RhoPartial<T>::Args{x, y, z, e.q, (int)e.n, apb, nbx, e.box, e.km, e.P,
                                        e.kpart};
}

template <typename T>
typename RhoFinal<T>::Args rho_final_args(Engine& e, unsigned long long* step_dev)
{
    int apb, nbx, nkt;
    rho_geometry(e, &apb, &nbx, &nkt);
    return typename RhoFinal<T>::Args{e.kpart, nbx, e.P, e.km, e.kw, e.box, e.alpha, e.qqr2e,
                                      e.rho, (KCoef<T>*)e.kcoef, e.kev, step_dev};
}

template <typename T, bool K>
typename KForce<T, K>::Args kforce_args(Engine& e, const double* x, const double* y,
                                        const double* z, void* fx, void* fy, void* fz,
                                        double* vx, double* vy, double* vz, const double* mass,
                                        double dt)
{
    return typename KForce<T, K>::Args{x,        y,      z,        e.q,    (int)e.n, e.box,
                                       (const KCoef<T>*)e.kcoef, e.P, (T*)fx, (T*)fy, (T*)fz,
                                       vx,       vy,     vz,       mass,   dt,       e.flags};
}

template <typename T>
hipError_t kspace_prepare_t(Batch& b)
{
    const size_t R = b.es.size();
    std::vector<typename RhoPartial<T>::Args> ra(R);
    std::vector<typename RhoFinal<T>::Args> fa(R);
    std::vector<typename KForce<T, true>::Args> ka(R);
    std::vector<int> nr(R), nf(R), nk(R);
    for (size_t r = 0; r < R; ++r) {
        Engine& e = *b.es[r];
        const RepState& st = b.st[r];
        ra[r] = rho_args<T>(e, st.x, st.y, st.z, &nr[r]);
        fa[r] = rho_final_args<T>(e, e.d_step);
        ka[r] = kforce_args<T, true>(e, st.x, st.y, st.z, st.fx, st.fy, st.fz, st.vx, st.vy,
                                     st.vz, st.mass, b.dt);
        nf[r] = e.P;
        nk[r] = (int)((e.n + kKfAtoms - 1) / kKfAtoms);
    }
    hipError_t err;
    if ((err = phase_upload(b.ph[BP_RHO], ra, nr, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_RHO_FINAL], fa, nf, 0)) != hipSuccess) return err;
    return phase_upload(b.ph[BP_KFORCE], ka, nk, 0);
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
    k_one<Sample><<<(P + 255) / 256, 256, 0, s>>>(sample_args(e, P, seed, step_host));
}

void launch_kweights_default(Engine& e, hipStream_t s)
{
    k_fill<<<(e.P + 255) / 256, 256, 0, s>>>(e.kw, e.P, e.S / (double)e.P);
}

void launch_rho(Engine& e, const double* x, const double* y, const double* z, hipStream_t s)
{
    int blocks;
    unsigned long long* step_dev = e.capturing ? e.d_step : nullptr;
    {
        ScopedTimer t(e, P_RHO, s);
        if (e.precision == 32) {
            const auto a = rho_args<float>(e, x, y, z, &blocks);
            k_one<RhoPartial<float>><<<blocks, kRhoBlock, 0, s>>>(a);
        } else {
            const auto a = rho_args<double>(e, x, y, z, &blocks);
            k_one<RhoPartial<double>><<<blocks, kRhoBlock, 0, s>>>(a);
        }
    }
    {
        ScopedTimer t(e, P_RHO_FINAL, s);
        if (e.precision == 32)
            k_one<RhoFinal<float>><<<e.P, 256, 0, s>>>(rho_final_args<float>(e, step_dev));
        else
            k_one<RhoFinal<double>><<<e.P, 256, 0, s>>>(rho_final_args<double>(e, step_dev));
    }
}

void launch_kforce(Engine& e, const double* x, const double* y, const double* z, void* fx,
                   void* fy, void* fz, double* vx, double* vy, double* vz, const double* mass,
                   double dt, hipStream_t s)
{
    ScopedTimer t(e, P_KFORCE, s);
    const int blocks = (int)((e.n + kKfAtoms - 1) / kKfAtoms);
    const bool kick = vx != nullptr;
#define RBMD_KF(T, K)                                                                          \
    k_one<KForce<T, K>><<<blocks, kKfThreads, 0, s>>>(                                        \
        kforce_args<T, K>(e, x, y, z, fx, fy, fz, vx, vy, vz, mass, dt))
    if (e.precision == 32) {
        if (kick) RBMD_KF(float, true);
        else RBMD_KF(float, false);
    } else {
        if (kick) RBMD_KF(double, true);
        else RBMD_KF(double, false);
    }
#undef RBMD_KF
}

hipError_t kspace_batch_prepare(Batch& b)
{
    return b.precision == 32 ? kspace_prepare_t<float>(b) : kspace_prepare_t<double>(b);
}

void kspace_batch_launch(Batch& b, int phase, hipStream_t s)
{
    const bool f32 = b.precision == 32;
    switch (phase) {
    case BP_RHO:
        if (f32) phase_launch<RhoPartial<float>>(b.ph[phase], s);
        else phase_launch<RhoPartial<double>>(b.ph[phase], s);
        break;
    case BP_RHO_FINAL:
        if (f32) phase_launch<RhoFinal<float>>(b.ph[phase], s);
        else phase_launch<RhoFinal<double>>(b.ph[phase], s);
        break;
    case BP_KFORCE:
        if (f32) phase_launch<KForce<float, true>>(b.ph[phase], s);
        else phase_launch<KForce<double, true>>(b.ph[phase], s);
        break;
    default: break;
    }
}

}  // namespace rbmd
