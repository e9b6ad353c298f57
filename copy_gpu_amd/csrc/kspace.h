// kspace.h -- k-space kernel bodies shared by the k-space launches and the
// fused MD-step kernels: the structure factor rho(k) = sum_j q_j exp(i k.r_j)
// (per-workgroup partial sums, then a fixed-order final reduction with the
// force coefficients and per-k energy / virial terms) and the per-atom k-space
// force.
//
// Reference (paths relative to the reference checkout):
//   rho(k)                 rbmd/framework/run/worklet/RunWorklet.cpp:1616-1677,
//                          rbmd/framework/run/ExecutionMD.cpp:412-448
//   RBE force              RunWorklet.cpp:1508-1548, ExecForceFunction.h:247-269
//   Ewald force / energy   ExecForceFunction.h:226-245, output/ThermoOutput.cpp:209-255
//
// Phases are formed from FP64 fractional coordinates, phi = frac(m . s) with
// s = (x - lo)/L, so FP32 sines keep full accuracy at any |k.r|.  FP32 mode sums
// each lane's products q cos, q sin over blocks of kRhoFlush atoms in FP32 and
// adds every block sum into an FP64 accumulator (the compensation: no rounding
// error grows past one block), and reduces the per-workgroup partials in FP64,
// deterministically.
#pragma once
#include "state.h"

namespace rbmd {

constexpr double kPi = 3.14159265358979323846;

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

// Fused multiply-add in T, written out: the k-space sums run with implicit
// contraction off (so kernels that share them give the same bits), and these
// are the FMAs they use on purpose.
__device__ inline float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return fma(a, b, c); }

// Atoms per block of a lane's T-precision rho sums before the block is added into
// FP64.  In FP32 a block of 16 unit-size products carries an error of a few 1e-7 at
// most; the FP64 sum of the blocks adds none that matters.
constexpr int kRhoFlush = 16;

// Charge in the staged atom record of the rho partials: FP64, or in FP32 mode the
// FP32 charge's bits in the low word (read back without a conversion per k).
template <typename T>
__device__ inline double rho_stage_q(double q)
{
    if constexpr (sizeof(T) == 4) return __hiloint2double(0, __float_as_int((float)q));
    return q;
}

template <typename T>
__device__ inline T rho_staged_q(double w)
{
    if constexpr (sizeof(T) == 4) return __int_as_float(__double2loint(w));
    return w;
}

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
    // LDS of one workgroup: kRhoStage staged atoms + the per-wave sums
    static constexpr size_t kLdsBytes =
        kRhoStage * sizeof(double4) + (kRhoBlock / 64) * 64 * 2 * sizeof(double);
    __device__ static void run(const Args& A, int blk, int)
    {
        __shared__ double4 lds[kLdsBytes / sizeof(double4)];
        run_lds(A, blk, lds);
    }
    __device__ static void run_lds(const Args& A, int blk, double4* st)
    {
#pragma clang fp contract(off)
        double(*red)[64][2] = (double(*)[64][2])(st + kRhoStage);
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
        double re = 0.0, im = 0.0;
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
                                                   (A.z[i] - box.lo[2]) * box.invL[2],
                                                   rho_stage_q<T>(A.q[i]));
                else
                    st[threadIdx.x] = make_double4(0, 0, 0, 0);
            }
            __syncthreads();
            const int cnt = (a1 - base) < kRhoStage ? (a1 - base) : kRhoStage;
            if (kok) {
                constexpr int kStep = kRhoBlock / 64;
                for (int b0 = wv; b0 < cnt; b0 += kStep * kRhoFlush) {
                    const int b1 = (b0 + kStep * kRhoFlush) < cnt ? (b0 + kStep * kRhoFlush) : cnt;
                    T br = 0, bi = 0;
                    for (int a = b0; a < b1; a += kStep) {
                        const double4 p = st[a];
                        const double phi = fma(mz, p.z, fma(my, p.y, mx * p.x));
                        T sn, cs;
                        sincos2pi<T>(phi, &sn, &cs);
                        const T qa = rho_staged_q<T>(p.w);
                        br = fma_t(qa, cs, br);
                        bi = fma_t(qa, sn, bi);
                    }
                    re += (double)br;
                    im += (double)bi;
                }
            }
        }
        red[wv][lane][0] = re;
        red[wv][lane][1] = im;
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
// FP64 sums of each k's atom-tile partials in a fixed order, then its force
// coefficients and energy / virial terms (rho_final_emit).
template <typename T>
struct RhoFinalArgs {
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

template <typename T>
__device__ inline void rho_final_emit(const RhoFinalArgs<T>& A, int k, double re, double im)
{
    const Box& box = A.box;
    A.rho[2 * k] = re;
    A.rho[2 * k + 1] = im;
    const double mm[3] = {(double)A.km[3 * k], (double)A.km[3 * k + 1], (double)A.km[3 * k + 2]};
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

// Beside the tile pass (TileRho): one wave per k, NT / 64 k's per workgroup
// (each lane strided over the partials, then a DPP wave sum), so that few
// workgroups share the tile pass's first dispatch round.
template <typename T, int NT>
struct RhoFinalT {
    static constexpr int kThreads = NT;
    using Args = RhoFinalArgs<T>;
    static constexpr int kPerBlock = NT / 64;
    __device__ static void run(const Args& A, int blk, int)
    {
        const int lane = threadIdx.x & 63;
        const int k = blk * kPerBlock + (int)(threadIdx.x >> 6);
        const int P = A.P;
        if (k >= P) return;
        double re = 0.0, im = 0.0;
        for (int b = lane; b < A.nbx; b += 64) {
            re += A.kpart[((size_t)b * P + k) * 2];
            im += A.kpart[((size_t)b * P + k) * 2 + 1];
        }
        re = wave_sum_dpp(re);
        im = wave_sum_dpp(im);
        if (lane == 0) rho_final_emit(A, k, re, im);
    }
};

// Workgroups of RhoFinalT over P k-vectors.
template <typename T, int NT>
inline int rho_final_blocks(int P)
{
    return (P + RhoFinalT<T, NT>::kPerBlock - 1) / RhoFinalT<T, NT>::kPerBlock;
}

// Standalone launch (the force pass, where the reduction's latency is exposed):
// one workgroup of 256 threads per k (strided sums, DPP wave sums, then the four
// waves in order), a shorter chain per lane than one wave per k.
template <typename T>
struct RhoFinal {
    static constexpr int kThreads = 256;
    using Args = RhoFinalArgs<T>;
    __device__ static void run(const Args& A, int blk, int)
    {
        __shared__ double wr[kThreads / 64][2];
        const int k = blk, t = threadIdx.x, lane = t & 63, wv = t >> 6;
        const int P = A.P;
        double r = 0.0, m = 0.0;
        for (int b = t; b < A.nbx; b += kThreads) {
            r += A.kpart[((size_t)b * P + k) * 2];
            m += A.kpart[((size_t)b * P + k) * 2 + 1];
        }
        r = wave_sum_dpp(r);
        m = wave_sum_dpp(m);
        if (lane == 0) {
            wr[wv][0] = r;
            wr[wv][1] = m;
        }
        __syncthreads();
        if (t == 0) {
            double re = 0.0, im = 0.0;
            for (int w = 0; w < kThreads / 64; ++w) {
                re += wr[w][0];
                im += wr[w][1];
            }
            rho_final_emit(A, k, re, im);
        }
    }
};

// ---------------------------------------------------------------- k-space force
// Workgroup = 8 waves = 8 / s atom groups of 64 atoms (lanes, original order) x s
// k-slices: the waves of a group take k = slice, slice + s, ... of the coefficients
// staged in LDS (every lane reads the same k: broadcast), then the s partial forces
// of each atom are summed in a fixed order; slice 0 completes and kicks the atoms.
// s = 8 (64 atoms per workgroup) unless P <= kKfWideP and that would give more
// than kKfTargetBlocks workgroups: then the fewest slices that still give that many
// (at 100,000 atoms and P = 50, s = 4), so fewer waves repeat the per-atom work.
// F_i += q_i sum_k k (cos(phi) A_k - sin(phi) B_k).
constexpr int kKfAtoms = 64;
constexpr int kKfSlices = 8;
constexpr int kKfThreads = kKfAtoms * kKfSlices;
constexpr int kKfStage = 512;
constexpr int kKfWideP = 128;
constexpr int kKfTargetBlocks = 512;  // two per CU on the 256 CUs of an MI355X

// k-slices per atom group of a k-space force launch over n atoms and P k-vectors
// (P = 0: the kernel only completes the forces / kicks), and its workgroups
inline int kf_slices(int64_t n, int P)
{
    if (P > kKfWideP) return kKfSlices;
    int s = 1;
    while (s < kKfSlices && (n * s + kKfThreads - 1) / kKfThreads < kKfTargetBlocks) s *= 2;
    return s;
}
inline int kf_blocks(int64_t n, int slices)
{
    const int64_t apb = (int64_t)kKfAtoms * (kKfSlices / slices);
    return (int)((n + apb - 1) / apb);
}

template <typename T>
struct KForceArgs {
    const double *x, *y, *z, *q;
    int n;
    Box box;
    const KCoef<T>* coef;
    int P;  // 0: no k-space term (the kernel only completes / kicks)
    T *fx, *fy, *fz;
    double *vx, *vy, *vz;
    const double* mass;
    double dt;
    uint32_t* flags;
    ThermoArgs th;  // velocity thermostat of the closing kick (KForce<T, true>)
    const double* rbl_mean = nullptr;  // random batch list: mean near force to subtract
    int slices = kKfSlices;            // k-slices per atom group (kf_slices)
};

template <typename T>
struct KForceLds {
    KCoef<T> st[kKfStage];
    T red[kKfSlices][3][kKfAtoms];
};

// The k-space force of the atoms of workgroup blk.  Each lane of an atom-owning
// wave (slice 0 of its atom group) first calls pre(i) for its atom i < n (loads
// that the completion needs, issued before the k-sum so that their latency
// overlaps it), then, after the k-sum, fin(i, f0, f1, f2, pre_value) with the
// complete force (pair force + k-space).
// fin runs with FP contraction off in the caller's body.
template <typename T, class Pre, class Fin>
__device__ inline void kforce_block(const KForceArgs<T>& A, int blk, KForceLds<T>& L, Pre pre,
                                    Fin fin)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = A.n, P = A.P;
    const int slices = uniform(A.slices);
    const int slice = wv % slices;
    const int i = blk * (kKfAtoms * (kKfSlices / slices)) + (wv / slices) * kKfAtoms + lane;
    const Box& box = A.box;
    const bool own = slice == 0 && i < n;
    double sx = 0, sy = 0, sz = 0;
    if (i < n && P > 0) {
        sx = (A.x[i] - box.lo[0]) * box.invL[0];
        sy = (A.y[i] - box.lo[1]) * box.invL[1];
        sz = (A.z[i] - box.lo[2]) * box.invL[2];
    }
    T f0 = 0, f1 = 0, f2 = 0, qi = 0;
    decltype(pre(0)) pv{};
    if (own) {
        f0 = A.fx[i];
        f1 = A.fy[i];
        f2 = A.fz[i];
        if (A.rbl_mean) {
            // the random batch list's mean-force correction (SumRBLCorrForceWorklet,
            // rbmd/framework/run/worklet/RunWorklet.cpp:1304-1321)
            f0 = (T)((double)f0 - A.rbl_mean[0]);
            f1 = (T)((double)f1 - A.rbl_mean[1]);
            f2 = (T)((double)f2 - A.rbl_mean[2]);
        }
        if (P > 0) qi = (T)A.q[i];
        pv = pre(i);
    }
    T ax = 0, ay = 0, az = 0;
    for (int base = 0; base < P; base += kKfStage) {
        __syncthreads();
        for (int k = threadIdx.x; k < kKfStage && base + k < P; k += kKfThreads)
            L.st[k] = A.coef[base + k];
        __syncthreads();
        const int cnt = (P - base) < kKfStage ? (P - base) : kKfStage;
        for (int k = slice; k < cnt; k += slices) {
            const KCoef<T>& c = L.st[k];
            const double phi = fma(c.m[2], sz, fma(c.m[1], sy, c.m[0] * sx));
            T sn, cs;
            sincos2pi<T>(phi, &sn, &cs);
            const T g = fma_t(cs, c.A, -(sn * c.B));
            ax = fma_t(g, c.k[0], ax);
            ay = fma_t(g, c.k[1], ay);
            az = fma_t(g, c.k[2], az);
        }
    }
    if (P > 0 && slices > 1) {
        L.red[wv][0][lane] = ax;
        L.red[wv][1][lane] = ay;
        L.red[wv][2][lane] = az;
        __syncthreads();
    }
    if (own) {
        if (P > 0) {
            T s0 = 0, s1 = 0, s2 = 0;
            if (slices > 1) {
                // the slices of this wave's atom group: waves wv .. wv + slices - 1
                for (int w = wv; w < wv + slices; ++w) {
                    s0 += L.red[w][0][lane];
                    s1 += L.red[w][1][lane];
                    s2 += L.red[w][2][lane];
                }
            } else {
                s0 += ax;
                s1 += ay;
                s2 += az;
            }
            f0 = f0 + qi * s0;
            f1 = f1 + qi * s1;
            f2 = f2 + qi * s2;
        }
        fin(i, f0, f1, f2, pv);
    }
}

// Velocity and mass of atom i for the closing half kick, and the thermostat's
// per-step inputs (Nose-Hoover friction, Langevin noise vector).
struct KickIn {
    double v[3], m;
};

// Thermostat coefficient from the kinetic energy sum of a step (sum m |v|^2 after
// the closing kick): T = sum / (dof kB), then the coefficient of the thermostat
// (ExecutionNVT::UpdateVelocityByTempConType, rbmd/framework/run/ExecutionNVT.cpp:256-291):
// the velocity scale of Berendsen (sqrt(1 + dt/Tdamp (T0/T - 1))) or of the plain
// rescale (sqrt(T0/T)), applied by the next opening kick, or the Nose-Hoover
// friction update xi += dt/2 (T/T0 - 1) / (dt/10); T is logged and the Langevin
// noise step advanced.  One thread.
__device__ inline double thermo_scale(const ThermoArgs& th, double sum)
{
    const double T = sum * th.inv_dof_kb;
    if (th.type == kThermoBerendsen) return sqrt(1.0 + th.dt_over_damp * (th.t0 / T - 1.0));
    if (th.type == kThermoRescale) return sqrt(th.t0 / T);
    return 1.0;
}

__device__ inline void thermo_finish(const ThermoArgs& th, double sum, double dt)
{
    const double T = sum * th.inv_dof_kb;
    double xi = th.st[1];
    if (th.type == kThermoNoseHoover) xi = xi + 0.5 * dt * (T / th.t0 - 1.0) / (0.1 * dt);
    th.st[0] = thermo_scale(th, sum);
    th.st[1] = xi;
    th.st[2] = T;
    const uint32_t pos = th.cnt[1];
    if (th.log && (int64_t)pos < th.log_len) th.log[pos] = T;
    th.cnt[1] = pos + 1u;
    if (th.type == kThermoLangevin) th.noise[0] += 1ull;
}

// Sum of the closing kernel's per-workgroup partials by a workgroup of NT threads
// in a fixed order (thread t takes j = t, t + NT, ...; then waves, then the NT/64
// wave sums in order): every workgroup that calls it gets the same bits.  All
// threads of the workgroup must call it.
template <int NT>
__device__ inline double thermo_sum(const ThermoArgs& th)
{
    __shared__ double red[NT / 64];
    const int t = threadIdx.x;
    double acc = 0;
    for (int j = t; j < th.nblk; j += NT) acc += th.part[j];
    acc = wave_sum(acc);
    if ((t & 63) == 0) red[t >> 6] = acc;
    __syncthreads();
    double s = 0;
    for (int w = 0; w < NT / 64; ++w) s += red[w];
    return s;
}

// The step's kinetic energy from the closing kernel's per-workgroup partials, summed
// by one workgroup in a fixed order (the same bits on every run), then
// thermo_finish.  Runs after the closing kernel, so no ordering between
// workgroups of one kernel is needed.
struct ThermoReduce {
    static constexpr int kThreads = 256;
    struct Args {
        ThermoArgs th;
        double dt;
    };
    __device__ static void run(const Args& a, int, int)
    {
        if (a.th.type == kThermoNone) return;  // a batch replica without a thermostat
        const double s = thermo_sum<kThreads>(a.th);
        if (threadIdx.x == 0) thermo_finish(a.th, s, a.dt);
    }
};

// F += k-space force; with KICK the closing half kick v += dt/2 F/m of the
// velocity-Verlet step follows (and the MD step epoch advances), with the
// thermostat of A.th: the Langevin force F += -m gamma v + sqrt(2 m gamma kBT/dt) g
// before the kick (UnderdampedLangevinWorklet, rbmd/framework/run/worklet/RunWorklet.cpp:1700-1728),
// the Nose-Hoover half step v += dt/2 (F/m - xi v) after it
// (UpdateVelocityNoseHooverWorklet, RunWorklet.cpp:1784-1804), and each
// workgroup's sum of m |v|^2 of the kicked velocities (reduced by ThermoReduce).
template <typename T, bool KICK>
struct KForce {
    static constexpr int kThreads = kKfThreads;
    using Args = KForceArgs<T>;
    __device__ static void run(const Args& A, int blk, int)
    {
#pragma clang fp contract(off)
        if (KICK && blk == 0 && threadIdx.x == 0) A.flags[3] += 1u;  // MD step epoch
        __shared__ KForceLds<T> L;
        const ThermoArgs& th = A.th;
        double ke = 0;
        kforce_block(
            A, blk, L,
            [&](int i) {
                KickIn k{};
                if (KICK) {
                    k.v[0] = A.vx[i];
                    k.v[1] = A.vy[i];
                    k.v[2] = A.vz[i];
                    k.m = A.mass ? A.mass[i] : 1.0;
                }
                return k;
            },
            [&](int i, T f0, T f1, T f2, const KickIn& k) {
#pragma clang fp contract(off)
                if (KICK && th.type == kThermoLangevin) {
                    double g[3];
                    langevin_gaussian(th.seed, th.noise[0], g);
                    const double mg = k.m * th.gamma;
                    const double amp = sqrt(2.0 * k.m * th.gamma * th.kbt / A.dt);
                    f0 = (T)((double)f0 - mg * k.v[0] + amp * g[0]);
                    f1 = (T)((double)f1 - mg * k.v[1] + amp * g[1]);
                    f2 = (T)((double)f2 - mg * k.v[2] + amp * g[2]);
                }
                A.fx[i] = f0;
                A.fy[i] = f1;
                A.fz[i] = f2;
                if (KICK) {
                    const double h = 0.5 * A.dt / k.m;
                    double v0 = k.v[0] + h * (double)f0;
                    double v1 = k.v[1] + h * (double)f1;
                    double v2 = k.v[2] + h * (double)f2;
                    if (th.type != kThermoNone) {
                        ke = k.m * (v0 * v0 + v1 * v1 + v2 * v2);
                        if (th.type == kThermoNoseHoover) {
                            const double xi = th.st[1], hd = 0.5 * A.dt;
                            v0 = v0 + hd * ((double)f0 / k.m - xi * v0);
                            v1 = v1 + hd * ((double)f1 / k.m - xi * v1);
                            v2 = v2 + hd * ((double)f2 / k.m - xi * v2);
                        }
                    }
                    A.vx[i] = v0;
                    A.vy[i] = v1;
                    A.vz[i] = v2;
                }
            });
        if (KICK && th.type != kThermoNone) {
            // one partial per workgroup: the waves holding atoms (wave 0 alone, or all
            // eight with few k-vectors; the others add zero) in wave order
            __shared__ double kes[kKfThreads / 64];
            const double s = wave_sum(ke);
            if ((threadIdx.x & 63) == 0) kes[threadIdx.x >> 6] = s;
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = 0.0;
                for (int w = 0; w < kKfThreads / 64; ++w) t += kes[w];
                th.part[blk] = t;
            }
        }
    }
};

template <typename T>
inline typename RhoPartial<T>::Args rho_args(Engine& e, const double* x, const double* y,
                                             const double* z, int* blocks)
{
    int apb, nbx, nkt;
    rho_geometry(e, &apb, &nbx, &nkt);
    *blocks = e.P > 0 ? nbx * nkt : 0;
    return typename RhoPartial<T>::Args{x, y, z, e.q, (int)e.n, apb, nbx, e.box, e.km, e.P,
                                        e.kpart};
}

template <typename T>
inline RhoFinalArgs<T> rho_final_args(Engine& e, unsigned long long* step_dev)
{
    int apb, nbx, nkt;
    rho_geometry(e, &apb, &nbx, &nkt);
    return RhoFinalArgs<T>{e.kpart, nbx,     e.P,   e.km,
                           e.kw,    e.box,   e.alpha, e.qqr2e,
                           e.rho,   (KCoef<T>*)e.kcoef, e.kev, step_dev};
}

// k-vectors a k-space force launch of the handle carries (kforce_args) and its
// workgroups
inline int kforce_P(const Engine& e) { return has_kspace(e) ? e.P : 0; }
inline int kforce_blocks(const Engine& e) { return kf_blocks(e.n, kf_slices(e.n, kforce_P(e))); }

template <typename T>
inline KForceArgs<T> kforce_args(Engine& e, const double* x, const double* y, const double* z,
                                 void* fx, void* fy, void* fz, double* vx, double* vy,
                                 double* vz, const double* mass, double dt, bool kspace = true)
{
    KForceArgs<T> a{x,        y,  z,  e.q,  (int)e.n, e.box, (const KCoef<T>*)e.kcoef,
                    kspace ? kforce_P(e) : 0, (T*)fx, (T*)fy, (T*)fz, vx, vy, vz, mass, dt,
                    e.flags};
    a.rbl_mean = e.rbl ? e.rbl_mean : nullptr;
    a.slices = kf_slices(e.n, a.P);
    return a;
}


}  // namespace rbmd
