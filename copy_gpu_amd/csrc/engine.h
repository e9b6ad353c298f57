// engine.h -- internal state and device helpers of the MI355X charged-LJ force engine.
//
// Layout in HBM (n atoms, C = list row capacity, P = k-vectors):
//   x_ref[3][n]  FP64  wrapped positions at the last list build (original atom order)
//   xs_ref[n]    double4 (x,y,z,q) of x_ref in cell-sorted order (list build input)
//   xs[n]        double4 (x,y,z,q) current positions, unwrapped against x_ref, cell-sorted
//   perm[n]      int32  sorted slot -> original atom id
//   scell[n]     int32  cell of each sorted slot
//   nlist[n*C]   uint16 row-major Verlet rows of tile-local indices: the LJ-range
//                       ("inner") entries, then the rest ("outer"), then one sentinel
//   ncount[n]    int2   (inner, outer) per row; (n, 0) with n >= C for a row that
//                       did not fit
//   cell_count/cell_start  int32 counting-sort histogram and CSR bounds
//   k-space: kvec (m, w), partial structure factors [blocks][P] (FP64), rho[P], per-k
//   force coefficients, per-k energy/virial terms
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <string>

#define RBMD_HD __host__ __device__

namespace rbmd {

constexpr int kWave = 64;
constexpr int kMaxTypes = 8;

// Orthorhombic periodic box passed by value to kernels.
struct Box {
    double lo[3];
    double L[3];
    double invL[3];
};

// Parameters of the pair pass, by value.
struct PairParams {
    double rcl2, rcc2;  // squared cutoffs (FP64 decisions)
    double alpha, sqrt_alpha, two_sqrt_alpha_over_pi, qqr2e;
    int ntypes;
    double eps24[kMaxTypes * kMaxTypes];  // 24 * sqrt(eps_i eps_j)
    double sig2[kMaxTypes * kMaxTypes];   // ((sig_i + sig_j)/2)^2
};

// Bins of the MD step's next rebuild, double-buffered by the step epoch
// (flags[3]): the tile pass of a step bins its home atoms' current positions into
// buffer (epoch + 1) & 1 and clears buffer epoch & 1, which the cell sort of the
// same step has consumed; the k-space force (or kick) kernel then advances the
// epoch.
// MD-step bins: an atom is binned under its cell-sorted slot at the binning, so
// the per-atom records are written in slot order (whole cache lines from a tile's
// consecutive home atoms) rather than scattered by atom id.
constexpr int kBinChunk = 256;  // cells per chunk count (the cell sort's prefix sums)
// Chunk counts are kept only above this many cells.  Below it each cell's start is
// summed over every earlier cell (a few us per rebuild at 13,824 cells), which is
// cheaper than the chunk atomics' cost on every step: the tile pass's wave-aggregated
// chunk additions land on ncell / 256 counters and serialise there (about 4 us per
// step at 729 and at 2,197 cells).
constexpr int kBinChunkMinCells = 16384;

struct MdBins {
    int32_t* count;        // [2][ncell]
    int32_t* chunk;        // [2][nchunk]: sums of count over kBinChunk consecutive cells
                           // (nchunk = 0: not kept, below kBinChunkMinCells cells)
    int32_t* bucket;       // [2][ncell][cap] slots at the binning
    int32_t* cell;         // [n] by slot: cell at the binning
    int32_t* vid;          // [n] by slot: atom id
    double *xw, *yw, *zw;  // [n] by slot: wrapped positions at the binning
    int ncell, cap, nchunk;
};

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11.  Round: (hi0,lo0) = M0*c0, (hi1,lo1) = M1*c2,
// c = (hi1^c1^k0, lo1, hi0^c3^k1, lo0), key bumped by the Weyl constants.
RBMD_HD inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                  uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Random batch list shell selection: whether the shell pair (i, j) -- original atom
// ids, i the atom whose row holds j -- is kept at list build `epoch`: Philox4x32-10
// with key = seed and counter (i, j, epoch, 0x52424C53), word 0 divisible by pice
// (probability 1/pice up to 2^-32 rounding).  The reference keeps every pice-th shell
// pair of its stencil sweep instead (ComputeRBLNeighboursOnceWorklet,
// rbmd/framework/run/worklet/RunWorklet.cpp:690-697), an order the cell sort's ties
// decide; a counter-based draw makes the sample reproducible.
RBMD_HD inline bool rbl_select(uint64_t seed, uint32_t epoch, int32_t i, int32_t j, uint32_t pice)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)j, epoch, 0x52424C53u};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    return r[0] % pice == 0u;
}

// 53-bit uniform in [0,1) from two 32-bit words (exact in FP64).
RBMD_HD inline double u53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Discrete-Gaussian inverse-CDF lookup: first index i with u < cdf[i]
// (cdf has len entries, cdf[len-1] == 1 up to rounding).
RBMD_HD inline int cdf_lookup(const double* cdf, int len, double u)
{
    int lo = 0, hi = len - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (u < cdf[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// The k-sampler draw shared by host and device: lattice vector of sample l at
// a given step.  Key = seed, counter = (l, attempt*2 + half, step_lo, step_hi ^ tag).
// cdf: three concatenated per-axis tables, each of length 2*M_d+1 for m in [-M_d, M_d].
RBMD_HD inline void sample_kvector(uint64_t seed, uint64_t step, uint32_t l, const double* cdf,
                                   const int* M, int m_out[3])
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t attempt = 0;; ++attempt) {
        uint32_t r0[4], r1[4];
        uint32_t c0[4] = {l, attempt * 2u, (uint32_t)step, (uint32_t)(step >> 32) ^ 0x6B537631u};
        uint32_t c1[4] = {l, attempt * 2u + 1u, (uint32_t)step,
                          (uint32_t)(step >> 32) ^ 0x6B537631u};
        philox4x32_10(c0, key, r0);
        philox4x32_10(c1, key, r1);
        const double u[3] = {u53(r0[0], r0[1]), u53(r0[2], r0[3]), u53(r1[0], r1[1])};
        int off = 0;
        for (int d = 0; d < 3; ++d) {
            int len = 2 * M[d] + 1;
            m_out[d] = cdf_lookup(cdf + off, len, u[d]) - M[d];
            off += len;
        }
        if (m_out[0] != 0 || m_out[1] != 0 || m_out[2] != 0) return;
        if (attempt > 1000000u) {  // unreachable for any sane box; keeps the loop finite
            m_out[0] = 1;
            return;
        }
    }
}

// Tabulated (learned) k-space proposal mixed with the analytic sampler:
// q_mix(m) = lambda * p_grid(m) [|m_d| <= Mg] + (1 - lambda) * a(m) / S,
// a(m) = exp(-k^2 / 4 alpha).  cdf is the inclusive running sum of p_grid over
// the (2Mg+1)^3 box in x-fastest order (closing at exactly 1), ratio[i] =
// a(m_i) / q_mix(m_i), and w_out = S / ((1 - lambda) P) is the weight of a draw
// outside the box.  cdf == nullptr selects the analytic sampler alone.
// guide[j] (j < 2^guide_bits) is the first index i with cdf[i] > j / 2^bits, so
// the inverse-CDF lookup starts there and walks ~1 entry (guide-table method,
// Chen & Asau 1974); with a power-of-two table u * 2^bits is exact.
struct GridProposal {
    const double* cdf = nullptr;
    const double* ratio = nullptr;
    const int32_t* guide = nullptr;
    int guide_bits = 0;
    int Mg = 0;
    double lambda = 0;
};

// First index i with u < cdf[i] (u in [0, 1), cdf closing at 1).
RBMD_HD inline int guided_lookup(const GridProposal& g, double u)
{
    int i = g.guide[(int)(u * (double)(1 << g.guide_bits))];
    while (u >= g.cdf[i]) ++i;
    return i;
}

// Draw of sample l (of P) and its importance weight w = a(m) / (P q(m)).
// Mixture choice and grid index come from counter (l, 0x80000000, step_lo,
// step_hi ^ 0x4C524E44); analytic draws use sample_kvector's counters.
RBMD_HD inline double draw_kvector(uint64_t seed, uint64_t step, uint32_t l, const double* cdf,
                                   const int* M, const GridProposal& g, int P, double w_analytic,
                                   double w_out, int m_out[3])
{
    if (!g.cdf) {
        sample_kvector(seed, step, l, cdf, M, m_out);
        return w_analytic;
    }
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t c[4] = {l, 0x80000000u, (uint32_t)step, (uint32_t)(step >> 32) ^ 0x4C524E44u};
    uint32_t r[4];
    philox4x32_10(c, key, r);
    const int side = 2 * g.Mg + 1;
    if (u53(r[0], r[1]) < g.lambda) {
        const int idx = guided_lookup(g, u53(r[2], r[3]));
        m_out[0] = idx % side - g.Mg;
        m_out[1] = (idx / side) % side - g.Mg;
        m_out[2] = idx / (side * side) - g.Mg;
        return g.ratio[idx] / (double)P;
    }
    sample_kvector(seed, step, l, cdf, M, m_out);
    if (m_out[0] < -g.Mg || m_out[0] > g.Mg || m_out[1] < -g.Mg || m_out[1] > g.Mg ||
        m_out[2] < -g.Mg || m_out[2] > g.Mg)
        return w_out;
    const int idx = (m_out[0] + g.Mg) + side * ((m_out[1] + g.Mg) + side * (m_out[2] + g.Mg));
    return g.ratio[idx] / (double)P;
}

// ---------------------------------------------------------------- NVT thermostats
// Velocity thermostats of ExecutionNVT::UpdateVelocityByTempConType and
// TempConTypeForce (rbmd/framework/run/ExecutionNVT.cpp:256-291, 735-747).
enum : int {
    kThermoNone = 0,
    kThermoBerendsen = 1,
    kThermoRescale = 2,
    kThermoNoseHoover = 3,
    kThermoLangevin = 4,
    kThermoMeasure = 5,  // no velocity change: the temperature is only measured and logged
};

// Constants and device state of a thermostat, by value in the closing-kick kernel
// and the reduction after it.  Each workgroup of the closing kernel stores sum
// m |v|^2 of its atoms after the closing kick in part[blk]; the reduction kernel
// sums the partials in a fixed order, forms T and the thermostat's coefficient,
// and logs T.
struct ThermoArgs {
    int type = kThermoNone;
    int nblk = 0;
    double* part = nullptr;             // [nblk] per-workgroup sum of m |v|^2
    uint32_t* cnt = nullptr;            // [1] temperature log position ([0] unused)
    double* st = nullptr;               // [0] velocity scale for the next opening kick,
                                        // [1] Nose-Hoover xi, [2] temperature of the last step
    unsigned long long* noise = nullptr;  // Langevin noise step
    double t0 = 0, dt_over_damp = 0, inv_dof_kb = 0, gamma = 0, kbt = 0;
    uint64_t seed = 0;
    double* log = nullptr;  // temperature per step (device), or null
    int64_t log_len = 0;
};

// Gaussian 3-vector of the Langevin force at a noise step (one vector per step for
// every atom, as the reference draws it): Philox counters (0x4C414E47, j, step_lo,
// step_hi) for j = 0, 1, then Box-Muller on (1 - u, u') pairs so the logarithm's
// argument lies in (0, 1].
RBMD_HD inline void langevin_gaussian(uint64_t seed, uint64_t step, double g[3])
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r0[4], r1[4];
    const uint32_t c0[4] = {0x4C414E47u, 0u, (uint32_t)step, (uint32_t)(step >> 32)};
    const uint32_t c1[4] = {0x4C414E47u, 1u, (uint32_t)step, (uint32_t)(step >> 32)};
    philox4x32_10(c0, key, r0);
    philox4x32_10(c1, key, r1);
    const double two_pi = 6.283185307179586;
    const double a = sqrt(-2.0 * log(1.0 - u53(r0[0], r0[1])));
    const double t = two_pi * u53(r0[2], r0[3]);
    const double b = sqrt(-2.0 * log(1.0 - u53(r1[0], r1[1])));
    const double t2 = two_pi * u53(r1[2], r1[3]);
    g[0] = a * cos(t);
    g[1] = a * sin(t);
    g[2] = b * cos(t2);
}

#ifdef __HIPCC__
// ---------------------------------------------------------------- wave64 helpers
// Minimum-image displacement along one axis of a periodic box.
__device__ inline double min_image(double t, double L, double invL)
{
    return t - L * rint(t * invL);
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Number of set bits of mask below this lane (ballot compaction offset).
__device__ inline uint32_t mbcnt64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename T>
__device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave64 sum delivered to every lane, with DPP lane moves instead of LDS
// permutes: quad xor 1, quad xor 2, half-row mirror, row mirror (each lane then
// holds its 16-lane row sum), row_bcast15 / row_bcast31 (lane 63 holds the
// total), then a readlane.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}

__device__ inline float wave_sum_dpp(float v)
{
    v += __int_as_float(dpp_i<0xB1>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x4E>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x141>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x140>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x142, 0xa>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x143, 0xc>(__float_as_int(v)));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ inline double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL, ROW_MASK>((int)(b & 0xffffffffll));
    const int hi = dpp_i<CTRL, ROW_MASK>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ inline double wave_sum_dpp(double v)
{
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    v += dpp_d<0x142, 0xa>(v);
    v += dpp_d<0x143, 0xc>(v);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
#endif

// ---------------------------------------------------------------- timing
enum Pass {
    P_CHECK = 0,
    P_BIN,
    P_SCAN,
    P_SCATTER,
    P_CELLSORT,
    P_LIST,
    P_GATHER,
    P_RHO,
    P_RHO_FINAL,
    P_PAIR,
    P_KFORCE,
    P_KICK,
    P_EV_FINAL,
    P_SAMPLE,
    P_NUM
};

}  // namespace rbmd
