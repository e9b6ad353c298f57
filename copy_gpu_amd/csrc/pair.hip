// pair.hip -- LJ + erfc real-space Coulomb pair pass and the FP64 energy/virial reduction.
//
// Physics (paths relative to the reference checkout):
//   LJ       rbmd/framework/forceFunction/ExecForceFunction.h:74-96 (energy :296-313)
//   Coulomb  ExecForceFunction.h:125-140 with Gnear :217-224 (energy :315-329)
//   per-atom full-list loop  rbmd/framework/run/worklet/RunWorklet.cpp:170-237
//   virial convention        rbmd-gpu src/force/src/ljforce_op/rocm/ljforce_op.hip.cu:555-575
//
// Mapping: one workgroup per cell tile (home cell + 27-cell stencil staged in
// LDS at the home cell's periodic image), kGroup lanes per home atom (four atoms
// per wave64); the lanes walk the atom's uint16 Verlet row kGroup x kRound
// entries per round (LJ-range entries come first, so only the first rounds
// evaluate LJ), then a DPP reduction within the lane group.
// Pair displacements are FP64 differences of the staged positions (no per-pair
// minimum image), the force maths runs in T (float or double), and cutoff
// decisions use the FP64 squared distance.
#include "half.h"
#include "kspace.h"
#include "state.h"
#include "tile.h"

namespace rbmd {

namespace {


// erfc(z) e^{z^2} for z = sqrt(alpha) r >= 0 in FP32, Abramowitz & Stegun 7.1.26
// (Hastings):
//   erfc(z) = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) e^{-z^2} + eps,
//   t = 1 / (1 + p z),  |eps| <= 1.5e-7 absolute.
// pa = p sqrt(alpha) is folded by the caller, so t = 1 / (1 + pa r).  The caller
// multiplies by e^{-z^2}, which the Gaussian term of G shares.
constexpr double kErfcP = 0.3275911;
__device__ inline float erfc_scaled_as(float r, float pa)
{
    const float t = __builtin_amdgcn_rcpf(fmaf(pa, r, 1.0f));
    float p = 1.061405429f;
    p = fmaf(p, t, -1.453152027f);
    p = fmaf(p, t, 1.421413741f);
    p = fmaf(p, t, -0.284496736f);
    p = fmaf(p, t, 0.254829592f);
    return p * t;
}

struct PairScalars {
    double rcl2, rcc2;
    double alpha, sqrt_alpha, tsp, qqr2e;
    double eps24, sig2;  // type 0 - type 0 (ntypes == 1)
    int ntypes;
    float nal2e;  // -alpha log2(e): e^{-alpha r^2} = 2^{nal2e r^2}
    float pa;     // kErfcP sqrt(alpha)
};

// Scalar force factor fs (force on i = fs * r_ij, r_ij = r_j - r_i) and, with
// EV, the two energies, for one pair at squared distance d2 (FP64 decisions).
template <typename T>
struct PairMath;

// FP32 maths without branches: every lane evaluates both terms and the cutoff
// tests (FP64 d2) select, so a wave needs no exec-mask regions per pair; lanes
// with ok == false (padding, d2 <= 1e-4) contribute exactly zero.
template <>
struct PairMath<float> {
    template <bool LJ>
    __device__ static inline float eval(double d2, float qq, float eps24, float sig2,
                                        const PairScalars& ps, bool ok, float* e_lj,
                                        float* e_re, bool ev)
    {
        const float r2 = (float)d2;
        const float rinv = __builtin_amdgcn_rsqf(r2);
        const float r2inv = rinv * rinv;
        float fs = 0.0f;
        if (LJ) {
            const float sr2 = sig2 * r2inv;
            const float sr6 = sr2 * sr2 * sr2;
            const bool in = ok && d2 < ps.rcl2;
            // -eps24 (2 sr6 - 1) sr6 / r^2, as (sr6 / r^2) (eps24 - 2 eps24 sr6)
            fs = in ? (sr6 * r2inv) * fmaf(-2.0f * eps24, sr6, eps24) : 0.0f;
            if (ev && in) *e_lj += 0.5f * (eps24 * (1.0f / 6.0f)) * (sr6 * sr6 - sr6);
        }
        // G(r) / r = e2 r2inv (P rinv + tsp), P = erfc(sqrt(alpha) r) e^{alpha r^2}
        const float r = r2 * rinv;
        const float e2 = __builtin_amdgcn_exp2f(ps.nal2e * r2);
        const float P = erfc_scaled_as(r, ps.pa);
        const float gr = e2 * r2inv * fmaf(P, rinv, (float)ps.tsp);
        const bool in = ok && d2 < ps.rcc2;
        fs = in ? fmaf(-qq, gr, fs) : fs;
        if (ev && in) *e_re += 0.5f * qq * (P * e2) * rinv;
        return fs;
    }
};

template <>
struct PairMath<double> {
    template <bool LJ>
    __device__ static inline double eval(double d2, double qq, double eps24, double sig2,
                                         const PairScalars& ps, bool ok, double* e_lj,
                                         double* e_re, bool ev)
    {
        double fs = 0.0;
        if (!ok) return fs;
        if (LJ && d2 < ps.rcl2) {
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

// Force factor fs of one pair (force on i = fs * r_ij, r_ij = r_j - r_i) and,
// with EV, its two energies.  qq already carries the Coulomb unit factor.
template <typename T, bool EV, bool MULTI, bool LJ>
__device__ inline T pair_fs(double d2, T qq, int ti, int tj, bool ok, const PairScalars& ps,
                            const PairParams* __restrict__ pp, T* e_lj, T* e_re)
{
    T eps24 = 0, sig2 = 0;
    if (LJ) {
        if (!MULTI) {
            eps24 = (T)ps.eps24;
            sig2 = (T)ps.sig2;
        } else {
            eps24 = (T)pp->eps24[ti * kMaxTypes + tj];
            sig2 = (T)pp->sig2[ti * kMaxTypes + tj];
        }
    }
    return PairMath<T>::template eval<LJ>(d2, qq, eps24, sig2, ps, ok, e_lj, e_re, EV);
}

// One pair: fs (with LJ when lj, a wave-uniform choice), then the force
// accumulated in FP64 (one FMA per component, no conversion of dx), and with EV
// the energies and the virial W_ab += -1/2 fs dx_a dx_b in FP64.  The two
// variants differ only in fs, so the accumulation follows the branch once.
// Random batch list weight of a pair at squared distance d2 (the reference's core
// and shell criteria, RunWorklet.cpp:686-703): 1 inside the core, pice for a selected
// shell pair, 0 otherwise (unselected shell pairs, the 1e-4 bands around r_core^2 and
// the cutoff, which the reference's two lists also leave out).
struct RblPair {
    double rs2, rc2;  // r_core^2, max(rc_lj, rc_coul)^2
    float pice;
};

__device__ inline float rbl_weight(const RblPair& r, double d2, bool sel)
{
    const bool core = (r.rs2 - d2 > 1e-4) && (d2 > 1e-4);
    const bool shell = sel && (r.rc2 - d2 > 1e-4) && (d2 - r.rs2 > 1e-4);
    return core ? 1.0f : (shell ? r.pice : 0.0f);
}

template <typename T, bool EV, bool MULTI, bool RBL = false>
__device__ inline void accumulate_pair(bool lj, double dx, double dy, double dz, T qq, int ti,
                                       int tj, bool valid, const PairScalars& ps,
                                       const PairParams* __restrict__ pp, double& fx, double& fy,
                                       double& fz, double* acc, const RblPair& rb = RblPair{},
                                       bool sel = false)
{
    const double d2 = dx * dx + dy * dy + dz * dz;
    const bool ok = valid && d2 > 1e-4;
    T e_lj = 0, e_re = 0;
    T fs = lj ? pair_fs<T, EV, MULTI, true>(d2, qq, ti, tj, ok, ps, pp, &e_lj, &e_re)
              : pair_fs<T, EV, MULTI, false>(d2, qq, ti, tj, ok, ps, pp, &e_lj, &e_re);
    if (RBL) {
        const T w = (T)rbl_weight(rb, d2, sel);
        fs *= w;
        e_lj *= w;
        e_re *= w;
    }
    const double f = (double)fs;
    fx = fma(f, dx, fx);
    fy = fma(f, dy, fy);
    fz = fma(f, dz, fz);
    if (EV && ok) {
        const double h = -0.5 * f;
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

// Charge in the staged record of the full-row pass: FP64, or in FP32 passes the
// FP32 charge's bits in the low word, so a pair's charge product is one FP32
// multiply (records read from global memory carry the FP64 charge).
template <typename T, bool STAGED>
__device__ inline T rec_q(double w)
{
    if constexpr (STAGED) return rho_staged_q<T>(w);
    return (T)w;
}

// LDS stage of the full-row pass: record u is (x, y) at xy[u] and (z, charge word)
// at zq[u], two arrays of 16-byte entries rather than one of 32-byte records, so a
// wave's random gathers of either half spread over all the LDS banks instead of the
// half that 32-byte strides reach.
struct FullStage {
    char* xy;       // LDS base of the (x, y) array
    uint32_t zoff;  // bytes from xy[u] to zq[u]
    __device__ double4 get(int u) const
    {
        const char* p = xy + 16u * (uint32_t)u;
        const double2 a = *(const double2*)p, b = *(const double2*)(p + zoff);
        return make_double4(a.x, a.y, b.x, b.y);
    }
    __device__ void put(int u, const double4& v) const
    {
        char* p = xy + 16u * (uint32_t)u;
        *(double2*)p = make_double2(v.x, v.y);
        *(double2*)(p + zoff) = make_double2(v.z, v.w);
    }
};

// Position of tile-local candidate u, already moved to the image next to the
// home cell (from LDS when the tile is staged, else from global memory).
// The row sentinel u = total reads the tile's dummy entry (staged at record total).
template <bool STAGED>
__device__ inline double4 tile_pos(const TileHdr& h, const FullStage& lds,
                                   const double4* __restrict__ xs, int u)
{
    if (STAGED) return lds.get(u);
    const int total = h.off[27];
    const int uc = u < total ? u : 0;
    const int st = tile_seg(h, uc);
    const double4 p = xs[h.gstart[st] + (uc - h.off[st])];
    const double4 d = h.dummy;
    return u < total ? make_double4(p.x - h.sx[st], p.y - h.sy[st], p.z - h.sz[st], p.w) : d;
}

template <bool STAGED>
__device__ inline int tile_type(const TileHdr& h, const int32_t* __restrict__ ltype,
                                const int32_t* __restrict__ ts, int u)
{
    if (STAGED) return ltype[u];
    const int total = h.off[27];
    const int uc = u < total ? u : 0;
    const int st = tile_seg(h, uc);
    const int t = ts[h.gstart[st] + (uc - h.off[st])];
    return u < total ? t : 0;
}

// Row chunks per atom per round, and lanes per home-atom row (16 or 32).  On the
// 32,768-ion workload, 16 lanes x 2 chunks measured fastest (3 chunks raise the
// register count past the occupancy the LDS stage allows; 32 lanes shorten the
// rows but leave more rounds per workgroup).
constexpr int kRound = 2;
constexpr int kGroup = 16;

struct PairArgs {
    const double4* xs;
    const int32_t* ts;
    const int32_t* perm;
    const uint16_t* nlist;
    const int2* ncount;
    int cap;
};

// Sum over the kGroup lanes of each lane group: quad xor 1, quad xor 2,
// half-row mirror, row mirror (row sums in all 16 lanes of each row), and for
// 32-lane groups row 1 += row 0, row 3 += row 2 (row broadcast 15).  The group
// sum is valid in lane kSumLane of the group; DPP lane moves, no LDS.
constexpr int kSumLane = kGroup == 32 ? 16 : 0;

template <typename T>
__device__ inline T group_sum(T v);

template <>
__device__ inline float group_sum<float>(float v)
{
    v += __int_as_float(dpp_i<0xB1>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x4E>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x141>(__float_as_int(v)));
    v += __int_as_float(dpp_i<0x140>(__float_as_int(v)));
    if (kGroup == 32) v += __int_as_float(dpp_i<0x142, 0xa>(__float_as_int(v)));
    return v;
}

template <>
__device__ inline double group_sum<double>(double v)
{
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    if (kGroup == 32) v += dpp_d<0x142, 0xa>(v);
    return v;
}

// Maximum of an int over the wave, for a value uniform over each lane group.
__device__ inline int wave_max_int(int v)
{
    int m = max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 32));
    if (kGroup == 16)
        m = max(m, max(__builtin_amdgcn_readlane(v, 16), __builtin_amdgcn_readlane(v, 48)));
    return m;
}

// One wave = 64 / kGroup lane groups, each walking one home atom's row (the
// workgroup's atoms are dealt to the groups in turn).  A round covers
// kRound x kGroup entries per atom: all index loads, then all LDS gathers, then
// the maths; rounds that can still hold LJ-range entries of any of the wave's
// atoms evaluate LJ, later rounds skip it.  When the wave's four rows all
// fitted their capacity (the common case) a lane reads slot min(v, n) of its
// row, n the row length: slots past the end read the sentinel, whose dummy
// entry contributes exactly zero, so the round needs no bounds tests.
// Otherwise (a row that overflowed, or lane groups past the tile's last atom)
// the round tests each slot, and rows that overflowed scan every tile candidate
// instead.
// Random batch list arguments of the pair pass: the pair weights, the shell
// selection the overflowed rows' candidate scan applies itself, and the per-tile
// sums of the near forces (for the mean-force correction).
struct RblArgs {
    RblPair w;
    uint64_t seed;
    uint32_t pice;
    double near2;
    double* part;  // [tile workgroups][3]
};

template <typename T, bool EV, bool MULTI, bool STAGED, bool RBL = false>
__device__ inline void pair_wave(const PairArgs& a, const TileHdr& h, const FullStage& lds,
                                 const int32_t* __restrict__ ltype, int a0, int a1, int h0, int wv,
                                 int lane, const PairScalars& ps,
                                 const PairParams* __restrict__ pp, T* __restrict__ fx,
                                 T* __restrict__ fy, T* __restrict__ fz, double* acc,
                                 const RblArgs& rbl = RblArgs{}, uint32_t epoch = 0,
                                 double* fsum = nullptr)
{
    const int total = h.off[27];
    const int cap = a.cap;
    constexpr int kPer = 64 / kGroup;  // atoms per wave
    const int g = lane / kGroup, gl = lane % kGroup;
    const uint16_t* __restrict__ rows = a.nlist + (size_t)uniform(h0) * cap;  // the tile's rows
    const char* __restrict__ rowb = (const char*)rows;
    for (int t0 = a0 + kPer * wv; t0 < a1; t0 += kPer * kTileWaves) {
        const int s = t0 + g;
        const bool live = s < a1;
        const int2 nc = live ? a.ncount[s] : make_int2(0, 0);
        const int n_in = nc.x, tot = nc.x + nc.y;
        const bool over = tot >= cap;
        const int tl = over ? 0 : tot;
        const int tmax = wave_max_int(tl);
        const int nin_max = wave_max_int(over ? 0 : n_in);
        const bool fast = __ballot(!live || over) == 0ull;
        const int ui = h.off[kHomeSeg] + (live ? s - h0 : 0);
        const double4 pi = STAGED ? lds.get(ui) : a.xs[live ? s : h0];
        const int ti = MULTI ? (STAGED ? ltype[ui] : a.ts[live ? s : h0]) : 0;
        const T qi = rec_q<T, STAGED>(pi.w) * (T)ps.qqr2e;
        const uint32_t rof = (uint32_t)(live ? s - h0 : 0) * (uint32_t)cap + (uint32_t)gl;
        double ax = 0, ay = 0, az = 0;
        auto rounds = [&](auto fast_tag) {
            constexpr bool FAST = decltype(fast_tag)::value;
            // row indices of round base: tile-local, the sentinel past the end
            // (FAST), -1 past the end otherwise
            // (32-bit byte offsets from the workgroup-uniform row base: one
            // address operation per load)
            auto load_u = [&](int base, int* u) {
#pragma unroll
                for (int k = 0; k < kRound; ++k) {
                    const int v = base + kGroup * k;
                    const uint32_t e = FAST ? rof + (uint32_t)min(v, tl - gl) : rof + (uint32_t)v;
                    if (FAST)
                        u[k] = (int)*(const uint16_t*)(rowb + 2u * e);
                    else
                        u[k] = v + gl < tl ? (int)*(const uint16_t*)(rowb + 2u * e) : -1;
                }
            };
            int un[kRound];
            load_u(0, un);
            for (int base = 0; base < tmax; base += kGroup * kRound) {
                // software pipeline: the next round's index loads are in flight
                // during this round's gathers and maths
                int u[kRound];
#pragma unroll
                for (int k = 0; k < kRound; ++k) u[k] = un[k];
                if (base + kGroup * kRound < tmax) load_u(base + kGroup * kRound, un);
                double4 pj[kRound];
                int tj[kRound] = {};
                bool sel[kRound] = {};
#pragma unroll
                for (int k = 0; k < kRound; ++k) {
                    if (RBL) {
                        // random batch rows: selected shell entries carry kRblSel
                        sel[k] = u[k] >= 0 && (u[k] & kRblSel) != 0;
                        if (u[k] >= 0) u[k] &= kMaxTile;
                    }
                    const int uu = FAST ? u[k] : (u[k] < 0 ? 0 : u[k]);
                    pj[k] = tile_pos<STAGED>(h, lds, a.xs, uu);
                    if (MULTI) tj[k] = tile_type<STAGED>(h, ltype, a.ts, uu);
                }
#pragma unroll
                for (int k = 0; k < kRound; ++k) {
                    if (base + kGroup * k < tmax)
                        accumulate_pair<T, EV, MULTI, RBL>(
                            base + kGroup * k < nin_max, pj[k].x - pi.x, pj[k].y - pi.y,
                            pj[k].z - pi.z, qi * rec_q<T, STAGED>(pj[k].w), ti, tj[k],
                            FAST || u[k] >= 0, ps, pp,
                            ax, ay, az, acc, rbl.w, sel[k]);
                }
            }
        };
        if (fast)
            rounds(std::true_type{});
        else
            rounds(std::false_type{});
        if (__ballot(live && over) != 0ull) {
            const int32_t id_i = RBL && live ? a.perm[s] : 0;
            for (int uu = gl; uu < total; uu += kGroup) {
                const double4 pj = tile_pos<STAGED>(h, lds, a.xs, uu);
                const int tj = MULTI ? tile_type<STAGED>(h, ltype, a.ts, uu) : 0;
                bool sel = false;
                if (RBL && live && over)
                    sel = rbl_select(rbl.seed, epoch, id_i, tile_atom_id(h, a.perm, uu),
                                     rbl.pice);
                accumulate_pair<T, EV, MULTI, RBL>(true, pj.x - pi.x, pj.y - pi.y, pj.z - pi.z,
                                                   qi * rec_q<T, STAGED>(pj.w), ti, tj,
                                                   live && over, ps, pp, ax,
                                                   ay, az, acc, rbl.w, sel);
            }
        }
        ax = group_sum(ax);
        ay = group_sum(ay);
        az = group_sum(az);
        if (live && gl == kSumLane) {
            const int i = a.perm[s];
            fx[i] = (T)ax;
            fy[i] = (T)ay;
            fz[i] = (T)az;
            if (RBL) {
                // the stored forces, summed for the mean-force correction
                fsum[0] += (double)(T)ax;
                fsum[1] += (double)(T)ay;
                fsum[2] += (double)(T)az;
            }
        }
    }
}

// Per-workgroup FP64 energy / virial partial: wave sums, then the 8 waves in order.
__device__ inline void block_ev(const double* acc, int lane, int wv, double* __restrict__ out)
{
    __shared__ double red[kTileWaves][8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double v = wave_sum_dpp(acc[k]);
        if (lane == 0) red[wv][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        double v = 0.0;
        for (int w = 0; w < kTileWaves; ++w) v += red[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// Per-workgroup FP64 sums of the near forces (random batch list): wave sums, then
// the 8 waves in order.
__device__ inline void block_fsum(const double* fs, int lane, int wv, double* __restrict__ out)
{
    __shared__ double red[kTileWaves][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double v = wave_sum_dpp(fs[k]);
        if (lane == 0) red[wv][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0.0;
        for (int w = 0; w < kTileWaves; ++w) v += red[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// Tile pass: one workgroup per (cell tile, part).  On a rebuild step it first
// builds the Verlet rows of its home atoms (list_wave over the tile's reference
// positions), then the pair pass: the tile's current positions are staged in
// LDS shifted to the home cell's image, with the charge (rec_q) and the types
// when there are several;
// each group of kGroup lanes walks one home atom's uint16 row -- LJ-range
// entries first, so later rounds skip the LJ code -- then reduces across its
// lanes.  Tiles larger than the LDS stage take the same steps from global
// memory.  The rows of a home atom are written and read by the same workgroup,
// so the list build and the pair pass need no launch in between.  In MD steps
// the workgroup finally bins its home atoms for the next rebuild; that comes
// last so its atomics' latency overlaps the other workgroups' pair work rather
// than delaying this one's.
template <typename T, bool EV, bool MULTI, bool RBL = false>
struct TilePass {
    static constexpr int kThreads = kTileThreads;
    struct Args {
        PairArgs a;
        const double4* xs_ref;
        const int32_t* cell_start;
        int split;
        int3 dims;
        Box box;
        int stage_cap;
        double rl2, rin2;
        PairScalars ps;
        const PairParams* pp;
        T *fx, *fy, *fz;
        double* ev_part;
        uint32_t* flags;
        unsigned long long* stats;
        uint16_t* nlist;
        int2* ncount;
        int32_t* cell_count;
        int md;        // MD step: bin the home atoms for the next rebuild
        double3 cell;  // cell edge lengths (binning)
        MdBins bins;
        const int32_t* order;  // workgroup -> cell * split + part (null: identity)
        RblArgs rbl;           // random batch list (RBL instantiation)
    };
    __device__ static void run(const Args& A, int blk, int)
    {
        const PairArgs& a = A.a;
        extern __shared__ double4 lds[];
        int32_t* ltype = (int32_t*)(lds + A.stage_cap);
        const FullStage fst{(char*)lds, 16u * (uint32_t)A.stage_cap};
        __shared__ TileHdr h;
        const int tb = A.order ? A.order[blk] : blk;
        const int c = tb / A.split, part = tb % A.split;
        const bool rebuild = A.flags[2] != 0u;
        if (threadIdx.x == 0) {
            // the build kernels of this step have read the request and the cell
            // histogram: clear both for the next step's binning
            if (blk == 0) A.flags[0] = 0u;
            if (part == 0) A.cell_count[c] = 0;
        }
        tile_setup(c, A.dims, A.box, A.cell_start, h);
        const int total = h.off[27];
        const bool staged = total < A.stage_cap;  // room for the sentinel's dummy entry
        const int lane = threadIdx.x & 63;
        const int wv = uniform(threadIdx.x >> 6);
        const int h0 = h.gstart[kHomeSeg], hc = h.cnt[kHomeSeg];
        const int a0 = h0 + (hc * part) / A.split, a1 = h0 + (hc * (part + 1)) / A.split;
        // list build number (the random batch list's selection epoch); the build
        // kernels before this one have counted the current build
        const uint32_t epoch = RBL ? (uint32_t)A.stats[0] : 0u;
        if (rebuild) {
            if (!staged && part == 0 && threadIdx.x == 0) atomicAdd(&A.stats[6], 1ull);
            const RefStage rs(lds, A.stage_cap);
            int32_t* ids = RBL && staged ? stage_ids(lds, A.stage_cap, a.cap) : nullptr;
            if (staged) {
                for (int u = threadIdx.x; u < total; u += kTileThreads) {
                    rs.put(u, stage_ref(h, A.xs_ref, u));
                    if (RBL) ids[u] = tile_atom_id(h, a.perm, u);
                }
                const int pad = ref_stage_len(total);  // the sweep reads whole chunks
                for (int u = total + (int)threadIdx.x; u < pad; u += kTileThreads)
                    rs.put(u, h.dummy);
            }
            __syncthreads();
            uint16_t* hb = hit_buffer(lds, A.stage_cap, a.cap, wv);
            const RblBuild rb{A.rbl.near2, A.rbl.seed, epoch, A.rbl.pice, a.perm, ids};
            if (staged)
                list_wave<true, RBL>(h, rs, A.xs_ref, a0 + wv, a1, h0, lane, A.rl2, A.rin2,
                                     a.cap, hb, A.nlist, A.ncount, A.stats, rb);
            else
                list_wave<false, RBL>(h, rs, A.xs_ref, a0 + wv, a1, h0, lane, A.rl2, A.rin2,
                                      a.cap, hb, A.nlist, A.ncount, A.stats, rb);
            __syncthreads();  // rows and LDS are reused below
        }
        if (staged) {
            // two entries per trip, both loads issued before either LDS write
            for (int u0 = threadIdx.x; u0 < total; u0 += 2 * kTileThreads) {
                int st[2], j[2];
                double4 p[2];
                int t[2] = {0, 0};
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int u = u0 + k * kTileThreads;
                    st[k] = tile_seg(h, u < total ? u : 0);
                    j[k] = h.gstart[st[k]] + ((u < total ? u : 0) - h.off[st[k]]);
                    p[k] = a.xs[j[k]];
                    if (MULTI) t[k] = a.ts[j[k]];
                }
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int u = u0 + k * kTileThreads;
                    if (u < total) {
                        fst.put(u, make_double4(p[k].x - h.sx[st[k]], p[k].y - h.sy[st[k]],
                                                p[k].z - h.sz[st[k]], rho_stage_q<T>(p[k].w)));
                        if (MULTI) ltype[u] = t[k];
                    }
                }
            }
            if (threadIdx.x == 0) {
                fst.put(total, h.dummy);
                if (MULTI) ltype[total] = 0;
            }
        }
        __syncthreads();
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double fsum[3] = {0, 0, 0};
        if (staged)
            pair_wave<T, EV, MULTI, true, RBL>(a, h, fst, ltype, a0, a1, h0, wv, lane, A.ps,
                                               A.pp, A.fx, A.fy, A.fz, acc, A.rbl, epoch, fsum);
        else
            pair_wave<T, EV, MULTI, false, RBL>(a, h, FullStage{}, nullptr, a0, a1, h0, wv, lane,
                                                A.ps, A.pp, A.fx, A.fy, A.fz, acc, A.rbl, epoch,
                                                fsum);
        if (EV) block_ev(acc, lane, wv, A.ev_part + (size_t)tb * 8);
        if (RBL) block_fsum(fsum, lane, wv, A.rbl.part + (size_t)tb * 3);
        if (A.md) {
            // bins of the next rebuild: clear this step's buffer (consumed by the
            // cell sort) for this cell, bin the current home atoms into the other
            const int b = (int)(A.flags[3] & 1u);
            if (part == 0 && threadIdx.x == 0) {
                A.bins.count[b * A.bins.ncell + c] = 0;
                if (A.bins.nchunk > 0 && c % kBinChunk == 0)
                    A.bins.chunk[b * A.bins.nchunk + c / kBinChunk] = 0;
            }
            for (int base = a0; base < a1; base += kTileThreads) {
                const int s = base + (int)threadIdx.x;
                const bool act = s < a1;
                double p[3] = {0.0, 0.0, 0.0};
                int v = 0;
                if (act) {
                    const double4 p4 = a.xs[s];
                    p[0] = p4.x;
                    p[1] = p4.y;
                    p[2] = p4.z;
                    v = a.perm[s];
                }
                md_bin_wave(A.bins, 1 - b, act, s, v, p, A.box, A.dims, A.cell);
            }
        }
    }
};

// ================================================================ half-shell pass
// Position / type of half-tile candidate u (LDS when staged; the row sentinel
// u = total reads the dummy).
template <bool STAGED>
__device__ inline double4 half_pos(const HalfHdr& h, const double4* __restrict__ lds,
                                   const double4* __restrict__ xs, int u)
{
    if (STAGED) return lds[u];
    const int total = h.off[kHalfSegs];
    const int uc = u < total ? u : 0;
    const int st = half_seg(h, uc);
    const double4 p = xs[h.gstart[st] + (uc - h.off[st])];
    return u < total ? make_double4(p.x - h.sx[st], p.y - h.sy[st], p.z - h.sz[st], p.w)
                     : h.dummy;
}

template <bool STAGED>
__device__ inline int half_type(const HalfHdr& h, const int32_t* __restrict__ ltype,
                                const int32_t* __restrict__ ts, int u)
{
    if (STAGED) return ltype[u];
    const int total = h.off[kHalfSegs];
    const int uc = u < total ? u : 0;
    const int st = half_seg(h, uc);
    const int t = ts[h.gstart[st] + (uc - h.off[st])];
    return u < total ? t : 0;
}

// LDS sums of the half pass: the forces on every candidate of the tile.
struct HalfAcc {
    double *x, *y, *z;
};

// One pair of a half row: the FP32 force factor as in the full pass, the force on
// i accumulated in FP64 registers, minus the same on j added to the LDS sums on the
// 2^-30 grid; with EV the pair's whole energy (both directed halves) and virial.
// Lanes on the row sentinel (u = total, past the row's end) and pairs beyond both
// cutoffs add nothing and skip the atomics (the sentinel lanes would all hit one LDS
// address and serialise); the sums are exact, so skipping a zero changes no bit.
template <bool EV, bool MULTI>
__device__ inline void accumulate_half(bool lj, double dx, double dy, double dz, double qq,
                                       int ti, int tj, bool valid, const PairScalars& ps,
                                       const PairParams* __restrict__ pp, double& fx,
                                       double& fy, double& fz, double* acc, const HalfAcc& J,
                                       int u, int total)
{
#pragma clang fp contract(off)
    const double d2 = dx * dx + dy * dy + dz * dz;
    const bool ok = valid && d2 > 1e-4;
    float e_lj = 0, e_re = 0;
    const float qf = (float)qq;
    const float fs = lj ? pair_fs<float, EV, MULTI, true>(d2, qf, ti, tj, ok, ps, pp, &e_lj, &e_re)
                        : pair_fs<float, EV, MULTI, false>(d2, qf, ti, tj, ok, ps, pp, &e_lj,
                                                           &e_re);
    const double f = (double)fs;
    fx = fma(f, dx, fx);
    fy = fma(f, dy, fy);
    fz = fma(f, dz, fz);
    if (valid && u < total && f != 0.0) {
        atomicAdd(&J.x[u], -half_grid(f * dx));
        atomicAdd(&J.y[u], -half_grid(f * dy));
        atomicAdd(&J.z[u], -half_grid(f * dz));
    }
    if (EV && ok) {
        const double h = -f;
        acc[0] += 2.0 * (double)e_lj;
        acc[1] += 2.0 * (double)e_re;
        acc[2] += h * dx * dx;
        acc[3] += h * dy * dy;
        acc[4] += h * dz * dz;
        acc[5] += h * dx * dy;
        acc[6] += h * dx * dz;
        acc[7] += h * dy * dz;
    }
}

// pair_wave over half rows: the same lane groups, rounds and sentinel handling;
// each row's force on its home atom goes to the LDS sums at the end of the row.
// Rows that overflowed scan the candidates after the home atom instead.  Rows are
// dealt to the lane groups in the order ord (home indices by decreasing row length,
// so a wave's four rows have about the same length; null: home order).
template <bool EV, bool MULTI, bool STAGED>
__device__ inline void pair_wave_half(const PairArgs& a, const HalfHdr& h,
                                      const double4* __restrict__ lds,
                                      const int32_t* __restrict__ ltype, int a0, int a1, int h0,
                                      int wv, int lane, const PairScalars& ps,
                                      const PairParams* __restrict__ pp, double* acc,
                                      const HalfAcc& J, const int32_t* __restrict__ ord)
{
    const int total = h.off[kHalfSegs];
    const int cap = a.cap;
    constexpr int kPer = 64 / kGroup;
    const int g = lane / kGroup, gl = lane % kGroup;
    const uint16_t* __restrict__ rows = a.nlist + (size_t)uniform(h0) * cap;
    for (int t0 = kPer * wv; t0 < a1 - a0; t0 += kPer * kTileWaves) {
        const int pos = t0 + g;
        const bool live = pos < a1 - a0;
        const int s = a0 + (live && ord ? ord[pos] : pos);
        const int2 nc = live ? a.ncount[s] : make_int2(0, 0);
        const int n_in = nc.x, tot = nc.x + nc.y;
        const bool over = tot >= cap;
        const int tl = over ? 0 : tot;
        const int tmax = wave_max_int(tl);
        const int nin_max = wave_max_int(over ? 0 : n_in);
        const bool fast = __ballot(!live || over) == 0ull;
        const int ui = live ? s - h0 : 0;
        const double4 pi = STAGED ? lds[ui] : half_pos<false>(h, lds, a.xs, ui);
        const int ti = MULTI ? half_type<STAGED>(h, ltype, a.ts, ui) : 0;
        const double qi = pi.w * ps.qqr2e;
        const uint32_t rof = (uint32_t)ui * (uint32_t)cap + (uint32_t)gl;
        double ax = 0, ay = 0, az = 0;
        auto rounds = [&](auto fast_tag) {
            constexpr bool FAST = decltype(fast_tag)::value;
            auto load_u = [&](int base, int* u) {
#pragma unroll
                for (int k = 0; k < kRound; ++k) {
                    const int v = base + kGroup * k;
                    if (FAST)
                        u[k] = (int)rows[rof + (uint32_t)min(v, tl - gl)];
                    else
                        u[k] = v + gl < tl ? (int)rows[rof + (uint32_t)v] : -1;
                }
            };
            int un[kRound];
            load_u(0, un);
            for (int base = 0; base < tmax; base += kGroup * kRound) {
                int u[kRound];
#pragma unroll
                for (int k = 0; k < kRound; ++k) u[k] = un[k];
                if (base + kGroup * kRound < tmax) load_u(base + kGroup * kRound, un);
                double4 pj[kRound];
                int tj[kRound] = {};
#pragma unroll
                for (int k = 0; k < kRound; ++k) {
                    const int uu = FAST ? u[k] : (u[k] < 0 ? total : u[k]);
                    pj[k] = half_pos<STAGED>(h, lds, a.xs, uu);
                    if (MULTI) tj[k] = half_type<STAGED>(h, ltype, a.ts, uu);
                }
#pragma unroll
                for (int k = 0; k < kRound; ++k) {
                    if (base + kGroup * k < tmax) {
                        const int uu = FAST ? u[k] : (u[k] < 0 ? total : u[k]);
                        accumulate_half<EV, MULTI>(base + kGroup * k < nin_max, pj[k].x - pi.x,
                                                   pj[k].y - pi.y, pj[k].z - pi.z, qi * pj[k].w,
                                                   ti, tj[k], FAST || u[k] >= 0, ps, pp, ax, ay,
                                                   az, acc, J, uu, total);
                    }
                }
            }
        };
        if (fast)
            rounds(std::true_type{});
        else
            rounds(std::false_type{});
        if (__ballot(live && over) != 0ull) {
            for (int uu = ui + 1 + gl; uu < total + kGroup; uu += kGroup) {
                const int uc = uu < total ? uu : total;
                const double4 pj = half_pos<STAGED>(h, lds, a.xs, uc);
                const int tj = MULTI ? half_type<STAGED>(h, ltype, a.ts, uc) : 0;
                accumulate_half<EV, MULTI>(true, pj.x - pi.x, pj.y - pi.y, pj.z - pi.z,
                                           qi * pj.w, ti, tj, live && over && uu < total, ps, pp,
                                           ax, ay, az, acc, J, uc, total);
            }
        }
        ax = group_sum(ax);
        ay = group_sum(ay);
        az = group_sum(az);
        if (live && gl == kSumLane) {
            atomicAdd(&J.x[ui], half_grid(ax));
            atomicAdd(&J.y[ui], half_grid(ay));
            atomicAdd(&J.z[ui], half_grid(az));
        }
    }
}

// Largest cell whose home rows the half pass orders by length (LDS ints).
constexpr int kHalfOrderMax = 256;

// Force planes of the half pass: [kHalfSegs][3][stride] floats, plane k holding the
// sums a tile made for the atoms of its segment k, at their cell-sorted slots.
struct HalfPlanes {
    float* p;
    size_t stride;
};

// Tile pass over half tiles (FP32): one workgroup per home cell (never split).  On
// a rebuild step the half rows are built first (list_wave_half), then the tile is
// staged, every pair of it evaluated once, and the LDS sums written to the planes.
// acc_cap bounds the tiles the LDS sums hold; a larger tile sets stats[7] and is
// skipped (the host then reports the error and returns the handle to full rows).
template <bool EV, bool MULTI>
struct TilePassHalf {
    static constexpr int kThreads = kTileThreads;
    struct Args {
        typename TilePass<float, EV, MULTI>::Args t;
        HalfPlanes planes;
        int acc_cap;
    };
    __device__ static void run(const Args& AH, int blk, int)
    {
        const auto& A = AH.t;
        const PairArgs& a = A.a;
        extern __shared__ double4 lds[];
        double* jbase = (double*)(lds + A.stage_cap);
        const HalfAcc J{jbase, jbase + AH.acc_cap, jbase + 2 * (size_t)AH.acc_cap};
        int32_t* ltype = (int32_t*)(jbase + 3 * (size_t)AH.acc_cap);
        __shared__ HalfHdr h;
        const int c = A.order ? A.order[blk] : blk;
        const bool rebuild = A.flags[2] != 0u;
        if (threadIdx.x == 0) {
            if (blk == 0) A.flags[0] = 0u;
            A.cell_count[c] = 0;
        }
        half_setup(c, A.dims, A.box, A.cell_start, h);
        const int total = h.off[kHalfSegs];
        const bool staged = total < A.stage_cap;
        const int lane = threadIdx.x & 63;
        const int wv = uniform(threadIdx.x >> 6);
        const int h0 = h.gstart[0], hc = h.cnt[0];
        const int a0 = h0, a1 = h0 + hc;
        if (total + 1 > AH.acc_cap) {
            if (threadIdx.x == 0) A.stats[7] = 1ull;
            return;
        }
        if (rebuild) {
            if (!staged && threadIdx.x == 0) atomicAdd(&A.stats[6], 1ull);
            const RefStage rs(lds, A.stage_cap);
            if (staged) {
                for (int u = threadIdx.x; u < total; u += kTileThreads)
                    rs.put(u, half_ref(h, A.xs_ref, u));
                const int pad = ref_stage_len(total);
                for (int u = total + (int)threadIdx.x; u < pad; u += kTileThreads)
                    rs.put(u, h.dummy);
            }
            __syncthreads();
            uint16_t* hb = hit_buffer(lds, A.stage_cap, a.cap, wv);
            if (staged)
                list_wave_half<true>(h, rs, A.xs_ref, a0 + wv, a1, h0, lane, A.rl2, A.rin2,
                                     a.cap, hb, A.nlist, A.ncount, A.stats);
            else
                list_wave_half<false>(h, rs, A.xs_ref, a0 + wv, a1, h0, lane, A.rl2, A.rin2,
                                      a.cap, hb, A.nlist, A.ncount, A.stats);
            __syncthreads();
        }
        if (staged) {
            for (int u = threadIdx.x; u < total; u += kTileThreads) {
                const int st = half_seg(h, u);
                const int j = h.gstart[st] + (u - h.off[st]);
                const double4 p = a.xs[j];
                lds[u] = make_double4(p.x - h.sx[st], p.y - h.sy[st], p.z - h.sz[st], p.w);
                if (MULTI) ltype[u] = a.ts[j];
            }
            if (threadIdx.x == 0) {
                lds[total] = h.dummy;
                if (MULTI) ltype[total] = 0;
            }
        }
        for (int u = threadIdx.x; u <= total; u += kTileThreads) {
            J.x[u] = 0.0;
            J.y[u] = 0.0;
            J.z[u] = 0.0;
        }
        // home atoms by decreasing half-row length (ties by index; an overflowed row
        // counts as longest), for cells of at most kHalfOrderMax atoms
        __shared__ int32_t ord[kHalfOrderMax], len[kHalfOrderMax];
        const bool sorted = hc <= kHalfOrderMax;
        if (sorted) {
            for (int t = threadIdx.x; t < hc; t += kTileThreads) {
                const int2 nc = a.ncount[a0 + t];
                len[t] = nc.x + nc.y;
            }
        }
        __syncthreads();
        if (sorted) {
            for (int t = threadIdx.x; t < hc; t += kTileThreads) {
                const int lt = len[t];
                int r = 0;
                for (int j = 0; j < hc; ++j) {
                    const int lj = len[j];
                    r += (lj > lt || (lj == lt && j < t)) ? 1 : 0;
                }
                ord[r] = t;
            }
        }
        __syncthreads();
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int32_t* po = sorted ? ord : nullptr;
        if (staged)
            pair_wave_half<EV, MULTI, true>(a, h, lds, ltype, a0, a1, h0, wv, lane, A.ps, A.pp,
                                            acc, J, po);
        else
            pair_wave_half<EV, MULTI, false>(a, h, nullptr, nullptr, a0, a1, h0, wv, lane, A.ps,
                                             A.pp, acc, J, po);
        __syncthreads();
        for (int u = threadIdx.x; u < total; u += kTileThreads) {
            const int st = half_seg(h, u);
            const size_t slot = (size_t)h.gstart[st] + (size_t)(u - h.off[st]);
            float* pl = AH.planes.p + (size_t)st * 3 * AH.planes.stride + slot;
            pl[0] = (float)J.x[u];
            pl[AH.planes.stride] = (float)J.y[u];
            pl[2 * AH.planes.stride] = (float)J.z[u];
        }
        if (EV) block_ev(acc, lane, wv, A.ev_part + (size_t)c * 8);
        if (A.md) {
            const int b = (int)(A.flags[3] & 1u);
            if (threadIdx.x == 0) {
                A.bins.count[b * A.bins.ncell + c] = 0;
                if (A.bins.nchunk > 0 && c % kBinChunk == 0)
                    A.bins.chunk[b * A.bins.nchunk + c / kBinChunk] = 0;
            }
            for (int base = a0; base < a1; base += kTileThreads) {
                const int s = base + (int)threadIdx.x;
                const bool act = s < a1;
                double p[3] = {0.0, 0.0, 0.0};
                int v = 0;
                if (act) {
                    const double4 p4 = a.xs[s];
                    p[0] = p4.x;
                    p[1] = p4.y;
                    p[2] = p4.z;
                    v = a.perm[s];
                }
                md_bin_wave(A.bins, 1 - b, act, s, v, p, A.box, A.dims, A.cell);
            }
        }
    }
};

// Per-atom sum of the 14 force planes in a fixed order (FP64), written to the
// force arrays in original atom order.
struct PlaneSum {
    static constexpr int kThreads = 256;
    struct Args {
        HalfPlanes planes;
        int n;
        const int32_t* perm;
        float *fx, *fy, *fz;
    };
    __device__ static void run(const Args& a, int blk, int)
    {
        const int s = blk * kThreads + (int)threadIdx.x;
        if (s >= a.n) return;
        const size_t st = a.planes.stride;
        double f[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < kHalfSegs; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) f[c] += (double)a.planes.p[((size_t)k * 3 + c) * st + s];
        const int i = a.perm[s];
        a.fx[i] = (float)f[0];
        a.fy[i] = (float)f[1];
        a.fz[i] = (float)f[2];
    }
};

// MD step over half tiles: TileRho with the half tile pass.
template <bool M>
struct TileRhoHalf {
    static constexpr int kThreads = kTileThreads;
    struct Args {
        typename RhoFinalT<float, kTileThreads>::Args fin;
        int nfin, ntile;
        int fin_first;
        typename TilePassHalf<false, M>::Args tile;
    };
    __device__ static void run(const Args& A, int blk, int)
    {
        const int t = A.fin_first ? blk - A.nfin : blk;
        if (t < 0 || t >= A.ntile) {
            RhoFinalT<float, kTileThreads>::run(A.fin, A.fin_first ? blk : blk - A.ntile, A.nfin);
            return;
        }
        TilePassHalf<false, M>::run(A.tile, t, A.ntile);
    }
};

// MD step: the tile pass and the rho(k) final reduction (nfin workgroups, one
// k per wave) in one grid.  Both need only the previous kernel's output
// (cell-sorted positions, rho partials), and the k-space force that follows
// needs both.  With the load-balanced tile order the tiles come first, so that
// workgroup t runs on CU t mod ncu as that order assumes, and the short
// reductions fill the slots that CUs with fewer tiles leave free; with many
// dispatch rounds the reductions come first and finish early.
template <typename T, bool EV, bool M, bool RBL = false>
struct TileRho {
    static constexpr int kThreads = kTileThreads;
    struct Args {
        typename RhoFinalT<T, kTileThreads>::Args fin;
        int nfin, ntile;
        int fin_first;
        typename TilePass<T, EV, M, RBL>::Args tile;
    };
    __device__ static void run(const Args& A, int blk, int)
    {
        const int t = A.fin_first ? blk - A.nfin : blk;
        if (t < 0 || t >= A.ntile) {
            RhoFinalT<T, kTileThreads>::run(A.fin, A.fin_first ? blk : blk - A.ntile, A.nfin);
            return;
        }
        TilePass<T, EV, M, RBL>::run(A.tile, t, A.ntile);
    }
};

// Mean near force of a random batch list pass: the per-tile sums in a fixed order,
// divided by n (ExecutionMD::ComputeRBLNearForce's corr_value,
// rbmd/framework/run/ExecutionMD.cpp:564-565); the k-space / kick kernel that follows
// subtracts it from every atom's near force (SumRBLCorrForceWorklet,
// rbmd/framework/run/worklet/RunWorklet.cpp:1304-1321).
struct RblMean {
    static constexpr int kThreads = 256;
    struct Args {
        const double* part;
        int ntile;
        double inv_n;
        double* mean;
    };
    __device__ static void run(const Args& a, int, int)
    {
        __shared__ double red[kThreads / 64][3];
        const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
        double v[3] = {0, 0, 0};
        for (int b = t; b < a.ntile; b += kThreads)
            for (int c = 0; c < 3; ++c) v[c] += a.part[(size_t)b * 3 + c];
        for (int c = 0; c < 3; ++c) {
            const double w = wave_sum_dpp(v[c]);
            if (lane == 0) red[wv][c] = w;
        }
        __syncthreads();
        if (t < 3) {
            double sum = 0.0;
            for (int w = 0; w < kThreads / 64; ++w) sum += red[w][t];
            a.mean[t] = sum * a.inv_n;
        }
    }
};

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

PairScalars pair_scalars(const Engine& e)
{
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
    ps.nal2e = (float)(-e.pp.alpha * 1.4426950408889634);
    ps.pa = (float)(kErfcP * e.pp.sqrt_alpha);
    return ps;
}

template <typename T, bool EV, bool M, bool RBL = false>
typename TilePass<T, EV, M, RBL>::Args tile_args(Engine& e, void* fx, void* fy, void* fz,
                                                 bool md)
{
    RblArgs rbl{};
    if (e.rbl) {
        const double rc = std::max(e.rc_lj, e.rc_coul), rn = e.r_core + e.skin;
        rbl.w = RblPair{e.r_core * e.r_core, rc * rc, (float)e.rbl_pice};
        rbl.seed = e.rbl_seed;
        rbl.pice = (uint32_t)e.rbl_pice;
        rbl.near2 = rn * rn;
        rbl.part = e.rbl_part;
    }
    return typename TilePass<T, EV, M, RBL>::Args{
        PairArgs{e.xs, e.ts, e.perm, e.nlist, e.ncount, e.cap},
        e.xs_ref, e.cell_start, e.split, make_int3(e.dims[0], e.dims[1], e.dims[2]), e.box,
        e.stage_cap, e.r_list * e.r_list, e.r_in * e.r_in, pair_scalars(e), e.d_pp, (T*)fx,
        (T*)fy, (T*)fz, e.ev_part, e.flags, e.stats, e.nlist, e.ncount, e.cell_count,
        md ? 1 : 0, make_double3(e.cell[0], e.cell[1], e.cell[2]), md_bins(e),
        tile_order_ncu(e.ncell, e.split, e.num_cu) > 0 ? e.tile_order : nullptr, rbl};
}

size_t pair_lds(const Engine& e)
{
    const size_t stage = (size_t)e.stage_cap * (sizeof(double4) + (e.pp.ntypes > 1 ? 4 : 0));
    const size_t build = list_lds_bytes(e.stage_cap, e.cap, e.rbl);
    return stage > build ? stage : build;
}

// LDS of a half tile pass: the staged positions (and types) and the force sums, or
// the list build's stage and hit buffers, whichever is larger.
size_t half_lds(const Engine& e)
{
    const size_t pass = (size_t)e.stage_cap * sizeof(double4) + (size_t)e.acc_cap * 24 +
                        (e.pp.ntypes > 1 ? (size_t)e.stage_cap * 4 : 0);
    const size_t build = list_lds_bytes(e.stage_cap, e.cap);
    return pass > build ? pass : build;
}

template <bool EV, bool M>
typename TilePassHalf<EV, M>::Args half_args(Engine& e, bool md)
{
    typename TilePassHalf<EV, M>::Args a;
    a.t = tile_args<float, EV, M, false>(e, nullptr, nullptr, nullptr, md);
    a.planes = HalfPlanes{e.planes, (size_t)std::max<int64_t>(e.n_max, 1)};
    a.acc_cap = e.acc_cap;
    return a;
}

void plane_sum_launch(Engine& e, void* fx, void* fy, void* fz, hipStream_t s)
{
    k_one<PlaneSum><<<(int)((e.n + PlaneSum::kThreads - 1) / PlaneSum::kThreads),
                      PlaneSum::kThreads, 0, s>>>(PlaneSum::Args{
        HalfPlanes{e.planes, (size_t)std::max<int64_t>(e.n_max, 1)}, (int)e.n, e.perm,
        (float*)fx, (float*)fy, (float*)fz});
}

template <bool EV, bool M>
void half_launch_t(Engine& e, void* fx, void* fy, void* fz, hipStream_t s)
{
    k_one<TilePassHalf<EV, M>><<<e.ncell, kTileThreads, half_lds(e), s>>>(half_args<EV, M>(e, false));
    plane_sum_launch(e, fx, fy, fz, s);
}

template <bool M>
void half_rho_launch_t(Engine& e, void* fx, void* fy, void* fz, hipStream_t s)
{
    typename TileRhoHalf<M>::Args a;
    const bool ks = has_kspace(e);
    if (ks) a.fin = rho_final_args<float>(e, e.capturing ? e.d_step : nullptr);
    else memset(&a.fin, 0, sizeof a.fin);
    a.nfin = ks ? rho_final_blocks<float, kTileThreads>(e.P) : 0;
    a.ntile = e.ncell;
    a.fin_first = tile_order_ncu(e.ncell, 1, e.num_cu) == 0;
    a.tile = half_args<false, M>(e, true);
    k_one<TileRhoHalf<M>><<<a.nfin + e.ncell, kTileThreads, half_lds(e), s>>>(a);
    plane_sum_launch(e, fx, fy, fz, s);
}

template <typename T, bool EV, bool M, bool RBL>
void pair_launch_t(Engine& e, void* fx, void* fy, void* fz, hipStream_t s)
{
    k_one<TilePass<T, EV, M, RBL>><<<e.ncell * e.split, kTileThreads, pair_lds(e), s>>>(
        tile_args<T, EV, M, RBL>(e, fx, fy, fz, false));
}

template <typename T, bool M, bool RBL = false>
typename TileRho<T, false, M, RBL>::Args tile_rho_args(Engine& e, void* fx, void* fy, void* fz,
                                                        unsigned long long* step_dev)
{
    typename TileRho<T, false, M, RBL>::Args a;
    const bool ks = has_kspace(e);
    if (ks) a.fin = rho_final_args<T>(e, step_dev);
    else memset(&a.fin, 0, sizeof a.fin);
    a.nfin = ks ? rho_final_blocks<T, kTileThreads>(e.P) : 0;
    a.ntile = e.ncell * e.split;
    a.fin_first = tile_order_ncu(e.ncell, e.split, e.num_cu) == 0;
    a.tile = tile_args<T, false, M, RBL>(e, fx, fy, fz, true);
    return a;
}

template <typename T, bool M, bool RBL>
void tile_rho_launch_t(Engine& e, void* fx, void* fy, void* fz, hipStream_t s)
{
    const auto a = tile_rho_args<T, M, RBL>(e, fx, fy, fz, e.capturing ? e.d_step : nullptr);
    k_one<TileRho<T, false, M, RBL>><<<a.nfin + e.ncell * e.split, kTileThreads, pair_lds(e),
                                       s>>>(a);
}

// Tile workgroups of a replica that fit one CU at once with its LDS stage.
inline int tile_wgs_per_cu(const Engine& e)
{
    const int per = (int)(pair_lds(e) + 2048);
    return std::max(1, std::min(Batch::kPairClasses, e.lds_cu / per));
}

// The batched tile pass: one launch per class of replicas that fit the same
// number of tile workgroups per CU, with that class's largest stage.  A single
// launch sized for the largest stage (small boxes, whose three-cell stencil spans
// the whole replica) would hold every replica to one workgroup per CU.
template <typename T, bool M>
hipError_t pair_prepare_t(Batch& b)
{
    const size_t R = b.es.size();
    std::vector<typename TileRho<T, false, M>::Args> ta(R);
    std::vector<int> wg(R);
    for (size_t r = 0; r < R; ++r) {
        Engine& e = *b.es[r];
        ta[r] = tile_rho_args<T, M>(e, b.st[r].fx, b.st[r].fy, b.st[r].fz, e.d_step);
        wg[r] = tile_wgs_per_cu(e);
    }
    // classes in order of decreasing occupancy; phase c holds class c's replicas
    int slot = 0;
    for (int w = Batch::kPairClasses; w >= 1; --w) {
        std::vector<int> nt(R, 0);
        size_t lds = 0;
        bool any = false;
        for (size_t r = 0; r < R; ++r) {
            if (wg[r] != w) continue;
            Engine& e = *b.es[r];
            nt[r] = ta[r].nfin + e.ncell * e.split;
            lds = std::max(lds, pair_lds(e));
            any = true;
        }
        if (!any) continue;
        BatchPhase& ph = slot == 0 ? b.ph[BP_PAIR] : b.pair_more[slot - 1];
        const hipError_t err = phase_upload(ph, ta, nt, lds);
        if (err != hipSuccess) return err;
        ++slot;
    }
    // classes no longer present launch nothing
    for (; slot < Batch::kPairClasses; ++slot) {
        BatchPhase& ph = slot == 0 ? b.ph[BP_PAIR] : b.pair_more[slot - 1];
        const hipError_t err = phase_upload(ph, ta, std::vector<int>(R, 0), 0);
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

}  // namespace

void launch_pair(Engine& e, void* fx, void* fy, void* fz, bool ev, hipStream_t s)
{
    if (e.n == 0) return;
    ScopedTimer t(e, P_PAIR, s);
    const bool multi = e.pp.ntypes > 1;
    if (e.half) {
        // the half-shell pass and the per-atom sum of its force planes
        if (ev) {
            if (multi) half_launch_t<true, true>(e, fx, fy, fz, s);
            else half_launch_t<true, false>(e, fx, fy, fz, s);
        } else {
            if (multi) half_launch_t<false, true>(e, fx, fy, fz, s);
            else half_launch_t<false, false>(e, fx, fy, fz, s);
        }
        return;
    }
#define RBMD_PAIR_M(T, EV)                                                                    \
    do {                                                                                      \
        if (e.rbl) {                                                                          \
            if (multi) pair_launch_t<T, EV, true, true>(e, fx, fy, fz, s);                    \
            else pair_launch_t<T, EV, false, true>(e, fx, fy, fz, s);                         \
        } else {                                                                              \
            if (multi) pair_launch_t<T, EV, true, false>(e, fx, fy, fz, s);                   \
            else pair_launch_t<T, EV, false, false>(e, fx, fy, fz, s);                        \
        }                                                                                     \
    } while (0)
    if (e.precision == 32) {
        if (ev) RBMD_PAIR_M(float, true);
        else RBMD_PAIR_M(float, false);
    } else {
        if (ev) RBMD_PAIR_M(double, true);
        else RBMD_PAIR_M(double, false);
    }
#undef RBMD_PAIR_M
}

void launch_tile_rho(Engine& e, void* fx, void* fy, void* fz, hipStream_t s)
{
    if (e.n == 0) return;
    ScopedTimer t(e, P_PAIR, s);
    const bool multi = e.pp.ntypes > 1;
    if (e.half) {
        if (multi) half_rho_launch_t<true>(e, fx, fy, fz, s);
        else half_rho_launch_t<false>(e, fx, fy, fz, s);
        return;
    }
#define RBMD_TR(T, M)                                                                         \
    do {                                                                                      \
        if (e.rbl) tile_rho_launch_t<T, M, true>(e, fx, fy, fz, s);                           \
        else tile_rho_launch_t<T, M, false>(e, fx, fy, fz, s);                                \
    } while (0)
    if (e.precision == 32) {
        if (multi) RBMD_TR(float, true);
        else RBMD_TR(float, false);
    } else {
        if (multi) RBMD_TR(double, true);
        else RBMD_TR(double, false);
    }
#undef RBMD_TR
}

void launch_rbl_mean(Engine& e, hipStream_t s)
{
    if (!e.rbl || e.n == 0) return;
    k_one<RblMean><<<1, RblMean::kThreads, 0, s>>>(
        RblMean::Args{e.rbl_part, e.ncell * e.split, 1.0 / (double)e.n, e.rbl_mean});
}

void launch_ev_final(Engine& e, bool with_pair, hipStream_t s)
{
    ScopedTimer t(e, P_EV_FINAL, s);
    const int nb = with_pair ? e.ncell * e.split : 0;  // one partial per tile workgroup
                                                        // (split 1 for half tiles)
    const int P = (e.kspace != 0) ? e.P : 0;
    k_ev_final<<<1, 256, 0, s>>>(e.ev_part, nb, e.kev, P, e.e_self, e.ev);
}

hipError_t pair_batch_prepare(Batch& b)
{
    if (b.precision == 32) return b.multi ? pair_prepare_t<float, true>(b) : pair_prepare_t<float, false>(b);
    return b.multi ? pair_prepare_t<double, true>(b) : pair_prepare_t<double, false>(b);
}

void pair_batch_launch(Batch& b, int phase, hipStream_t s)
{
#define RBMD_PB(T, M)                                                                         \
    do {                                                                                      \
        if (phase == BP_PAIR) {                                                               \
            phase_launch<TileRho<T, false, M>>(b.ph[BP_PAIR], s);                           \
            for (const BatchPhase& ph : b.pair_more) phase_launch<TileRho<T, false, M>>(ph, s); \
        }                                                                                     \
    } while (0)
    if (b.precision == 32) {
        if (b.multi) RBMD_PB(float, true);
        else RBMD_PB(float, false);
    } else {
        if (b.multi) RBMD_PB(double, true);
        else RBMD_PB(double, false);
    }
#undef RBMD_PB
}

}  // namespace rbmd
