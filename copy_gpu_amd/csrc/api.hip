// api.hip -- C ABI of the MI355X charged-LJ force engine (declared in include/rbmd_force.h).
//
// Host orchestration of one force evaluation (the work of Force::Execute,
// rbmd-gpu src/force/src/lj_cut_coul_kspace_force.cpp:73-78, and of
// ExecutionNVT::ComputeAllForce, rbmd/framework/run/ExecutionNVT.cpp:185-222):
//   check -> [bin, scan, scatter, cellsort, list]* -> gather -> [rho, rho_final]
//   -> pair -> [kforce] -> [ev_final]
// (* the build kernels run every call but exit at once unless the displacement
// check or the host asked for a rebuild, so a step needs no host round trip and
// can be captured in a HIP graph).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "../../include/rbmd_force.h"
#include "half.h"
#include "kspace.h"
#include "proposal.h"
#include "state.h"
#include "tile.h"

namespace rbmd {

void build_kcdf(double alpha, const double* L, std::vector<double>& cdf, int M[3]);
double rbe_S(double alpha, const double* L);

// An MD run in which some atom moved more than half the Verlet skin in one step: a
// rebuild references the positions binned one step earlier, so such a step can miss
// pairs (flags[4], set by the drift).
static const char* kStepTooLong =
    "an atom moved more than half the Verlet skin in one step: the list rebuilds may "
    "have missed pairs; use a larger skin or a smaller dt";

static const char* kPassNames[P_NUM] = {"check", "bin",  "scan",   "scatter",   "cellsort",
                                        "list",  "gather", "rho",  "rho_final", "pair",
                                        "kforce", "kick", "ev_final", "sample"};

void timer_begin(Engine& e, int pass, hipStream_t s, hipEvent_t* ev)
{
    *ev = nullptr;
    if (!e.timing || e.capturing) return;
    hipEvent_t a;
    if (hipEventCreate(&a) != hipSuccess) return;
    hipEventRecord(a, s);
    *ev = a;
}

void timer_end(Engine& e, int pass, hipStream_t s, hipEvent_t a)
{
    if (!a) return;
    hipEvent_t b;
    if (hipEventCreate(&b) != hipSuccess) {
        hipEventDestroy(a);
        return;
    }
    hipEventRecord(b, s);
    e.pending[pass].push_back({a, b});
}

static void timer_drain(Engine& e)
{
    for (int p = 0; p < P_NUM; ++p) {
        for (auto& r : e.pending[p]) {
            float ms = 0.f;
            hipEventSynchronize(r.b);
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                e.t_sum[p] += ms;
                e.t_cnt[p] += 1;
            }
            hipEventDestroy(r.a);
            hipEventDestroy(r.b);
        }
        e.pending[p].clear();
    }
}

}  // namespace rbmd

using namespace rbmd;

struct rbmd_force_s {
    Engine e;
};

static thread_local std::string g_create_err;

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _st = (expr);                                                         \
        if (_st != hipSuccess) {                                                         \
            char _b[256];                                                                \
            snprintf(_b, sizeof _b, "%s failed: %s", #expr, hipGetErrorString(_st));     \
            E.err = _b;                                                                  \
            return RBMD_ERR_HIP;                                                         \
        }                                                                                \
    } while (0)

static int fail(Engine& E, int code, const char* msg)
{
    E.err = msg;
    return code;
}

template <typename P>
static hipError_t dalloc(P** p, size_t bytes)
{
    return hipMalloc((void**)p, bytes ? bytes : 16);
}

static void free_all(Engine& E)
{
    void* ptrs[] = {E.xr,       E.yr,    E.zr,     E.xs_ref, E.xs,        E.perm,     E.scell,
                    E.ts,       E.cell_count, E.cell_start, E.atom_cell, E.atom_rank, E.nlist,
                    E.tile_order,
                    E.md_count, E.md_bucket, E.md_cell, E.md_vid, E.md_w,
                    E.ncount,   E.stats, E.q,      E.type,      E.d_pp,     E.km,
                    E.kw,       E.kpart, E.rho,    E.kcoef,  E.kev,       E.kcdf,     E.d_step,
                    E.ev_part,  E.ev,    E.gp_cdf,  E.gp_ratio, E.gp_guide, E.thermo_part,
                    E.thermo_st, E.rbl_part, E.rbl_mean, E.planes};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (E.h_stats) hipHostFree(E.h_stats);
    for (auto& G : E.graphs) {
        if (G.exec) hipGraphExecDestroy(G.exec);
        if (G.graph) hipGraphDestroy(G.graph);
    }
    if (E.own_stream) hipStreamDestroy(E.own_stream);
    if (E.order_event) hipEventDestroy(E.order_event);
}

// The list build keeps the tile's reference stage and one hit buffer of cap
// slots per wave in LDS: shrink the stage until both fit a workgroup (larger
// tiles take the global-memory path).  False if even the smallest stage does not.
static bool fit_stage(Engine& E)
{
    const size_t lim = (size_t)E.lds_max - 4096;
    while (E.stage_cap > 64 && list_lds_bytes(E.stage_cap, E.cap, E.rbl) > lim) E.stage_cap -= 64;
    return list_lds_bytes(E.stage_cap, E.cap, E.rbl) <= lim;
}

static int alloc_list(Engine& E, int cap)
{
    if (E.nlist) hipFree(E.nlist);
    E.nlist = nullptr;
    E.cap = cap;
    if (!fit_stage(E))
        return fail(E, RBMD_ERR_CAPACITY, "Verlet rows of this length do not fit the LDS");
    HIP_TRY(dalloc(&E.nlist, (size_t)E.n_max * (size_t)cap * sizeof(uint16_t)));
    return RBMD_OK;
}

static int alloc_kspace(Engine& E, int P)
{
    // k-vector arrays, coefficients and per-k terms for P slots; the partial
    // structure factors are bounded by the rho geometry: blocks_x * P <=
    // max(2048 * 64, P) for any n and P (at most 2048 / ceil(P/64) atom tiles).
    if (P <= E.kcap && E.kpart) return RBMD_OK;
    void* olds[] = {E.km, E.kw, E.rho, E.kcoef, E.kev, E.kpart};
    for (void* p : olds)
        if (p) hipFree(p);
    E.km = nullptr;
    E.kw = nullptr;
    E.rho = nullptr;
    E.kcoef = nullptr;
    E.kev = nullptr;
    E.kpart = nullptr;
    const int cap = std::max(P, E.kcap);
    E.kcap = cap;
    HIP_TRY(dalloc(&E.km, (size_t)cap * 3 * sizeof(int32_t)));
    HIP_TRY(dalloc(&E.kw, (size_t)cap * sizeof(double)));
    HIP_TRY(dalloc(&E.rho, (size_t)cap * 2 * sizeof(double)));
    HIP_TRY(dalloc(&E.kcoef, (size_t)cap * 64));
    HIP_TRY(dalloc(&E.kev, (size_t)cap * 7 * sizeof(double)));
    E.kpart_cap = 2 * (int64_t)std::max(2048 * 64, cap);
    HIP_TRY(dalloc(&E.kpart, (size_t)E.kpart_cap * sizeof(double)));
    return RBMD_OK;
}

// Half-space Ewald lattice: m with exp(-k^2/4alpha) >= tol, weight 2 exp(-k^2/4alpha).
static void ewald_lattice(const Engine& E, std::vector<int32_t>& m, std::vector<double>& w)
{
    const double kmax2 = -4.0 * E.alpha * log(E.ewald_tol);
    int M[3];
    for (int d = 0; d < 3; ++d) M[d] = (int)floor(sqrt(kmax2) * E.box.L[d] / (2.0 * M_PI)) + 1;
    for (int a = 0; a <= M[0]; ++a)
        for (int b = -M[1]; b <= M[1]; ++b)
            for (int g = -M[2]; g <= M[2]; ++g) {
                if (a == 0 && (b < 0 || (b == 0 && g <= 0))) continue;
                const double kx = 2 * M_PI * a / E.box.L[0], ky = 2 * M_PI * b / E.box.L[1],
                             kz = 2 * M_PI * g / E.box.L[2];
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 > kmax2) continue;
                m.push_back(a);
                m.push_back(b);
                m.push_back(g);
                w.push_back(2.0 * exp(-k2 / (4.0 * E.alpha)));
            }
}

// Cell-tile geometry for n atoms: workgroups per tile (enough tiles to load
// every CU several times) and the LDS stage, 1.25x the mean stencil population
// (the rare larger tile reads its candidates from global memory).
// Identity tile order (workgroup t computes tile t); the next list build
// replaces it with the load-balanced order.  Synchronous, outside capture.
static hipError_t reset_tile_order(Engine& E)
{
    if (!E.tile_order) return hipSuccess;
    std::vector<int32_t> id((size_t)E.ncell * 4);
    for (size_t k = 0; k < id.size(); ++k) id[k] = (int32_t)k;
    return hipMemcpy(E.tile_order, id.data(), id.size() * 4, hipMemcpyHostToDevice);
}

// Half-shell tiles (home cell + 13 forward cells): the LDS holds the staged
// positions (stage_cap entries) and the force sums (acc_cap entries, 24 bytes
// each); the stage has room for a tile 6 standard deviations above the mean, the
// sums for any tile that size or the stage's, whichever is larger, and a tile is
// never split over several workgroups.
static void half_geometry(Engine& E, int64_t n)
{
    const double V = E.box.L[0] * E.box.L[1] * E.box.L[2];
    const double cellv = E.cell[0] * E.cell[1] * E.cell[2];
    const double mean = 14.0 * (double)n / V * cellv;
    const int per = (int)sizeof(double4) + (E.pp.ntypes > 1 ? 4 : 0);
    const int want = ((int)(mean + 6.0 * std::sqrt(mean) + 64.0) + 63) / 64 * 64;
    int cap = want;
    if (E.tune_stage_scale > 0.0) cap = ((int)(E.tune_stage_scale * mean + 64.0) + 63) / 64 * 64;
    if (E.stage_override > 0) cap = E.stage_override;
    E.stage_cap = std::max(64, cap);
    E.acc_cap = std::max(E.stage_cap, want) + 64;
    const int64_t budget = (int64_t)E.lds_max - 4096;
    auto bytes = [&]() {
        return (int64_t)E.stage_cap * per + (int64_t)E.acc_cap * 24;
    };
    while (E.stage_cap > 64 && (bytes() > budget ||
                                (int64_t)list_lds_bytes(E.stage_cap, E.cap) > budget))
        E.stage_cap -= 64;
    E.split = 1;
}

static void tile_geometry(Engine& E, int64_t n)
{
    if (E.half) {
        half_geometry(E, n);
        return;
    }
    const double V = E.box.L[0] * E.box.L[1] * E.box.L[2];
    const double cellv = E.cell[0] * E.cell[1] * E.cell[2];
    const double mean = 27.0 * (double)n / V * cellv;
    const int per = (int)sizeof(double4) + (E.pp.ntypes > 1 ? 4 : 0);
    const int max_entries = (E.lds_max - 4096) / per;
    int cap;
    if (E.tune_stage_scale > 0.0) {
        cap = ((int)(E.tune_stage_scale * mean + 64.0) + 63) / 64 * 64;
    } else {
        // occupancy-aware: room for a tile 5 standard deviations (Poisson) above
        // the mean, then the largest stage that keeps that many workgroups per CU
        const int want = ((int)(mean + 5.0 * std::sqrt(mean) + 64.0) + 63) / 64 * 64;
        const int wg = std::max(1, E.lds_cu / (std::min(want, max_entries) * per + 2048));
        cap = std::max(want, ((E.lds_cu / wg - 2048) / per) / 64 * 64);
    }
    if (E.stage_override > 0) cap = E.stage_override;
    E.stage_cap = std::max(64, std::min(cap, max_entries));
    fit_stage(E);
    // tiles resident at once (LDS-limited); split a tile over several
    // workgroups only when there are too few tiles to fill that
    const int lds_wg = E.stage_cap * per + 2048;
    const int per_cu = std::max(1, std::min(4, E.lds_cu / lds_wg));
    const int resident = std::max(E.num_cu, 1) * per_cu;
    E.split = 1;
    if (4 * E.ncell < 3 * resident)
        E.split = std::min(4, (resident + E.ncell - 1) / E.ncell);
    if (E.tune_split > 0) E.split = std::min(4, E.tune_split);
}

extern "C" {

const char* rbmd_force_last_error(rbmd_force_t h)
{
    if (!h) return g_create_err.c_str();
    return h->e.err.c_str();
}

int rbmd_force_create(const rbmd_force_params* p, int64_t n_max, rbmd_force_t* out)
{
    g_create_err.clear();
    if (!p || !out || n_max < 0 || n_max > (int64_t)INT32_MAX / 2) {
        g_create_err = "invalid arguments to rbmd_force_create";
        return RBMD_ERR_INVALID;
    }
    *out = nullptr;
    if (p->precision != 32 && p->precision != 64) {
        g_create_err = "precision must be 32 or 64";
        return RBMD_ERR_INVALID;
    }
    if (p->ntypes < 1 || p->ntypes > kMaxTypes || !p->eps || !p->sigma) {
        g_create_err = "ntypes must be in [1, 8] with host eps/sigma arrays";
        return RBMD_ERR_INVALID;
    }
    if (!(p->rc_lj > 0) || !(p->rc_coul >= 0) || !(p->skin >= 0) ||
        (p->kspace != RBMD_KSPACE_NONE && !(p->alpha > 0)) || p->kspace < 0 || p->kspace > 2 ||
        (p->kspace == RBMD_KSPACE_RBE && p->p_max < 1)) {
        g_create_err = "invalid cutoff, skin, alpha, kspace or p_max";
        return RBMD_ERR_INVALID;
    }
    for (int d = 0; d < 3; ++d)
        if (!(p->box_len[d] > 0)) {
            g_create_err = "box lengths must be positive";
            return RBMD_ERR_INVALID;
        }
    rbmd_force_t h = new rbmd_force_s();
    Engine& E = h->e;
    for (int d = 0; d < 3; ++d) {
        E.box.lo[d] = p->box_lo[d];
        E.box.L[d] = p->box_len[d];
        E.box.invL[d] = 1.0 / p->box_len[d];
    }
    E.rc_lj = p->rc_lj;
    E.rc_coul = p->rc_coul;
    E.skin = p->skin;
    E.alpha = p->alpha;
    E.qqr2e = p->qqr2e;
    E.precision = p->precision;
    E.kspace = p->kspace;
    E.p_max = p->p_max;
    E.ewald_tol = p->ewald_tol > 0 ? p->ewald_tol : 1e-16;
    E.r_list = std::max(p->rc_lj, p->rc_coul) + p->skin;
    E.r_in = p->rc_lj + p->skin;
    if (p->neighbor_type != RBMD_NEIGHBOR_VERLET && p->neighbor_type != RBMD_NEIGHBOR_RBL) {
        g_create_err = "neighbor_type must be RBMD_NEIGHBOR_VERLET or RBMD_NEIGHBOR_RBL";
        delete h;
        return RBMD_ERR_INVALID;
    }
    if (p->neighbor_type == RBMD_NEIGHBOR_RBL) {
        if (!(p->r_core > 0) || !(p->r_core < std::max(p->rc_lj, p->rc_coul)) ||
            p->rbl_pice < 1 || p->rbl_pice > 65536) {
            g_create_err = "RBL needs 0 < r_core < max(rc_lj, rc_coul) and 1 <= rbl_pice <= 65536";
            delete h;
            return RBMD_ERR_INVALID;
        }
        E.rbl = true;
        E.r_core = p->r_core;
        E.rbl_pice = p->rbl_pice;
        E.rbl_seed = p->rbl_seed;
    }
    E.n_max = n_max;
    E.stage_override = p->tile_stage;
    for (int d = 0; d < 3; ++d) {
        E.dims[d] = (int)(E.box.L[d] / E.r_list);
        E.cell[d] = E.box.L[d] / (E.dims[d] > 0 ? E.dims[d] : 1);
    }
    // cells at least one list radius wide: boxes narrower than three cells take
    // the same 27-cell stencil, whose wrapped cells then repeat with different
    // periodic shifts, so every image within the list radius is a candidate (the
    // reference's locator does the same, ContPointLocator.h:44-53,
    // ExecPointLocator.h:303-319)
    if (E.dims[0] < 1 || E.dims[1] < 1 || E.dims[2] < 1) {
        g_create_err = "box must be at least one list radius (max cutoff + skin) on every axis";
        delete h;
        return RBMD_ERR_BOX;
    }
    E.ncell = E.dims[0] * E.dims[1] * E.dims[2];
    // the half-shell pair pass, when the caller asks for it: FP32, full Verlet list,
    // every neighbouring cell pair distinct (three or more cells per axis)
    E.half = p->half_shell != 0 && p->precision == 32 && !E.rbl && E.dims[0] >= 3 &&
             E.dims[1] >= 3 && E.dims[2] >= 3;
    // pair parameters
    PairParams& pp = E.pp;
    pp.rcl2 = p->rc_lj * p->rc_lj;
    pp.rcc2 = p->rc_coul * p->rc_coul;
    pp.alpha = p->alpha;
    pp.sqrt_alpha = sqrt(p->alpha);
    pp.two_sqrt_alpha_over_pi = 2.0 * sqrt(p->alpha / M_PI);
    pp.qqr2e = p->qqr2e;
    pp.ntypes = p->ntypes;
    for (int a = 0; a < p->ntypes; ++a)
        for (int b = 0; b < p->ntypes; ++b) {
            const double s = 0.5 * (p->sigma[a] + p->sigma[b]);
            pp.eps24[a * kMaxTypes + b] = 24.0 * sqrt(p->eps[a] * p->eps[b]);
            pp.sig2[a * kMaxTypes + b] = s * s;
        }
    if (p->ntypes == 1) {
        pp.eps24[0] = 24.0 * p->eps[0];
        pp.sig2[0] = p->sigma[0] * p->sigma[0];
    }
    // capacity: 1.35x the mean row at the density of a full handle, >= 128
    const double V = E.box.L[0] * E.box.L[1] * E.box.L[2];
    const double mean = (double)n_max / V * 4.0 / 3.0 * M_PI * pow(E.r_list, 3);
    // half-shell rows hold half the neighbours on average, but not at most: the row of
    // an atom at the forward corner of its cell holds the 7/8 of its sphere that lies in
    // forward cells plus the later home atoms of the last 1/8, so they get the same
    // capacity as full rows
    int cap = (int)ceil(1.35 * mean + 32.0);
    cap = std::max(128, ((cap + 63) / 64) * 64);
    if (p->list_capacity > 0) cap = ((p->list_capacity + 63) / 64) * 64;

    auto bail = [&](int code) {
        g_create_err = E.err;
        free_all(E);
        delete h;
        return code;
    };
    int rc;
#define CR(expr)                                                  \
    if (hipError_t _s = (expr); _s != hipSuccess) {               \
        E.err = std::string(#expr) + ": " + hipGetErrorString(_s); \
        return bail(RBMD_ERR_HIP);                                \
    }
    const size_t N = (size_t)std::max<int64_t>(n_max, 1);
    {
        int dev = 0, v = 0;
        CR(hipGetDevice(&dev));
        CR(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        E.lds_max = v;
        CR(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        E.num_cu = v;
        CR(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
        E.lds_cu = v > 0 ? v : E.lds_max;
    }
    tile_geometry(E, n_max);
    CR(dalloc(&E.xr, N * 8));
    CR(dalloc(&E.yr, N * 8));
    CR(dalloc(&E.zr, N * 8));
    CR(dalloc(&E.xs_ref, N * sizeof(double4)));
    CR(dalloc(&E.xs, N * sizeof(double4)));
    CR(dalloc(&E.perm, N * 4));
    CR(dalloc(&E.scell, N * 4));
    CR(dalloc(&E.ts, N * 4));
    CR(dalloc(&E.atom_cell, N * 4));
    CR(dalloc(&E.atom_rank, N * 4));
    CR(dalloc(&E.cell_count, (size_t)E.ncell * 4));
    CR(dalloc(&E.tile_order, (size_t)E.ncell * 4 * 4));
    CR(reset_tile_order(E));
    {
        // per-cell buckets of the MD step's binning: 4x the mean population
        // (a cell beyond it is gathered by a scan instead)
        const double mean = (double)N / (double)E.ncell;
        E.bucket_cap = std::min(2048, std::max(64, ((int)(4.0 * mean) + 32 + 63) / 64 * 64));
        CR(dalloc(&E.md_count, md_count_bytes(E.ncell)));
        CR(dalloc(&E.md_bucket, (size_t)2 * E.ncell * E.bucket_cap * 4));
        CR(dalloc(&E.md_cell, N * 4));
        CR(dalloc(&E.md_vid, N * 4));
        CR(dalloc(&E.md_w, N * 3 * 8));
    }
    CR(dalloc(&E.cell_start, ((size_t)E.ncell + 1) * 4));
    CR(dalloc(&E.ncount, N * sizeof(int2)));
    // the build statistics (128 bytes) and the flags (64) share one block, so a run
    // reads both back with one copy
    CR(dalloc(&E.stats, 192));
    E.flags = (uint32_t*)((char*)E.stats + 128);
    CR(dalloc(&E.q, N * 8));
    CR(dalloc(&E.type, N * 4));
    CR(dalloc(&E.d_pp, sizeof(PairParams)));
    CR(dalloc(&E.d_step, 64));
    // thermostat: one partial per closing-kick workgroup, then scale / xi / T,
    // the log position and the noise step in one 64-byte block
    E.thermo_part_cap = std::max<int64_t>(((int64_t)N + kKfAtoms - 1) / kKfAtoms, 1);
    CR(dalloc(&E.thermo_part, (size_t)E.thermo_part_cap * 8));
    CR(dalloc(&E.thermo_st, 64));
    E.thermo_cnt = (uint32_t*)(E.thermo_st + 4);
    E.thermo_noise = (unsigned long long*)(E.thermo_st + 6);
    CR(dalloc(&E.ev, 16 * 8));
    E.ev_part_blocks = E.ncell * 4;  // tile workgroups (split <= 4)
    CR(dalloc(&E.ev_part, (size_t)E.ev_part_blocks * 8 * 8));
    if (E.half) CR(dalloc(&E.planes, (size_t)kHalfSegs * 3 * N * sizeof(float)));
    if (E.rbl) {
        CR(dalloc(&E.rbl_part, (size_t)E.ev_part_blocks * 3 * 8));
        CR(dalloc(&E.rbl_mean, 4 * 8));
        CR(hipMemset(E.rbl_mean, 0, 4 * 8));
    }
    CR(hipHostMalloc((void**)&E.h_stats, 192, hipHostMallocDefault));
    CR(hipStreamCreateWithFlags(&E.own_stream, hipStreamNonBlocking));
    CR(hipEventCreateWithFlags(&E.order_event, hipEventDisableTiming));
    CR(hipMemset(E.flags, 0, 64));
    CR(hipMemset(E.stats, 0, 128));
    CR(hipMemset(E.d_step, 0, 64));
    CR(hipMemset(E.thermo_st, 0, 64));
    CR(hipMemset(E.ev, 0, 16 * 8));
    CR(hipMemset(E.cell_start, 0, ((size_t)E.ncell + 1) * 4));
    CR(hipMemset(E.cell_count, 0, (size_t)E.ncell * 4));
    CR(hipMemset(E.md_count, 0, md_count_bytes(E.ncell)));
    CR(hipMemset(E.xr, 0, N * 8));
    CR(hipMemset(E.yr, 0, N * 8));
    CR(hipMemset(E.zr, 0, N * 8));
    CR(hipMemcpy(E.d_pp, &E.pp, sizeof(PairParams), hipMemcpyHostToDevice));
    memset(E.h_stats, 0, 64);
    if ((rc = alloc_list(E, cap)) != RBMD_OK) return bail(rc);
    if (E.kspace == RBMD_KSPACE_RBE) {
        build_kcdf(E.alpha, E.box.L, E.h_kcdf, E.kM);
        E.S = rbe_S(E.alpha, E.box.L);
        CR(dalloc(&E.kcdf, E.h_kcdf.size() * sizeof(double)));
        CR(hipMemcpy(E.kcdf, E.h_kcdf.data(), E.h_kcdf.size() * sizeof(double),
                     hipMemcpyHostToDevice));
        if ((rc = alloc_kspace(E, E.p_max)) != RBMD_OK) return bail(rc);
    } else if (E.kspace == RBMD_KSPACE_EWALD) {
        std::vector<int32_t> m;
        std::vector<double> w;
        ewald_lattice(E, m, w);
        const int P = (int)w.size();
        if ((rc = alloc_kspace(E, P)) != RBMD_OK) return bail(rc);
        CR(hipMemcpy(E.km, m.data(), m.size() * 4, hipMemcpyHostToDevice));
        CR(hipMemcpy(E.kw, w.data(), w.size() * 8, hipMemcpyHostToDevice));
        E.P = P;
    }
#undef CR
    *out = h;
    return RBMD_OK;
}

void rbmd_force_destroy(rbmd_force_t h)
{
    if (!h) return;
    hipDeviceSynchronize();
    free_all(h->e);
    delete h;
}

int rbmd_force_set_atoms(rbmd_force_t h, int64_t n, const double* q, const int32_t* type,
                         void* stream)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (n < 0 || n > E.n_max) return fail(E, RBMD_ERR_CAPACITY, "n exceeds n_max");
    if (n > 0 && !q) return fail(E, RBMD_ERR_INVALID, "charges required");
    hipStream_t s = (hipStream_t)stream;
    E.n = n;
    const int split = E.split;
    tile_geometry(E, n);
    if (E.split != split) HIP_TRY(reset_tile_order(E));
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(E.q, q, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
        if (type) {
            HIP_TRY(hipMemcpyAsync(E.type, type, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        } else {
            HIP_TRY(hipMemsetAsync(E.type, 0, (size_t)n * 4, s));
        }
        // self energy -sqrt(alpha/pi) sum q^2 (output/ThermoOutput.cpp:196-203), on the host
        std::vector<double> hq((size_t)n);
        HIP_TRY(hipMemcpyAsync(hq.data(), q, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        double sq = 0.0;
        for (double v : hq) sq += v * v;
        E.e_self = E.kspace ? -sqrt(E.alpha / M_PI) * sq * E.qqr2e : 0.0;
    }
    // the inverse permutation is read by the fused gather of the MD step before
    // the first build writes it: start it in range
    if (n > 0) HIP_TRY(hipMemsetAsync(E.atom_rank, 0, (size_t)n * 4, s));
    launch_force_rebuild(E, s);
    if (E.kspace == RBMD_KSPACE_RBE || E.kspace == RBMD_KSPACE_EWALD) {
        const int rc = alloc_kspace(E, E.kspace == RBMD_KSPACE_RBE ? std::max(E.p_max, E.P) : E.P);
        if (rc != RBMD_OK) return rc;
    }
    E.atoms_set = true;
    return RBMD_OK;
}

int rbmd_force_set_kvectors(rbmd_force_t h, int32_t P, const int32_t* m, const double* w,
                            void* stream)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (E.kspace != RBMD_KSPACE_RBE) return fail(E, RBMD_ERR_STATE, "handle has no RBE k-space");
    if (P < 1 || P > E.p_max || !m) return fail(E, RBMD_ERR_CAPACITY, "P out of [1, p_max]");
    hipStream_t s = (hipStream_t)stream;
    E.P = P;
    int rc = alloc_kspace(E, P);
    if (rc != RBMD_OK) return rc;
    HIP_TRY(hipMemcpyAsync(E.km, m, (size_t)P * 12, hipMemcpyDeviceToDevice, s));
    if (w) {
        HIP_TRY(hipMemcpyAsync(E.kw, w, (size_t)P * 8, hipMemcpyDeviceToDevice, s));
    } else {
        launch_kweights_default(E, s);
    }
    return RBMD_OK;
}

int rbmd_force_set_proposal(rbmd_force_t h, rbmd_proposal_t p)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (E.kspace != RBMD_KSPACE_RBE) return fail(E, RBMD_ERR_STATE, "handle has no RBE k-space");
    E.gp_gen += 1;
    if (!p) {
        E.gp = GridProposal{};
        return RBMD_OK;
    }
    for (int d = 0; d < 3; ++d)
        if (p->L[d] != E.box.L[d])
            return fail(E, RBMD_ERR_INVALID, "proposal box differs from the handle's box");
    if (p->alpha != E.alpha) return fail(E, RBMD_ERR_INVALID, "proposal alpha differs");
    const int64_t G = (int64_t)p->cdf.size();
    const int64_t K = (int64_t)p->guide.size();
    HIP_TRY(hipDeviceSynchronize());
    if (G > E.gp_cap || K > E.gp_guide_cap) {
        if (E.gp_cdf) hipFree(E.gp_cdf);
        if (E.gp_ratio) hipFree(E.gp_ratio);
        if (E.gp_guide) hipFree(E.gp_guide);
        E.gp_cdf = E.gp_ratio = nullptr;
        E.gp_guide = nullptr;
        E.gp_cap = E.gp_guide_cap = 0;
        E.gp = GridProposal{};
        HIP_TRY(hipMalloc((void**)&E.gp_cdf, (size_t)G * 8));
        HIP_TRY(hipMalloc((void**)&E.gp_ratio, (size_t)G * 8));
        HIP_TRY(hipMalloc((void**)&E.gp_guide, (size_t)K * 4));
        E.gp_cap = G;
        E.gp_guide_cap = K;
    }
    HIP_TRY(hipMemcpy(E.gp_cdf, p->cdf.data(), (size_t)G * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(E.gp_ratio, p->ratio.data(), (size_t)G * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(E.gp_guide, p->guide.data(), (size_t)K * 4, hipMemcpyHostToDevice));
    E.gp.cdf = E.gp_cdf;
    E.gp.ratio = E.gp_ratio;
    E.gp.guide = E.gp_guide;
    E.gp.guide_bits = p->guide_bits;
    E.gp.Mg = p->Mg;
    E.gp.lambda = p->lambda;
    return RBMD_OK;
}

int rbmd_force_sample_kvectors(rbmd_force_t h, int32_t P, uint64_t seed, uint64_t step,
                               void* stream)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (E.kspace != RBMD_KSPACE_RBE) return fail(E, RBMD_ERR_STATE, "handle has no RBE k-space");
    if (P < 1 || P > E.p_max) return fail(E, RBMD_ERR_CAPACITY, "P out of [1, p_max]");
    E.P = P;
    int rc = alloc_kspace(E, P);
    if (rc != RBMD_OK) return rc;
    launch_sample(E, P, seed, &step, (hipStream_t)stream);
    return RBMD_OK;
}

// One force evaluation at fixed positions: displacement check, list build (on
// request), gather, rho(k), tile pass, k-space force.  One stream: running the
// structure factor on a second stream concurrently with the list build and pair
// pass measured slower (cross-queue joins cost more than the overlap gains at
// these kernel sizes).
static int force_pass(Engine& E, const double* x, const double* y, const double* z, void* fx,
                      void* fy, void* fz, bool ev, hipStream_t s)
{
    launch_check(E, x, y, z, s);
    const bool ks = has_kspace(E);
    launch_build(E, x, y, z, s);
    launch_gather(E, x, y, z, s);
    if (ks) launch_rho(E, x, y, z, s);
    launch_pair(E, fx, fy, fz, ev, s);
    // a random batch list's near forces have their mean subtracted by the k-space
    // force kernel (which runs without k-vectors when there is no k-space)
    if (E.rbl) launch_rbl_mean(E, s);
    if (ks || E.rbl) launch_kforce(E, x, y, z, fx, fy, fz, s);
    if (ev) launch_ev_final(E, true, s);
    return RBMD_OK;
}

// Grow the list rows if the last build overflowed (rare: the pair pass is
// still exact through its stencil fallback, only slower).
static int maintain(Engine& E, hipStream_t s)
{
    if (E.half && E.h_stats[7] != 0ull) {
        // a half tile larger than its LDS force sums was skipped (stats[7], mirrored
        // after every call): report it once and return the handle to full rows (no
        // such limit) for the following calls
        unsigned long long f7 = 0;
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(&f7, E.stats + 7, 8, hipMemcpyDeviceToHost));
        if (f7) {
            E.half = false;
            tile_geometry(E, E.n_max);
            int rc = alloc_list(E, (int)std::max<int64_t>(2 * E.cap, 128));
            if (rc != RBMD_OK) return rc;
            HIP_TRY(hipMemset(E.stats + 7, 0, 8));
            E.h_stats[7] = 0ull;
            launch_force_rebuild(E, s);
            return fail(E, RBMD_ERR_CAPACITY,
                        "a half-shell tile exceeded its LDS force sums; the last forces are "
                        "invalid and the handle now uses full rows");
        }
    }
    if (E.h_stats[2] == 0ull) return RBMD_OK;
    HIP_TRY(hipStreamSynchronize(s));
    unsigned long long st[8];
    HIP_TRY(hipMemcpy(st, E.stats, 64, hipMemcpyDeviceToHost));
    if (st[2] == 0ull) {
        E.h_stats[2] = 0;
        return RBMD_OK;
    }
    // longest row of the last build (stats[4]) plus 20% headroom
    const int want = (int)(1.2 * (double)st[4]) + 16;
    const int newcap = (std::max((int)(E.cap * 1.5), want) + 63) / 64 * 64;
    int rc = alloc_list(E, newcap);
    if (rc != RBMD_OK) return rc;
    E.h_stats[2] = 0;
    launch_force_rebuild(E, s);
    return RBMD_OK;
}

int rbmd_force_compute(rbmd_force_t h, const double* x, const double* y, const double* z,
                       void* fx, void* fy, void* fz, int32_t flags, void* stream)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (!E.atoms_set) return fail(E, RBMD_ERR_STATE, "rbmd_force_set_atoms not called");
    if (E.kspace == RBMD_KSPACE_RBE && E.P < 1)
        return fail(E, RBMD_ERR_STATE, "no k-vectors: call set_kvectors or sample_kvectors");
    if (E.n == 0) return RBMD_OK;
    if (!x || !y || !z || !fx || !fy || !fz) return fail(E, RBMD_ERR_INVALID, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    int rc = maintain(E, s);
    if (rc != RBMD_OK) return rc;
    force_pass(E, x, y, z, fx, fy, fz, (flags & RBMD_EV) != 0, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(E.h_stats, E.stats, 64, hipMemcpyDeviceToHost, s));
    return RBMD_OK;
}

int rbmd_force_time_pair(rbmd_force_t h, const double* x, const double* y, const double* z,
                         void* fx, void* fy, void* fz, int32_t reps, double* mean_ms,
                         void* stream)
{
    if (!h || !mean_ms) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (reps < 1 || reps > 10000) return fail(E, RBMD_ERR_INVALID, "reps out of [1, 10000]");
    *mean_ms = 0.0;
    // two evaluations: the first may rebuild the rows, the second leaves no rebuild
    // pending, so the launches below are the pair pass alone
    for (int k = 0; k < 2; ++k) {
        const int rc = rbmd_force_compute(h, x, y, z, fx, fy, fz, 0, stream);
        if (rc != RBMD_OK) return rc;
    }
    if (E.n == 0) return RBMD_OK;
    hipStream_t s = (hipStream_t)stream;
    const bool timing = E.timing;
    E.timing = false;  // no per-launch events between the two below
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, s));
    for (int r = 0; r < reps; ++r) launch_pair(E, fx, fy, fz, false, s);
    HIP_TRY(hipEventRecord(b, s));
    HIP_TRY(hipEventSynchronize(b));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    HIP_TRY(hipEventDestroy(a));
    HIP_TRY(hipEventDestroy(b));
    E.timing = timing;
    *mean_ms = (double)ms / (double)reps;
    // the arrays end holding a whole evaluation again (pair plus k-space forces)
    return rbmd_force_compute(h, x, y, z, fx, fy, fz, 0, stream);
}

int rbmd_force_get_energy_virial(rbmd_force_t h, rbmd_energy_virial* out, void* stream)
{
    if (!h || !out) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    double v[10];
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    HIP_TRY(hipMemcpy(v, E.ev, sizeof v, hipMemcpyDeviceToHost));
    out->e_lj = v[0];
    out->e_real = v[1];
    out->e_self = v[2];
    out->e_kspace = v[3];
    for (int k = 0; k < 6; ++k) out->virial[k] = v[4 + k];
    return RBMD_OK;
}

// One run segment of `steps` velocity-Verlet steps: the opening kick/drift, then
// per step the rho partials + cell sort, the rho final + tile pass, and the
// k-space force with the closing half kick, fused with the next step's opening
// kick/drift except after the last step.  step_host: the first step's number
// (eager launches), or null under graph capture, where the sampler reads the
// device step counter that the rho final reduction advances.
static void md_segment(Engine& E, const RepState& st, double dt, uint64_t seed,
                       const uint64_t* step_host, int steps, hipStream_t s)
{
    auto sa_of = [&](int k) {
        SampleArgs sa;
        if (E.kspace != RBMD_KSPACE_RBE) return sa;
        if (!step_host) return sample_args(E, E.P, seed, nullptr);
        const uint64_t v = *step_host + (uint64_t)k;
        return sample_args(E, E.P, seed, &v);
    };
    if (E.thermo_type != kThermoNone) {
        // a thermostat needs the kinetic temperature of every atom between the
        // closing kick of a step and the opening kick of the next: each step is a
        // segment of its own (opening kick, structure factor, tile pass, closing
        // kick) and the temperature is reduced before the next opening kick's
        // velocities are used
        // (for up to kFuseThermoPartials closing-kick workgroups the opening kick of
        // step k > 0 reduces step k-1's temperature itself, every workgroup re-reading
        // the partials, and only the segment's last step needs the reduction kernel;
        // above that the re-reads would grow as n^2 and every step launches it)
        constexpr int kFuseThermoPartials = 2048;
        const bool fuse = std::max(kforce_blocks(E), 1) <= kFuseThermoPartials;
        for (int k = 0; k < steps; ++k) {
            launch_md_open(E, st, dt, sa_of(k), s, fuse && k > 0);
            launch_md_sort_rho(E, st, s);
            launch_tile_rho(E, st.fx, st.fy, st.fz, s);
            launch_rbl_mean(E, s);
            launch_md_close(E, st, dt, s);
            if (!fuse || k + 1 == steps) launch_thermo_reduce(E, dt, s);
        }
        return;
    }
    launch_md_open(E, st, dt, sa_of(0), s);
    for (int k = 0; k < steps; ++k) {
        launch_md_sort_rho(E, st, s);
        launch_tile_rho(E, st.fx, st.fy, st.fz, s);
        launch_rbl_mean(E, s);  // random batch list only
        if (k + 1 < steps) launch_md_close_open(E, st, dt, sa_of(k + 1), s);
        else launch_md_close(E, st, dt, s);
    }
}

int rbmd_md_run(rbmd_force_t h, double* x, double* y, double* z, double* vx, double* vy,
                double* vz, void* fx, void* fy, void* fz, int32_t* img, const double* mass,
                double dt, int64_t nsteps, uint64_t seed, uint64_t step0, int32_t use_graph,
                void* stream)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (!E.atoms_set) return fail(E, RBMD_ERR_STATE, "rbmd_force_set_atoms not called");
    if (E.kspace == RBMD_KSPACE_RBE && E.P < 1)
        return fail(E, RBMD_ERR_STATE, "no k-vectors: sample once to fix P");
    if (E.n == 0 || nsteps <= 0) return RBMD_OK;
    // Every launch and graph replay goes to the caller's stream (the null stream
    // included).  Capture needs a non-null stream: with a null caller stream the
    // graphs are captured on the handle's own stream (capture runs nothing, so no
    // ordering is needed) and replayed on the null stream.
    hipStream_t s = (hipStream_t)stream;
    hipStream_t cap = s ? s : E.own_stream;
    int rc = maintain(E, s);
    if (rc != RBMD_OK) return rc;
    // bin the current positions for a rebuild in the first step (later steps'
    // bins come from the tile pass of the step before); the binning kernel also
    // clears the long-step flag and sets the graphs' step counter to step0
    launch_md_bin(E, x, y, z, step0, s);
    if (E.thermo_type != kThermoNone) {
        if (!(thermo_dof(E) > 0))
            return fail(E, RBMD_ERR_STATE, "thermostat: no degrees of freedom (dof <= 0)");
        launch_thermo_begin(E, s);
    }
    RepState rs;
    rs.x = x; rs.y = y; rs.z = z;
    rs.vx = vx; rs.vy = vy; rs.vz = vz;
    rs.fx = fx; rs.fy = fy; rs.fz = fz;
    rs.img = img;
    rs.mass = mass;
    if (!use_graph) {
        // segments of at most 2^30 steps; the kernels give the same bits however
        // a run is cut into segments
        constexpr int64_t kEagerSegment = int64_t(1) << 30;
        for (int64_t k = 0; k < nsteps; k += kEagerSegment) {
            const uint64_t st = step0 + (uint64_t)k;
            md_segment(E, rs, dt, seed, &st, (int)std::min<int64_t>(kEagerSegment, nsteps - k), s);
        }
        HIP_TRY(hipGetLastError());
    } else {
        std::vector<const void*> key = {x, y, z, vx, vy, vz, fx, fy, fz, img, mass,
                                        (const void*)(uintptr_t)E.P, (const void*)(uintptr_t)E.n,
                                        (const void*)(uintptr_t)E.cap};
        // dt and seed are baked into the captured kernels: key them too
        uint64_t dtbits;
        memcpy(&dtbits, &dt, 8);
        key.push_back((const void*)(uintptr_t)dtbits);
        key.push_back((const void*)(uintptr_t)seed);
        // the proposal's table pointers, lambda and P-dependent weights are baked too
        key.push_back((const void*)(uintptr_t)E.gp_gen);
        // and the thermostat's constants
        key.push_back((const void*)(uintptr_t)E.thermo_gen);
        // and the launch shapes the autotuner may change
        for (int64_t v : {(int64_t)E.split, (int64_t)E.stage_cap, (int64_t)E.ncell,
                          (int64_t)E.rho_apb, (int64_t)E.kpart_cap})
            key.push_back((const void*)(uintptr_t)v);
        // cached graphs: kGraphSteps steps (graphs[0]) and the remainder r < kGraphSteps
        // (graphs[r]), captured on first use
        const int rem = (int)(nsteps % kGraphSteps);
        for (int g : {0, rem}) {
            Engine::StepGraph& G = E.graphs[g];
            const int steps = g == 0 ? kGraphSteps : g;
            if (g == 0 && nsteps < kGraphSteps) continue;
            if (g != 0 && rem == 0) continue;
            if (G.exec && key == G.key) continue;
            if (G.exec) hipGraphExecDestroy(G.exec);
            if (G.graph) hipGraphDestroy(G.graph);
            G.exec = nullptr;
            G.graph = nullptr;
            E.capturing = true;
            HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
            md_segment(E, rs, dt, seed, nullptr, steps, cap);
            hipError_t st = hipStreamEndCapture(cap, &G.graph);
            E.capturing = false;
            if (st != hipSuccess) {
                E.err = std::string("graph capture failed: ") + hipGetErrorString(st);
                return RBMD_ERR_HIP;
            }
            HIP_TRY(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
            G.key = key;
        }
        // the device step counter holds step0 (set by the binning kernel); the
        // graph's rho_final kernel advances it
        for (int64_t k = 0; k + kGraphSteps <= nsteps; k += kGraphSteps)
            HIP_TRY(hipGraphLaunch(E.graphs[0].exec, s));
        if (rem > 0) HIP_TRY(hipGraphLaunch(E.graphs[rem].exec, s));
    }
    if (E.thermo_type != kThermoNone) launch_thermo_end(E, rs, s);
    // the statistics and the flags after them: one copy, one wait
    HIP_TRY(hipMemcpyAsync(E.h_stats, E.stats, 128 + 5 * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const uint32_t* fl = (const uint32_t*)((const char*)E.h_stats + 128);
    if (fl[1]) return fail(E, RBMD_ERR_INVALID, "non-finite positions during the run");
    if (fl[4])
        return fail(E, RBMD_ERR_STATE, kStepTooLong);
    return RBMD_OK;
}

int rbmd_force_get_neighbors(rbmd_force_t h, int64_t* offsets, int32_t* idx, int64_t cap,
                             int64_t* count)
{
    if (!h || !offsets || !count) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    HIP_TRY(hipDeviceSynchronize());
    const int64_t n = E.n;
    std::vector<int32_t> perm((size_t)n), scell((size_t)n), cs((size_t)E.ncell + 1);
    std::vector<uint16_t> rows((size_t)n * E.cap);
    std::vector<int2> nc((size_t)n);
    HIP_TRY(hipMemcpy(perm.data(), E.perm, (size_t)n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(scell.data(), E.scell, (size_t)n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cs.data(), E.cell_start, cs.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(nc.data(), E.ncount, (size_t)n * sizeof(int2), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rows.data(), E.nlist, rows.size() * 2, hipMemcpyDeviceToHost));
    // rows hold tile-local indices of the atom's cell tile (stencil order of tile.h;
    // half tiles: the home cell, then the forward cells in stencil order, half.h)
    const int nseg = E.half ? kHalfSegs : 27;
    auto tile = [&](int c, int off[28], int gstart[27]) {
        const int dx = E.dims[0], dy = E.dims[1], dz = E.dims[2];
        const int cz = c / (dx * dy), cy = (c / dx) % dy, cx = c % dx;
        int order[27], k = 0;
        if (E.half) {
            order[k++] = kHomeSeg;
            for (int t = 0; t < 27; ++t) {
                const int a = t / 9 - 1, b = (t / 3) % 3 - 1, g = t % 3 - 1;
                if (g > 0 || (g == 0 && (b > 0 || (b == 0 && a > 0)))) order[k++] = t;
            }
        } else {
            for (int t = 0; t < 27; ++t) order[k++] = t;
        }
        int o = 0;
        for (int q = 0; q < nseg; ++q) {
            const int t = order[q];
            const int ax = (cx + t / 9 - 1 + dx) % dx, ay = (cy + (t / 3) % 3 - 1 + dy) % dy,
                      az = (cz + t % 3 - 1 + dz) % dz;
            const int cc = ax + ay * dx + az * dx * dy;
            off[q] = o;
            gstart[q] = cs[(size_t)cc];
            o += cs[(size_t)cc + 1] - cs[(size_t)cc];
        }
        off[nseg] = o;
    };
    std::vector<int64_t> slot_of((size_t)n);
    for (int64_t s = 0; s < n; ++s) slot_of[(size_t)perm[(size_t)s]] = s;
    bool overflow = false;
    int off[28], gst[27], last_c = -1;
    // the rows in atom ids; a half-shell row's pair (i, j) also goes to j's row
    std::vector<std::vector<int32_t>> lists((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t s = slot_of[(size_t)i];
        const int ni = nc[(size_t)s].x, no = nc[(size_t)s].y;
        if (ni + no >= E.cap) overflow = true;
        const int c = scell[(size_t)s];
        if (c != last_c) {
            tile(c, off, gst);
            last_c = c;
        }
        const uint16_t* row = rows.data() + (size_t)s * E.cap;
        for (int k = 0; k < ni + no && !overflow; ++k) {
            const int u = row[k] & kMaxTile;  // a random batch list's selection tag aside
            int st = nseg - 1;
            while (off[st] > u) --st;
            const int32_t j = perm[(size_t)(gst[st] + (u - off[st]))];
            lists[(size_t)i].push_back(j);
            if (E.half) lists[(size_t)j].push_back((int32_t)i);
        }
    }
    int64_t e = 0;
    offsets[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int32_t j : lists[(size_t)i]) {
            if (idx && e < cap) idx[e] = j;
            ++e;
        }
        offsets[i + 1] = e;
    }
    *count = e;
    if (overflow) return fail(E, RBMD_ERR_CAPACITY, "list rows overflowed at the last build");
    return RBMD_OK;
}

int rbmd_force_get_kspace(rbmd_force_t h, int32_t* P, int32_t* m, double* w, double* re,
                          double* im)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    HIP_TRY(hipDeviceSynchronize());
    if (P) *P = E.P;
    if (E.P == 0) return RBMD_OK;
    if (m) HIP_TRY(hipMemcpy(m, E.km, (size_t)E.P * 12, hipMemcpyDeviceToHost));
    if (w) HIP_TRY(hipMemcpy(w, E.kw, (size_t)E.P * 8, hipMemcpyDeviceToHost));
    if (re || im) {
        std::vector<double> r((size_t)E.P * 2);
        HIP_TRY(hipMemcpy(r.data(), E.rho, r.size() * 8, hipMemcpyDeviceToHost));
        for (int k = 0; k < E.P; ++k) {
            if (re) re[k] = r[2 * (size_t)k];
            if (im) im[k] = r[2 * (size_t)k + 1];
        }
    }
    return RBMD_OK;
}

int rbmd_md_set_thermostat(rbmd_force_t h, const rbmd_thermostat* t)
{
    if (!h || !t) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (t->type < RBMD_THERMO_NONE || t->type > RBMD_THERMO_MEASURE)
        return fail(E, RBMD_ERR_INVALID, "thermostat: unknown type");
    if (t->type != RBMD_THERMO_NONE) {
        if (!(t->t_target > 0) && t->type != RBMD_THERMO_LANGEVIN &&
            t->type != RBMD_THERMO_MEASURE)
            return fail(E, RBMD_ERR_INVALID, "thermostat: t_target must be > 0");
        if (t->type == RBMD_THERMO_BERENDSEN && !(t->t_damp > 0))
            return fail(E, RBMD_ERR_INVALID, "thermostat: BERENDSEN needs t_damp > 0");
        if (!(t->dof >= 0) || !(t->kb >= 0))
            return fail(E, RBMD_ERR_INVALID, "thermostat: dof and kb must be >= 0");
        if (t->type == RBMD_THERMO_LANGEVIN && (!(t->gamma >= 0) || !(t->kbt >= 0)))
            return fail(E, RBMD_ERR_INVALID, "thermostat: LANGEVIN needs gamma, kbt >= 0");
    }
    E.thermo_type = t->type;
    E.thermo_t0 = t->t_target;
    E.thermo_tdamp = t->t_damp;
    E.thermo_dof = t->dof;
    E.thermo_kb = t->kb > 0 ? t->kb : 1.0;
    E.thermo_gamma = t->gamma;
    E.thermo_kbt = t->kbt;
    E.thermo_seed = t->seed;
    ++E.thermo_gen;
    // device state: scale 1, xi, T 0, counters 0, noise step
    double st[8] = {1.0, t->xi, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    memcpy(&st[6], &t->noise_step, 8);
    HIP_TRY(hipMemcpy(E.thermo_st, st, sizeof st, hipMemcpyHostToDevice));
    return RBMD_OK;
}

int rbmd_md_get_thermostat(rbmd_force_t h, double* temperature, double* xi, uint64_t* noise_step)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    double st[8];
    HIP_TRY(hipMemcpy(st, E.thermo_st, sizeof st, hipMemcpyDeviceToHost));
    if (temperature) *temperature = st[2];
    if (xi) *xi = st[1];
    if (noise_step) memcpy(noise_step, &st[6], 8);
    return RBMD_OK;
}

int rbmd_md_set_temperature_log(rbmd_force_t h, double* log, int64_t len)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (len < 0 || (len > 0 && !log)) return fail(E, RBMD_ERR_INVALID, "temperature log: bad length");
    E.thermo_log = len > 0 ? log : nullptr;
    E.thermo_log_len = len;
    ++E.thermo_gen;
    return RBMD_OK;
}

int rbmd_langevin_noise_host(uint64_t seed, uint64_t step, double* g)
{
    if (!g) return RBMD_ERR_INVALID;
    langevin_gaussian(seed, step, g);
    return RBMD_OK;
}

int rbmd_force_get_stats(rbmd_force_t h, int64_t* out)
{
    if (!h || !out) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    HIP_TRY(hipDeviceSynchronize());
    unsigned long long st[8];
    HIP_TRY(hipMemcpy(st, E.stats, 64, hipMemcpyDeviceToHost));
    unsigned long long ls[3] = {0, 0, 0};
    if (E.n > 0) {
        launch_list_stats(E, E.stats + 8, nullptr);
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipMemcpy(ls, E.stats + 8, sizeof ls, hipMemcpyDeviceToHost));
    }
    out[0] = (int64_t)st[0];
    out[1] = (int64_t)ls[0] * (E.half ? 2 : 1);  // full-list entries (a half row holds each pair once)
    out[2] = E.cap;
    out[3] = (int64_t)st[2];
    out[4] = E.ncell;
    out[5] = (int64_t)ls[1] * (E.half ? 2 : 1);
    out[6] = E.kspace == RBMD_KSPACE_EWALD ? E.P : 0;
    out[7] = E.n;
    out[8] = (int64_t)st[6];
    out[9] = E.stage_cap;
    // largest tile: stencil populations from the cell bounds
    std::vector<int32_t> cs((size_t)E.ncell + 1);
    HIP_TRY(hipMemcpy(cs.data(), E.cell_start, cs.size() * 4, hipMemcpyDeviceToHost));
    int64_t big = 0;
    const int dx = E.dims[0], dy = E.dims[1], dz = E.dims[2];
    for (int c = 0; c < E.ncell; ++c) {
        const int cz = c / (dx * dy), cy = (c / dx) % dy, cx = c % dx;
        int64_t tot = 0;
        for (int t = 0; t < 27; ++t) {
            const int ax = (cx + t / 9 - 1 + dx) % dx, ay = (cy + (t / 3) % 3 - 1 + dy) % dy,
                      az = (cz + t % 3 - 1 + dz) % dz;
            const int cc = ax + ay * dx + az * dx * dy;
            tot += cs[(size_t)cc + 1] - cs[(size_t)cc];
        }
        big = std::max(big, tot);
    }
    out[10] = big;
    out[11] = E.split;
    {
        int apb = 0, nbx = 0, nkt = 0;
        if (E.P > 0) rho_geometry(E, &apb, &nbx, &nkt);
        out[12] = apb;
    }
    out[13] = (int64_t)(E.tuned_ms * 1e3);
    return RBMD_OK;
}

int rbmd_force_set_timing(rbmd_force_t h, int32_t enable)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    timer_drain(E);
    for (int p = 0; p < P_NUM; ++p) {
        E.t_sum[p] = 0;
        E.t_cnt[p] = 0;
    }
    E.timing = enable != 0;
    return RBMD_OK;
}

int rbmd_force_get_timing(rbmd_force_t h, const char* pass, double* mean_ms, int64_t* launches)
{
    if (!h || !pass) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    timer_drain(E);
    for (int p = 0; p < P_NUM; ++p)
        if (strcmp(pass, kPassNames[p]) == 0) {
            if (mean_ms) *mean_ms = E.t_cnt[p] ? E.t_sum[p] / (double)E.t_cnt[p] : 0.0;
            if (launches) *launches = E.t_cnt[p];
            return RBMD_OK;
        }
    return fail(E, RBMD_ERR_INVALID, "unknown pass name");
}

// ---------------------------------------------------------------- autotuner

}  // extern "C"

// Tile shapes searched by the autotuner.
struct TilePolicy {
    int split;          // workgroups per cell tile (0: heuristic)
    double stage_scale; // LDS stage = scale * mean stencil population + 64
    int rho_apb;        // atoms per rho(k) workgroup (0: heuristic)
};

static void apply_policy(Engine& E, const TilePolicy& p)
{
    const int split = E.split;
    E.tune_split = p.split;
    E.tune_stage_scale = p.stage_scale;
    E.rho_apb = p.rho_apb;
    tile_geometry(E, E.n);
    if (E.split != split) (void)reset_tile_order(E);
}

// Expected share of steps that rebuild the list (weights the build's cost).
static constexpr double kRebuildShare = 0.15;

// Search order: (split, stage) with the rho(k) heuristic, then the rho(k)
// workgroup size with the best tile shape.  measure(policy) returns the cost
// in ms (negative: failure).
template <class Measure>
static TilePolicy autotune_search(const TilePolicy& start, Measure measure, double* best_ms)
{
    TilePolicy best = start;
    double bc = measure(best);
    for (int split : {1, 2, 3, 4})
        for (double sc : {0.0, 1.1, 1.25, 1.5}) {
            const TilePolicy p{split, sc, start.rho_apb};
            const double c = measure(p);
            if (c >= 0 && c < bc) {
                bc = c;
                best = p;
            }
        }
    for (int apb : {64, 128, 256, 512, 1024, 2048, 4096}) {
        TilePolicy p = best;
        p.rho_apb = apb;
        const double c = measure(p);
        if (c >= 0 && c < bc) {
            bc = c;
            best = p;
        }
    }
    *best_ms = bc;
    return best;
}

struct TuneEvents {
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ok = true;
    TuneEvents()
    {
        for (auto& e : ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    }
    ~TuneEvents()
    {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    // times of the last rebuild pass (ev0 -> ev1) and plain pass (ev2 -> ev3), ms
    // (negative on error)
    double read(double* tb, double* tf)
    {
        float a = 0, b = 0;
        if (hipEventSynchronize(ev[3]) != hipSuccess ||
            hipEventElapsedTime(&a, ev[0], ev[1]) != hipSuccess ||
            hipEventElapsedTime(&b, ev[2], ev[3]) != hipSuccess)
            return -1.0;
        *tb = a;
        *tf = b;
        return 0.0;
    }
};

extern "C" {

int rbmd_force_autotune(rbmd_force_t h, const double* x, const double* y, const double* z,
                        void* fx, void* fy, void* fz, int32_t reps, void* stream)
{
    if (!h) return RBMD_ERR_INVALID;
    Engine& E = h->e;
    if (!E.atoms_set) return fail(E, RBMD_ERR_STATE, "rbmd_force_set_atoms not called");
    if (E.kspace == RBMD_KSPACE_RBE && E.P < 1)
        return fail(E, RBMD_ERR_STATE, "no k-vectors: call set_kvectors or sample_kvectors");
    if (E.n == 0) return RBMD_OK;
    if (!x || !y || !z || !fx || !fy || !fz) return fail(E, RBMD_ERR_INVALID, "null pointer");
    if (reps < 1 || reps > 100) return fail(E, RBMD_ERR_INVALID, "reps out of [1, 100]");
    hipStream_t s = (hipStream_t)stream;
    int rc = maintain(E, s);
    if (rc != RBMD_OK) return rc;
    TuneEvents te;
    if (!te.ok) return fail(E, RBMD_ERR_HIP, "event creation failed");
    const bool ks = (E.kspace == RBMD_KSPACE_RBE || E.kspace == RBMD_KSPACE_EWALD) && E.P > 0;
    auto pass = [&]() {
        launch_build(E, x, y, z, s);
        launch_gather(E, x, y, z, s);
        if (ks) launch_rho(E, x, y, z, s);
        launch_pair(E, fx, fy, fz, false, s);
        if (ks) launch_kforce(E, x, y, z, fx, fy, fz, s);
    };
    auto measure = [&](const TilePolicy& p) -> double {
        apply_policy(E, p);
        double sb = 0, sf = 0;
        for (int r = 0; r <= reps; ++r) {  // the first round warms up
            // a pass that rebuilds (into a zeroed histogram: the per-step check
            // kernel clears it in ordinary passes), then one that does not
            (void)hipMemsetAsync(E.cell_count, 0, (size_t)E.ncell * sizeof(int32_t), s);
            launch_force_rebuild(E, s);
            (void)hipEventRecord(te.ev[0], s);
            pass();
            (void)hipEventRecord(te.ev[1], s);
            (void)hipEventRecord(te.ev[2], s);
            pass();
            (void)hipEventRecord(te.ev[3], s);
            double tb, tf;
            if (te.read(&tb, &tf) < 0) return -1.0;
            if (r > 0) {
                sb += tb;
                sf += tf;
            }
        }
        return (sf + kRebuildShare * (sb - sf)) / reps;
    };
    const TilePolicy start{E.tune_split, E.tune_stage_scale, E.rho_apb};
    double best_ms = 0;
    const TilePolicy best = autotune_search(start, measure, &best_ms);
    apply_policy(E, best);
    launch_force_rebuild(E, s);  // the next pass builds with the chosen shapes
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    E.tuned_ms = best_ms;
    return RBMD_OK;
}

// ---------------------------------------------------------------- replica batches

}  // extern "C"

struct rbmd_batch_s {
    std::vector<rbmd_force_t> reps;
    Batch b;
    hipStream_t own_stream = nullptr;
    hipEvent_t order_event = nullptr;
    Engine::StepGraph graphs[kGraphSteps];  // [0] kGraphSteps steps, [r] r steps
    std::string err;
};

static int bfail(rbmd_batch_t b, int code, const std::string& msg)
{
    b->err = msg;
    return code;
}

#define BHIP_TRY(expr)                                                                   \
    do {                                                                                 \
        hipError_t _st = (expr);                                                         \
        if (_st != hipSuccess) return bfail(b, RBMD_ERR_HIP, std::string(#expr) + ": " + \
                                                                 hipGetErrorString(_st)); \
    } while (0)

// A batched run segment of `steps` velocity-Verlet steps: the kernel sequence of
// md_segment, every phase one launch over all replicas.  The replicas' samplers
// read their device step counters (advanced by the rho final reduction).
static void batch_segment(rbmd_batch_t b, int steps, hipStream_t s)
{
    Batch& B = b->b;
    if (B.thermo) {
        // as md_segment with a thermostat: every step is a segment of its own, and
        // the kinetic temperature of each thermostatted replica is reduced after
        // its closing kick (one workgroup per replica) before the next opening kick
        // applies the coefficient; replicas without a thermostat take the same
        // steps with no reduction (their opening and closing kicks carry none)
        for (int k = 0; k < steps; ++k) {
            md_batch_launch(B, BP_KICKDRIFT, s);
            md_batch_launch(B, BP_SORT_RHO, s);
            pair_batch_launch(B, BP_PAIR, s);
            md_batch_launch(B, BP_KFORCE, s);
            md_batch_launch(B, BP_THERMO_REDUCE, s);
        }
        return;
    }
    md_batch_launch(B, BP_KICKDRIFT, s);
    for (int k = 0; k < steps; ++k) {
        md_batch_launch(B, BP_SORT_RHO, s);
        pair_batch_launch(B, BP_PAIR, s);
        md_batch_launch(B, k + 1 < steps ? BP_KFORCE_DRIFT : BP_KFORCE, s);
    }
}

extern "C" {

int rbmd_batch_create(int32_t nrep, const rbmd_force_params* params, const int64_t* n_max,
                      rbmd_batch_t* out)
{
    if (!out || nrep < 1 || !params || !n_max) return RBMD_ERR_INVALID;
    *out = nullptr;
    rbmd_batch_t b = new rbmd_batch_s();
    for (int32_t r = 0; r < nrep; ++r) {
        const rbmd_force_params& p = params[r];
        if (p.neighbor_type == RBMD_NEIGHBOR_RBL) {
            g_create_err = "batched replicas use the full Verlet list (RBL is single-handle only)";
            rbmd_batch_destroy(b);
            return RBMD_ERR_INVALID;
        }
        if (p.precision != params[0].precision || p.kspace != params[0].kspace ||
            (p.ntypes > 1) != (params[0].ntypes > 1)) {
            g_create_err = "replicas must share precision, k-space mode and single/multi type";
            rbmd_batch_destroy(b);
            return RBMD_ERR_INVALID;
        }
        rbmd_force_t h = nullptr;
        // replicas share the batched full-row tile pass
        rbmd_force_params pr = p;
        pr.half_shell = 0;
        const int rc = rbmd_force_create(&pr, n_max[r], &h);
        if (rc != RBMD_OK) {
            rbmd_batch_destroy(b);
            return rc;
        }
        // the batch fills the GPU with whole tiles, so a tile is never split
        // over several workgroups
        h->e.tune_split = 1;
        b->reps.push_back(h);
        b->b.es.push_back(&h->e);
    }
    b->b.precision = params[0].precision;
    b->b.kspace = params[0].kspace;
    b->b.multi = params[0].ntypes > 1;
    const size_t out_bytes = (size_t)nrep * kStatsOutWords * sizeof(unsigned long long);
    if (hipMalloc((void**)&b->b.d_out, out_bytes) != hipSuccess ||
        hipHostMalloc((void**)&b->b.h_out, out_bytes, hipHostMallocDefault) != hipSuccess) {
        g_create_err = "batch readback allocation failed";
        rbmd_batch_destroy(b);
        return RBMD_ERR_HIP;
    }
    b->b.out_cap = (size_t)nrep;
    if (hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->order_event, hipEventDisableTiming) != hipSuccess) {
        g_create_err = "stream/event creation failed";
        rbmd_batch_destroy(b);
        return RBMD_ERR_HIP;
    }
    *out = b;
    return RBMD_OK;
}

void rbmd_batch_destroy(rbmd_batch_t b)
{
    if (!b) return;
    for (auto& G : b->graphs) {
        if (G.exec) hipGraphExecDestroy(G.exec);
        if (G.graph) hipGraphDestroy(G.graph);
    }
    for (auto& ph : b->b.ph) ph.release();
    for (auto& ph : b->b.pair_more) ph.release();
    for (rbmd_force_t h : b->reps) rbmd_force_destroy(h);
    if (b->b.d_out) hipFree(b->b.d_out);
    if (b->b.h_out) hipHostFree(b->b.h_out);
    if (b->own_stream) hipStreamDestroy(b->own_stream);
    if (b->order_event) hipEventDestroy(b->order_event);
    delete b;
}

const char* rbmd_batch_last_error(rbmd_batch_t b)
{
    if (!b) return g_create_err.c_str();
    return b->err.c_str();
}

rbmd_force_t rbmd_batch_replica(rbmd_batch_t b, int32_t r)
{
    if (!b || r < 0 || r >= (int32_t)b->reps.size()) return nullptr;
    return b->reps[(size_t)r];
}

int rbmd_batch_md_run(rbmd_batch_t b, const rbmd_replica_state* st, double dt, int64_t nsteps,
                      const uint64_t* seeds, uint64_t step0, int32_t use_graph, void* stream)
{
    if (!b || !st || !seeds) return RBMD_ERR_INVALID;
    Batch& B = b->b;
    const size_t R = b->reps.size();
    for (size_t r = 0; r < R; ++r) {
        const Engine& E = *B.es[r];
        const rbmd_replica_state& q = st[r];
        if (!E.atoms_set) return bfail(b, RBMD_ERR_STATE, "replica without set_atoms");
        if (E.kspace == RBMD_KSPACE_RBE && E.P < 1)
            return bfail(b, RBMD_ERR_STATE, "replica without k-vectors: sample once to fix P");
        if (E.n > 0 && (!q.x || !q.y || !q.z || !q.vx || !q.vy || !q.vz || !q.fx || !q.fy ||
                        !q.fz))
            return bfail(b, RBMD_ERR_INVALID, "null state pointer");
        if (E.thermo_type != kThermoNone && !(thermo_dof(E) > 0))
            return bfail(b, RBMD_ERR_STATE, "replica " + std::to_string(r) +
                                                ": thermostat without degrees of freedom");
    }
    if (nsteps <= 0) return RBMD_OK;
    // launches and replays on the caller's stream; capture on a non-null stream
    // (the batch's own when the caller's is null; capture runs nothing)
    hipStream_t s = (hipStream_t)stream;
    hipStream_t cap = s ? s : b->own_stream;
    for (size_t r = 0; r < R; ++r) {
        int rc = maintain(*B.es[r], s);
        if (rc != RBMD_OK) return bfail(b, rc, "replica " + std::to_string(r) + ": " + B.es[r]->err);
    }
    // arguments are read from device memory by the batched kernels: wait for
    // earlier work on the stream before overwriting them
    BHIP_TRY(hipStreamSynchronize(s));
    B.st.resize(R);
    B.sa.assign(R, SampleArgs{});
    B.dt = dt;
    B.thermo = false;
    for (const Engine* E : B.es) B.thermo = B.thermo || E->thermo_type != kThermoNone;
    for (size_t r = 0; r < R; ++r) {
        const rbmd_replica_state& q = st[r];
        RepState& o = B.st[r];
        o.x = q.x; o.y = q.y; o.z = q.z;
        o.vx = q.vx; o.vy = q.vy; o.vz = q.vz;
        o.fx = q.fx; o.fy = q.fy; o.fz = q.fz;
        o.img = q.img;
        o.mass = q.mass;
        Engine& E = *B.es[r];
        if (E.kspace == RBMD_KSPACE_RBE) B.sa[r] = sample_args(E, E.P, seeds[r], nullptr);
    }
    // the binning kernel sets every replica's step counter to step0
    B.step0 = step0;
    BHIP_TRY(nlist_batch_prepare(B));
    BHIP_TRY(pair_batch_prepare(B));
    BHIP_TRY(md_batch_prepare(B));
    // bins of the current positions for a rebuild in the first step: the counts
    // cleared, then binned (which also clears the long-step flags and sets the step
    // counters), one grid each over all replicas
    nlist_batch_launch(B, BP_MDCLEAR, s);
    nlist_batch_launch(B, BP_MDBIN, s);
    nlist_batch_launch(B, BP_CHUNKSUM, s);
    if (B.thermo) md_batch_launch(B, BP_THERMO_BEGIN, s);
    if (!use_graph) {
        constexpr int64_t kEagerSegment = int64_t(1) << 30;
        for (int64_t k = 0; k < nsteps; k += kEagerSegment)
            batch_segment(b, (int)std::min<int64_t>(kEagerSegment, nsteps - k), s);
        BHIP_TRY(hipGetLastError());
    } else {
        std::vector<const void*> key;
        auto key_phase = [&](const BatchPhase& ph) {
            key.push_back(ph.args);
            key.push_back(ph.map);
            key.push_back((const void*)(uintptr_t)ph.items);
            key.push_back((const void*)(uintptr_t)ph.lds);
        };
        for (const BatchPhase& ph : B.ph) key_phase(ph);
        for (const BatchPhase& ph : B.pair_more) key_phase(ph);
        key.push_back((const void*)(uintptr_t)(B.thermo ? 1 : 0));
        const int rem = (int)(nsteps % kGraphSteps);
        for (int g : {0, rem}) {
            Engine::StepGraph& G = b->graphs[g];
            const int steps = g == 0 ? kGraphSteps : g;
            if (g == 0 && nsteps < kGraphSteps) continue;
            if (g != 0 && rem == 0) continue;
            if (G.exec && key == G.key) continue;
            if (G.exec) hipGraphExecDestroy(G.exec);
            if (G.graph) hipGraphDestroy(G.graph);
            G.exec = nullptr;
            G.graph = nullptr;
            BHIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
            batch_segment(b, steps, cap);
            hipError_t e = hipStreamEndCapture(cap, &G.graph);
            if (e != hipSuccess)
                return bfail(b, RBMD_ERR_HIP, std::string("graph capture failed: ") +
                                                  hipGetErrorString(e));
            BHIP_TRY(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
            G.key = key;
        }
        for (int64_t k = 0; k + kGraphSteps <= nsteps; k += kGraphSteps)
            BHIP_TRY(hipGraphLaunch(b->graphs[0].exec, s));
        if (rem > 0) BHIP_TRY(hipGraphLaunch(b->graphs[rem].exec, s));
    }
    if (B.thermo) md_batch_launch(B, BP_THERMO_END, s);
    // every replica's statistics and flags gathered into one block: one copy, one wait
    nlist_batch_launch(B, BP_STATS_OUT, s);
    BHIP_TRY(hipMemcpyAsync(B.h_out, B.d_out, R * kStatsOutWords * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, s));
    BHIP_TRY(hipStreamSynchronize(s));
    for (size_t r = 0; r < R; ++r) {
        Engine& E = *B.es[r];
        const unsigned long long* o = B.h_out + r * kStatsOutWords;
        memcpy(E.h_stats, o, 64);
        uint32_t fl[5];
        for (int k = 0; k < 5; ++k) fl[k] = (uint32_t)o[8 + k];
        if (fl[1])
            return bfail(b, RBMD_ERR_INVALID,
                         "non-finite positions in replica " + std::to_string(r));
        if (fl[4])
            return bfail(b, RBMD_ERR_STATE, std::string(kStepTooLong) + " (replica " +
                                                std::to_string(r) + ")");
    }
    return RBMD_OK;
}

int rbmd_batch_autotune(rbmd_batch_t b, const rbmd_replica_state* st, int32_t reps,
                        void* stream)
{
    if (!b || !st) return RBMD_ERR_INVALID;
    if (reps < 1 || reps > 100) return bfail(b, RBMD_ERR_INVALID, "reps out of [1, 100]");
    Batch& B = b->b;
    const size_t R = b->reps.size();
    for (size_t r = 0; r < R; ++r) {
        const Engine& E = *B.es[r];
        if (!E.atoms_set) return bfail(b, RBMD_ERR_STATE, "replica without set_atoms");
        if (E.kspace == RBMD_KSPACE_RBE && E.P < 1)
            return bfail(b, RBMD_ERR_STATE, "replica without k-vectors: sample once to fix P");
        if (!st[r].x || !st[r].vx || !st[r].fx) return bfail(b, RBMD_ERR_INVALID, "null state");
    }
    hipStream_t s = (hipStream_t)stream;
    for (size_t r = 0; r < R; ++r) {
        int rc = maintain(*B.es[r], s);
        if (rc != RBMD_OK) return bfail(b, rc, B.es[r]->err);
    }
    BHIP_TRY(hipStreamSynchronize(s));
    TuneEvents te;
    if (!te.ok) return bfail(b, RBMD_ERR_HIP, "event creation failed");
    // the tuning passes run the force phases with dt = 0: the fused closing kick
    // then leaves the velocities unchanged
    B.st.resize(R);
    B.sa.assign(R, SampleArgs{});
    B.dt = 0.0;
    for (size_t r = 0; r < R; ++r) {
        const rbmd_replica_state& q = st[r];
        RepState& o = B.st[r];
        o.x = q.x; o.y = q.y; o.z = q.z;
        o.vx = q.vx; o.vy = q.vy; o.vz = q.vz;
        o.fx = q.fx; o.fy = q.fy; o.fz = q.fz;
        o.img = q.img;
        o.mass = q.mass;
    }
    const bool ks = B.kspace == RBMD_KSPACE_RBE || B.kspace == RBMD_KSPACE_EWALD;
    auto measure = [&](const TilePolicy& p) -> double {
        for (Engine* E : B.es) apply_policy(*E, p);
        if (nlist_batch_prepare(B) != hipSuccess || pair_batch_prepare(B) != hipSuccess ||
            md_batch_prepare(B) != hipSuccess || (ks && kspace_batch_prepare(B) != hipSuccess))
            return -1.0;
        auto pass = [&]() {
            for (int ph : {BP_BIN, BP_SCAN, BP_SCATTER, BP_CELLSORT}) nlist_batch_launch(B, ph, s);
            nlist_batch_launch(B, BP_GATHER, s);
            if (ks) kspace_batch_launch(B, BP_RHO, s);
            pair_batch_launch(B, BP_PAIR, s);  // rho final + tile pass
            md_batch_launch(B, BP_KFORCE, s);
        };
        double sb = 0, sf = 0;
        for (int r = 0; r <= reps; ++r) {
            for (Engine* E : B.es) {
                // a rebuild bins into a zeroed histogram (the kick/check kernel
                // clears it in ordinary steps)
                (void)hipMemsetAsync(E->cell_count, 0, (size_t)E->ncell * sizeof(int32_t), s);
                launch_force_rebuild(*E, s);
            }
            (void)hipEventRecord(te.ev[0], s);
            pass();
            (void)hipEventRecord(te.ev[1], s);
            (void)hipEventRecord(te.ev[2], s);
            pass();
            (void)hipEventRecord(te.ev[3], s);
            double tb, tf;
            if (te.read(&tb, &tf) < 0) return -1.0;
            if (r > 0) {
                sb += tb;
                sf += tf;
            }
        }
        return (sf + kRebuildShare * (sb - sf)) / reps;
    };
    const Engine& E0 = *B.es[0];
    const TilePolicy start{E0.tune_split, E0.tune_stage_scale, E0.rho_apb};
    double best_ms = 0;
    const TilePolicy best = autotune_search(start, measure, &best_ms);
    for (Engine* E : B.es) {
        apply_policy(*E, best);
        E->tuned_ms = best_ms;
        launch_force_rebuild(*E, s);
    }
    BHIP_TRY(hipGetLastError());
    BHIP_TRY(hipStreamSynchronize(s));
    return RBMD_OK;
}

int rbmd_util_copy(void* dst, const void* src, int64_t nbytes, void* stream)
{
    if (!dst || !src || nbytes < 0 || (nbytes % 16) != 0) return RBMD_ERR_INVALID;
    launch_copy(dst, src, nbytes, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? RBMD_OK : RBMD_ERR_HIP;
}

}  // extern "C"
