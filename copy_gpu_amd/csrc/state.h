// state.h -- host-side engine state (device buffers, parameters) and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "engine.h"
#include "launch.h"
#include "../../include/rbmd_force.h"

namespace rbmd {

// Per-k coefficients of the k-space force: phase from the integer lattice
// vector m (FP64) and k = 2 pi m / L, A = c Im(rho), B = c Re(rho) with
// c = w * (-4 pi) / (V k^2) * qqr2e.
template <typename T>
struct KCoef {
    double m[3];
    T k[3];
    T A, B;
};

// k-sampler arguments of the fused kick/drift kernel (P = 0: no sampling).
struct SampleArgs {
    int P = 0;
    uint64_t seed = 0;
    uint64_t step = 0;
    const unsigned long long* step_dev = nullptr;  // overrides step when set
    const double* cdf = nullptr;
    int3 M{};
    double w = 0;       // analytic weight S / P
    double w_out = 0;   // weight of a mixture draw outside the grid
    GridProposal gp{};  // learned proposal (gp.cdf == nullptr: analytic only)
    int32_t* km = nullptr;
    double* kw = nullptr;
};

// MD steps captured in one graph.  Launching a graph leaves a gap of about 8 us
// on the device and every segment opens and closes with its own kick kernels
// (about 4 us more than a fused step), so segments are long: a run of K steps
// replays the kGraphSteps-step graph K / kGraphSteps times and one graph of the
// remaining K % kGraphSteps steps.
constexpr int kGraphSteps = 32;

struct Engine {
    // ---- parameters
    Box box{};
    double rc_lj = 0, rc_coul = 0, skin = 0, alpha = 0, qqr2e = 1, ewald_tol = 1e-16;
    double r_list = 0, r_in = 0;
    // random batch list (neighbor type RBL): core radius, shell stride, selection seed
    bool rbl = false;
    double r_core = 0;
    int rbl_pice = 1;
    uint64_t rbl_seed = 0;
    double* rbl_part = nullptr;  // [tile workgroups][3] near-force sums of each tile
    double* rbl_mean = nullptr;  // [4] mean near force over all atoms
    // half-shell pair pass (FP32, full Verlet list, at least three cells per axis):
    // every pair evaluated once; force planes [14][3][n_max] summed per atom
    bool half = false;
    int acc_cap = 0;             // LDS force sums of a half tile (candidates + sentinel)
    float* planes = nullptr;
    int precision = 32, kspace = 0, p_max = 0;
    PairParams pp{};
    int64_t n_max = 0, n = 0;
    int dims[3] = {0, 0, 0};
    double cell[3] = {0, 0, 0};
    int ncell = 0;
    int cap = 0;        // list row capacity, multiple of 64
    int split = 1;      // workgroups per cell tile
    int stage_cap = 0;  // LDS tile capacity (candidates)
    int stage_override = 0;
    // tile shapes chosen by the autotuner (0 / default: heuristics)
    int tune_split = 0;              // workgroups per cell tile
    // LDS stage = scale * mean stencil population + 64; 0: the largest stage that
    // keeps the most workgroups per CU for a tile 5 sigma above the mean
    double tune_stage_scale = 0.0;
    int rho_apb = 0;                 // atoms per rho(k) workgroup (multiple of 64)
    double tuned_ms = 0;             // cost of the chosen shapes at the last autotune
    int lds_max = 0;    // device LDS bytes per workgroup
    int lds_cu = 0;     // device LDS bytes per CU
    int num_cu = 0;
    double e_self = 0;
    bool atoms_set = false;

    // ---- device buffers
    double *xr = nullptr, *yr = nullptr, *zr = nullptr;
    double4 *xs_ref = nullptr, *xs = nullptr;
    int32_t *perm = nullptr, *scell = nullptr, *ts = nullptr;
    int32_t *cell_count = nullptr, *cell_start = nullptr, *atom_cell = nullptr,
            *atom_rank = nullptr;
    // tile workgroup -> (cell * split + part): load-balanced when the tiles fit about
    // one dispatch round (rebuilt with the cells; identity until then)
    int32_t* tile_order = nullptr;
    // bins of the MD step's next rebuild (double-buffered by the epoch flags[3])
    int32_t* md_count = nullptr;   // [2][ncell] per-cell counts, then [2][nchunk] chunk counts
    int32_t* md_bucket = nullptr;  // [2][ncell][bucket_cap]
    int32_t* md_cell = nullptr;    // [n] by slot at binning
    int32_t* md_vid = nullptr;     // [n] by slot at binning: atom id
    double* md_w = nullptr;        // [3][n] by slot: wrapped positions at binning
    int bucket_cap = 0;
    uint16_t* nlist = nullptr;  // tile-local indices
    int2* ncount = nullptr;
    uint32_t* flags = nullptr;  // [0] rebuild request, [1] NaN seen, [2] rebuilding this step,
                                // [3] step epoch of the MD bins
    unsigned long long* stats = nullptr;    // [0] builds [1] entries [2] overflow rows [3] inner [4] longest overflowing row [6] big tiles [8..10] list stats
    double* q = nullptr;
    int32_t* type = nullptr;
    PairParams* d_pp = nullptr;

    // k-space
    int P = 0;            // current number of k-vectors
    int kcap = 0;         // allocated k-vector slots
    int32_t* km = nullptr;
    double* kw = nullptr;
    double* kpart = nullptr;  // [blocks_x][P][2] partial structure factors
    int64_t kpart_cap = 0;    // doubles
    double* rho = nullptr;    // [P][2]
    void* kcoef = nullptr;    // KCoef<T>[P]
    double* kev = nullptr;    // [P][7] per-k energy / virial terms
    double* kcdf = nullptr;   // sampler CDF tables
    int kM[3] = {0, 0, 0};
    std::vector<double> h_kcdf;
    double S = 0;             // RBE normaliser H_x H_y H_z - 1
    unsigned long long* d_step = nullptr;  // device step counter for graph replays
    bool kvec_weights_default = true;
    // learned proposal tables (device), active when gp.cdf != nullptr
    GridProposal gp{};
    double* gp_cdf = nullptr;
    double* gp_ratio = nullptr;
    int32_t* gp_guide = nullptr;
    int64_t gp_cap = 0;  // entries allocated
    int64_t gp_guide_cap = 0;
    uint64_t gp_gen = 0; // bumped by every set_proposal (keys the MD graph)

    // energy / virial
    double* ev_part = nullptr;  // [blocks][8]
    int ev_part_blocks = 0;
    double* ev = nullptr;       // [10]

    // NVT thermostat of rbmd_md_run (rbmd_md_set_thermostat)
    int thermo_type = 0;
    double thermo_t0 = 0, thermo_tdamp = 0, thermo_dof = 0, thermo_kb = 1;
    double thermo_gamma = 0, thermo_kbt = 0;
    uint64_t thermo_seed = 0;
    double* thermo_part = nullptr;    // [thermo_part_cap] closing-kick workgroup partials
    int64_t thermo_part_cap = 0;
    double* thermo_st = nullptr;      // [4] scale, xi, last T, noise step (as counter below)
    uint32_t* thermo_cnt = nullptr;   // [1] temperature log position
    unsigned long long* thermo_noise = nullptr;  // [1] Langevin noise step
    double* thermo_log = nullptr;     // caller's device array of T per step, or null
    int64_t thermo_log_len = 0;
    uint64_t thermo_gen = 0;          // bumped by every thermostat change (keys the MD graph)

    // host pinned mirror of stats
    unsigned long long* h_stats = nullptr;

    hipStream_t own_stream = nullptr;
    hipEvent_t order_event = nullptr;  // orders own_stream after the caller's null stream

    // timing
    bool timing = false;
    struct Rec {
        hipEvent_t a, b;
    };
    std::vector<Rec> pending[P_NUM];
    double t_sum[P_NUM] = {};
    int64_t t_cnt[P_NUM] = {};
    std::vector<hipEvent_t> event_pool;

    // MD graph cache: [0] kGraphSteps steps, [r] r steps (a run's remainder)
    struct StepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<const void*> key;
    };
    StepGraph graphs[kGraphSteps];
    bool capturing = false;

    std::string err;
};

// ---- batched replicas: device pointers of one replica's MD state
struct RepState {
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    void *fx = nullptr, *fy = nullptr, *fz = nullptr;
    int32_t* img = nullptr;
    const double* mass = nullptr;
};

// Words per replica in a batch's readback block: the 8 build statistics, then the
// flags flags[0..4] one per word.
constexpr int kStatsOutWords = 16;

// Phases of the batched MD step, one launch each over all replicas.
enum BatchPh {
    BP_KICKDRIFT = 0,  // segment opening kick/drift (md.hip)
    BP_BIN,            // force-pass list build (nlist.hip)
    BP_SCAN,
    BP_SCATTER,
    BP_CELLSORT,
    BP_MDCLEAR,        // MD bin counts cleared before BP_MDBIN (nlist.hip)
    BP_MDBIN,
    BP_CHUNKSUM,       // chunk counts of the MD bins after BP_MDBIN (nlist.hip)
    BP_GATHER,
    BP_RHO,            // rho partials of the force pass (kspace.hip)
    BP_SORT_RHO,       // MD step: rho partials + cell sort (md.hip)
    BP_PAIR,           // rho final + tile pass (pair.hip)
    BP_KFORCE,         // k-space force + closing kick (md.hip)
    BP_KFORCE_DRIFT,   // ... fused with the next step's opening kick/drift (md.hip)
    BP_THERMO_BEGIN,   // thermostatted replicas: state reset before the first step
    BP_THERMO_REDUCE,  // ... temperature and coefficient after each closing kick
    BP_THERMO_END,     // ... last velocity scale after the last step
    BP_STATS_OUT,      // every replica's build statistics and flags into one block (nlist.hip)
    BP_NUM
};

struct Batch {
    std::vector<Engine*> es;
    std::vector<RepState> st;
    std::vector<SampleArgs> sa;  // per replica, device step counters
    double dt = 0;
    int precision = 32;
    int kspace = 0;
    bool multi = false;
    bool thermo = false;  // some replica has a thermostat: one segment per step
    unsigned long long step0 = 0;  // first step of the run (MdBin sets the counters)
    // per replica kStatsOutWords words: the 8 build statistics, then flags[0..4];
    // device block and its pinned host mirror, one copy per run
    unsigned long long* d_out = nullptr;
    unsigned long long* h_out = nullptr;
    size_t out_cap = 0;  // replicas the blocks hold
    BatchPhase ph[BP_NUM];
    // the tile pass runs one launch per LDS class of replicas: ph[BP_PAIR] holds
    // the first class, pair_more the others (pair.hip pair_prepare_t)
    static constexpr int kPairClasses = 4;
    BatchPhase pair_more[kPairClasses - 1];
};

// Batched phases: prepare packs and uploads every replica's arguments (host,
// synchronous, outside graph capture); launch issues one grid per phase.
hipError_t nlist_batch_prepare(Batch& b);
void nlist_batch_launch(Batch& b, int phase, hipStream_t s);
hipError_t pair_batch_prepare(Batch& b);
void pair_batch_launch(Batch& b, int phase, hipStream_t s);
hipError_t kspace_batch_prepare(Batch& b);
void kspace_batch_launch(Batch& b, int phase, hipStream_t s);
hipError_t md_batch_prepare(Batch& b);
void md_batch_launch(Batch& b, int phase, hipStream_t s);

inline bool has_kspace(const Engine& e)
{
    return (e.kspace == RBMD_KSPACE_RBE || e.kspace == RBMD_KSPACE_EWALD) && e.P > 0;
}

// Degrees of freedom of the kinetic temperature: the thermostat's own, or 3n - 3
// (ExecutionNVT::ComputeTempe, rbmd/framework/run/ExecutionNVT.cpp:772-775).
inline double thermo_dof(const Engine& e)
{
    return e.thermo_dof > 0 ? e.thermo_dof : 3.0 * (double)e.n - 3.0;
}

inline int md_nchunk(int ncell)
{
    return ncell > kBinChunkMinCells ? (ncell + kBinChunk - 1) / kBinChunk : 0;
}

inline MdBins md_bins(const Engine& e)
{
    const size_t n = (size_t)std::max<int64_t>(e.n_max, 1);
    return MdBins{e.md_count, e.md_count + 2 * (size_t)e.ncell, e.md_bucket, e.md_cell,
                  e.md_vid,   e.md_w, e.md_w + n, e.md_w + 2 * n, e.ncell, e.bucket_cap,
                  md_nchunk(e.ncell)};
}

// Bytes of the MD bin counts: per-cell and per-chunk, both double-buffered.
inline size_t md_count_bytes(int ncell)
{
    return (size_t)2 * (size_t)(ncell + md_nchunk(ncell)) * sizeof(int32_t);
}

// ---- launchers (nlist.hip)
void launch_check(Engine& e, const double* x, const double* y, const double* z, hipStream_t s);
void launch_force_rebuild(Engine& e, hipStream_t s);
void launch_build(Engine& e, const double* x, const double* y, const double* z, hipStream_t s);
void launch_gather(Engine& e, const double* x, const double* y, const double* z, hipStream_t s);
void launch_md_bin(Engine& e, const double* x, const double* y, const double* z,
                   unsigned long long step0, hipStream_t s);
void launch_list_stats(Engine& e, unsigned long long* out, hipStream_t s);
void launch_copy(void* dst, const void* src, int64_t nbytes, hipStream_t s);

// ---- launchers (pair.hip)
void launch_pair(Engine& e, void* fx, void* fy, void* fz, bool ev, hipStream_t s);
void launch_tile_rho(Engine& e, void* fx, void* fy, void* fz, hipStream_t s);
void launch_ev_final(Engine& e, bool with_pair, hipStream_t s);
// random batch list: the mean near force the next k-space / kick kernel subtracts
void launch_rbl_mean(Engine& e, hipStream_t s);

// ---- launchers (kspace.hip)
SampleArgs sample_args(const Engine& e, int P, uint64_t seed, const uint64_t* step_host);
void rho_geometry(const Engine& e, int* atoms_per_block, int* nbx, int* nkt);
void launch_sample(Engine& e, int P, uint64_t seed, const uint64_t* step_host, hipStream_t s);
void launch_kweights_default(Engine& e, hipStream_t s);
void launch_rho(Engine& e, const double* x, const double* y, const double* z, hipStream_t s);
void launch_kforce(Engine& e, const double* x, const double* y, const double* z, void* fx,
                   void* fy, void* fz, hipStream_t s);

// ---- launchers (md.hip)
// fuse_thermo: this opening kick also reduces the temperature of the step before
// (whose closing kick then needs no ThermoReduce launch)
void launch_md_open(Engine& e, const RepState& st, double dt, const SampleArgs& sa,
                    hipStream_t s, bool fuse_thermo = false);
void launch_md_sort_rho(Engine& e, const RepState& st, hipStream_t s);
void launch_md_close(Engine& e, const RepState& st, double dt, hipStream_t s);
// thermostatted runs: state reset before the first step, last velocity scale after
// the last (no kernel without a rescaling thermostat)
void launch_thermo_begin(Engine& e, hipStream_t s);
// after each closing kick: the kinetic temperature and the thermostat's coefficient
void launch_thermo_reduce(Engine& e, double dt, hipStream_t s);
void launch_thermo_end(Engine& e, const RepState& st, hipStream_t s);
void launch_md_close_open(Engine& e, const RepState& st, double dt, const SampleArgs& sa,
                          hipStream_t s);

// timing helpers (api.hip)
void timer_begin(Engine& e, int pass, hipStream_t s, hipEvent_t* ev);
void timer_end(Engine& e, int pass, hipStream_t s, hipEvent_t ev);

struct ScopedTimer {
    Engine& e;
    int pass;
    hipStream_t s;
    hipEvent_t a = nullptr;
    ScopedTimer(Engine& e_, int p, hipStream_t s_) : e(e_), pass(p), s(s_) { timer_begin(e, pass, s, &a); }
    ~ScopedTimer() { timer_end(e, pass, s, a); }
};

}  // namespace rbmd
