/*
 * rbmd_force.h -- C ABI of the MI355X charged-LJ force engine.
 *
 * One handle = one force evaluator for one periodic orthorhombic system:
 * cell list + Verlet list, LJ + erfc real-space Coulomb pair pass, Random
 * Batch Ewald (or full Ewald) k-space, FP64 energy and virial.  All array
 * arguments are DEVICE pointers unless stated otherwise; streams are
 * hipStream_t passed as void* (NULL = the null stream).  Every call is
 * stream-ordered; functions documented as "synchronising" wait for the
 * stream.  Return value: RBMD_OK (0) or a negative RBMD_ERR_* code; the
 * message of the last error is available from rbmd_force_last_error().
 * A handle is not thread-safe; use one per host thread / stream.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   Force::Init / Force::Execute      rbmd-gpu-clone/rbmd-gpu/src/force/include/force.h:8-25
 *   LJCutCoulKspaceForce::Init/Execute rbmd-gpu-clone/rbmd-gpu/src/force/src/lj_cut_coul_kspace_force.cpp:37-78
 *   ExecutionNVT::ComputeAllForce     rbmd/framework/run/ExecutionNVT.cpp:185-222
 *   ExecutionNVT::Solve (VV step)     rbmd/framework/run/ExecutionNVT.cpp:133-159
 */
#ifndef RBMD_FORCE_H
#define RBMD_FORCE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBMD_OK 0
#define RBMD_ERR_INVALID (-1)   /* bad argument or parameter                    */
#define RBMD_ERR_HIP (-2)       /* HIP runtime error                            */
#define RBMD_ERR_BOX (-3)       /* box shorter than one list radius on some axis */
#define RBMD_ERR_PRECISION (-4) /* force dtype does not match handle precision */
#define RBMD_ERR_STATE (-5)     /* call order (e.g. compute before set_atoms)   */
#define RBMD_ERR_CAPACITY (-6)  /* n or P above the sizes given at create       */

#define RBMD_KSPACE_NONE 0  /* LJ + real-space only                           */
#define RBMD_KSPACE_RBE 1   /* Random Batch Ewald over P sampled k-vectors    */
#define RBMD_KSPACE_EWALD 2 /* full Ewald lattice sum (FP64 physics oracle)   */

#define RBMD_EV 1           /* compute_flags: accumulate energy and virial    */

#define RBMD_NEIGHBOR_VERLET 0 /* full Verlet list (reference "VERLET-LIST")     */
#define RBMD_NEIGHBOR_RBL 1    /* random batch list: core + sampled shell ("RBL") */

/* Parameters.  Key names follow the reference JSON (res/json_files/LJ_salt/rbmd.json:
 * cut_off, alpha, coulomb_sample_num, ...). */
typedef struct {
    double box_lo[3];   /* box origin                                         */
    double box_len[3];  /* box edge lengths                                   */
    double rc_lj;       /* LJ cutoff ("cut_off")                              */
    double rc_coul;     /* real-space Coulomb cutoff                          */
    double skin;        /* Verlet skin; list radius = max(rc_lj, rc_coul)+skin */
    double alpha;       /* Ewald splitting: erfc(sqrt(alpha) r), exp(-k^2/4alpha) */
    double qqr2e;       /* Coulomb unit factor (1 in LJ units)                */
    int32_t ntypes;     /* number of atom types (>= 1)                        */
    const double* eps;  /* HOST array [ntypes], LJ epsilon per type          */
    const double* sigma;/* HOST array [ntypes], LJ sigma per type            */
    int32_t precision;  /* 32: FP32 pair/k-space maths, 64: FP64 throughout   */
    int32_t kspace;     /* RBMD_KSPACE_*                                      */
    int32_t p_max;      /* RBE: largest P used with this handle               */
    double ewald_tol;   /* EWALD: keep k with exp(-k^2/4alpha) >= tol (e.g. 1e-16) */
    int32_t list_capacity; /* Verlet row capacity (multiple of 64); 0 = automatic */
    int32_t tile_stage;    /* LDS stage of a cell tile, in atoms (>= 64); 0 = automatic.
                              Tiles with more candidates read global memory instead. */
    /* Near-field list ("neighbor.type").  RBL (the random batch list of
     * ComputeRBLNeighboursOnceWorklet / ComputeNearForceRBLWorklet,
     * rbmd/framework/run/worklet/RunWorklet.cpp:634-708, 957-1048, with the mean-force
     * correction of ExecutionMD::ComputeRBLNearForce, run/ExecutionMD.cpp:460-566):
     * pairs closer than r_core count fully; each pair in the shell
     * r_core < r < max(rc_lj, rc_coul) is kept with probability 1/rbl_pice and then
     * counts rbl_pice times (the reference keeps every pice_num-th shell pair of its
     * sweep); then the mean of the near forces over all atoms is subtracted.  The
     * shell pairs of atom i are chosen by Philox4x32-10 (key rbl_seed, counter
     * (id_i, id_j, list build number, 0x52424C53)), word 0 mod rbl_pice == 0, drawn
     * again at every list build. */
    int32_t neighbor_type; /* RBMD_NEIGHBOR_*                                     */
    double r_core;         /* RBL core radius ("r_core")                           */
    int32_t rbl_pice;      /* RBL shell stride (pice_num, >= 1)                    */
    uint64_t rbl_seed;     /* RBL shell selection stream                           */
    /* 0 (default): every row holds all neighbours and each pair is evaluated from
     * both sides, as the reference does.  1: with FP32 maths, the full Verlet list
     * and at least three cells per axis, the pair pass evaluates every pair once
     * (half-shell rows, Newton's third law) and adds the force to both atoms (LDS
     * FP64 sums on a 2^-30 grid, order independent); otherwise as 0.  Forces agree
     * to FP32 rounding; neighbour lists and energies are the same.  The half-shell
     * pass is slower on MI355X than full rows (its rows differ in length by the
     * atom's place in its cell, and every pair adds to LDS three times), so it is
     * opt-in. */
    int32_t half_shell;
} rbmd_force_params;

/* FP64 energies (LJ units x qqr2e for Coulomb terms) and virial
 * W_ab = 1/2 sum_i sum_j (r_i - r_j)_a F_ij,b + k-space part. */
typedef struct {
    double e_lj, e_real, e_self, e_kspace;
    double virial[6]; /* xx, yy, zz, xy, xz, yz */
} rbmd_energy_virial;

typedef struct rbmd_force_s* rbmd_force_t;

/* Create a handle sized for up to n_max atoms.  Replaces Force::Init()
 * (force.h:19, lj_cut_coul_kspace_force.cpp:37-71): reads cutoffs, alpha,
 * P and unit factors, allocates the cell grid, list and k-space scratch. */
int rbmd_force_create(const rbmd_force_params* params, int64_t n_max, rbmd_force_t* out);

/* Free the handle and all device memory it owns. */
void rbmd_force_destroy(rbmd_force_t h);

/* Message of the last error on this handle (or of the last failed create
 * when h is NULL).  Never NULL. */
const char* rbmd_force_last_error(rbmd_force_t h);

/* Bind per-atom constants: charge (FP64, n) and type (int32, n, nullable =>
 * all type 0).  Replaces the DeviceData charge/type arrays the Force reads
 * (rbmd-gpu src/data_manager/include/model/device_data.h:8-156).  Forces a
 * list rebuild at the next compute.  Copies the data. */
int rbmd_force_set_atoms(rbmd_force_t h, int64_t n, const double* q, const int32_t* type,
                         void* stream);

/* Use the given k-vectors for the next computes: m (int32, 3*P, integer
 * lattice vectors, k = 2 pi m / L) and w (FP64, P, per-k estimator weight;
 * NULL => the RBE weight S/P of RBEPSAMPLE, model/RBEPSample.h:38-61).
 * Replaces the host k-list of ExecutionNVT::PreForce (run/ExecutionNVT.cpp:331-353).
 * m must not be (0,0,0). */
int rbmd_force_set_kvectors(rbmd_force_t h, int32_t P, const int32_t* m, const double* w,
                            void* stream);

/* Draw P k-vectors on the device from the exact discrete Gaussian
 * p(m) ~ exp(-k^2/(4 alpha)), m != 0 (the target of RBEPSAMPLE::Fetch_P_Sample,
 * model/RBEPSample.h:143-175), counter-based (Philox4x32-10, key = seed,
 * counter = (l, attempt, step)), weights S/P.  Deterministic in (seed, step). */
int rbmd_force_sample_kvectors(rbmd_force_t h, int32_t P, uint64_t seed, uint64_t step,
                               void* stream);

/* One force evaluation.  x, y, z: FP64 positions (n, any periodic image).
 * fx, fy, fz: float* when the handle precision is 32, double* when 64.
 * flags: RBMD_EV to also accumulate energy and virial.
 * Replaces Force::Execute() (force.h:20; lj_cut_coul_kspace_force.cpp:73-78)
 * and ExecutionNVT::ComputeAllForce (run/ExecutionNVT.cpp:185-222). */
int rbmd_force_compute(rbmd_force_t h, const double* x, const double* y, const double* z,
                       void* fx, void* fy, void* fz, int32_t flags, void* stream);

/* Energies and virial of the last compute with RBMD_EV (synchronising). */
int rbmd_force_get_energy_virial(rbmd_force_t h, rbmd_energy_virial* out, void* stream);

/* Velocity-Verlet steps with the force evaluation fused into the step:
 * kick (v += dt/2 F/m), drift (x += v dt, wrap into the box, image flags),
 * list check/rebuild, force, kick -- the order of ExecutionNVT::Solve
 * (run/ExecutionNVT.cpp:133-159) with UpdateVelocityWorklet and
 * UpdatePositionFlagWorklet (run/worklet/RunWorklet.cpp:1764-1782,1854-1877).
 * x,y,z,vx,vy,vz FP64 (n); f* as in compute and holding the forces of the
 * current positions on entry; img int32 (3n, xyz interleaved) or NULL;
 * mass FP64 (n) or NULL (=1).  With RBE, k-vectors are resampled every step
 * from (seed, step0 + s).  use_graph != 0 captures 32 steps (and the
 * nsteps % 32 remainder) in HIP graphs, cached per handle, and replays them;
 * graph and eager runs give the same bits.  A list rebuild takes its reference
 * positions from the binning one step earlier, so the lists are exact while no
 * atom moves more than skin/2 in one step; a run in which one did returns
 * RBMD_ERR_STATE (use a larger skin or a smaller dt).  Positions/velocities/forces are updated in place. */
int rbmd_md_run(rbmd_force_t h, double* x, double* y, double* z, double* vx, double* vy,
                double* vz, void* fx, void* fy, void* fz, int32_t* img, const double* mass,
                double dt, int64_t nsteps, uint64_t seed, uint64_t step0, int32_t use_graph,
                void* stream);

/* NVT velocity thermostats of rbmd_md_run ("temp_ctrl_type"), in LJ units
 * (fmt2v = mvv2e = 1).  Replaces ExecutionNVT::ComputeTempe,
 * UpdateVelocityByTempConType and TempConTypeForce (run/ExecutionNVT.cpp:749-787,
 * 256-291, 735-747) with rbmd-gpu's temperature controllers
 * (src/simulate_pipeline/src/controller/temperature_controller/temperature_controller_op/rocm/update_temperature_op.hip.cu:9-95).
 * After each step's closing kick the kinetic temperature T = sum m v^2 / (dof kB)
 * is formed on the device; then
 *   BERENDSEN    v *= sqrt(1 + dt/t_damp (t_target/T - 1))
 *   TEMP_RESCALE v *= sqrt(t_target/T)
 *   NOSE_HOOVER  v += dt/2 (F/m - xi v), then xi += dt/2 (T/t_target - 1) / (dt/10)
 *   LANGEVIN     F += -m gamma v + sqrt(2 m gamma kbt / dt) g before the closing
 *                kick, g one Gaussian 3-vector per step shared by all atoms (noise
 *                stream (seed, noise_step), advanced once per step).
 * A thermostatted run launches four kernels per step instead of three (the
 * opening kick also sums the step before's temperature), plus a one-workgroup
 * temperature reduction after the last step of each segment. */
#define RBMD_THERMO_NONE 0
#define RBMD_THERMO_BERENDSEN 1
#define RBMD_THERMO_RESCALE 2
#define RBMD_THERMO_NOSE_HOOVER 3
#define RBMD_THERMO_LANGEVIN 4
#define RBMD_THERMO_MEASURE 5   /* velocities untouched; T of every step measured and
                                   logged, as ComputeTempe does for the thermo output */
typedef struct {
    int32_t type;        /* RBMD_THERMO_*                                          */
    double t_target;     /* "temperature"[0]                                       */
    double t_damp;       /* "temperature"[2] (BERENDSEN)                           */
    double dof;          /* degrees of freedom; 0 = 3n - 3 (3n for the in-built lattice) */
    double kb;           /* Boltzmann constant; 0 = 1                              */
    double xi;           /* NOSE_HOOVER friction at the start (the reference: 0)  */
    double gamma;        /* LANGEVIN friction (the reference: 100)                 */
    double kbt;          /* LANGEVIN kB T (the reference: 1)                       */
    uint64_t seed;       /* LANGEVIN noise stream                                  */
    uint64_t noise_step; /* LANGEVIN noise step of the next step                   */
} rbmd_thermostat;
int rbmd_md_set_thermostat(rbmd_force_t h, const rbmd_thermostat* t);
/* Temperature of the last step (after its closing kick, before the rescale), the
 * Nose-Hoover friction and the next Langevin noise step.  Synchronising. */
int rbmd_md_get_thermostat(rbmd_force_t h, double* temperature, double* xi,
                           uint64_t* noise_step);
/* Device array receiving T of step j of each following thermostatted run
 * (j < len); log NULL or len 0 stops logging. */
int rbmd_md_set_temperature_log(rbmd_force_t h, double* log, int64_t len);
/* The Langevin noise vector g (3 doubles) of a noise step, on the host. */
int rbmd_langevin_noise_host(uint64_t seed, uint64_t step, double* g);

/* Current Verlet list in ORIGINAL atom ids as CSR (HOST arrays):
 * offsets[n+1], idx[cap]; rows unsorted.  An atom j appears once per periodic image
 * of j within the list radius (several times only in boxes narrower than two list
 * radii).  A random batch list holds the near band and the selected shell pairs.  Returns the entry count in *count
 * (if larger than cap only the offsets are written).  Synchronising.
 * Replaces the list of FullNeighborListBuilder::Build
 * (rbmd-gpu src/neighbor_list/src/neighbor_list_builder/full_neighbor_list_builder.cpp:31-45). */
int rbmd_force_get_neighbors(rbmd_force_t h, int64_t* offsets, int32_t* idx, int64_t cap,
                             int64_t* count);

/* k-vectors (HOST m[3P], w[P]) and structure factor (HOST re[P], im[P]) of
 * the last compute.  Synchronising.  Any output pointer may be NULL. */
int rbmd_force_get_kspace(rbmd_force_t h, int32_t* P, int32_t* m, double* w, double* re,
                          double* im);

/* Diagnostic counters (HOST out[12], synchronising): [0] list builds so far,
 * [1] list entries (full, at the list radius), [2] row capacity, [3] rows
 * that overflowed at the last build, [4] cells, [5] inner (LJ-range) entries,
 * [6] Ewald k-vectors (EWALD mode), [7] n, [8] cell tiles larger than the LDS
 * stage at the last build, [9] LDS stage (atoms), [10] largest tile (atoms),
 * [11] workgroups per tile, [12] atoms per rho(k) workgroup, [13] cost of the tile
 * shapes at the last autotune (microseconds, 0 if never tuned).  out has 14 entries. */
int rbmd_force_get_stats(rbmd_force_t h, int64_t* out);

/* Tile-shape autotuner (no counterpart in the reference, whose launch shapes are
 * fixed by VTK-m / the rbmd-gpu kernels).  Times the list build and the force pass
 * (gather, rho(k), pair pass, k-space force) with HIP events on the given
 * configuration for candidate workgroups per cell tile (1-4), LDS stage sizes and
 * rho(k) workgroup sizes, reps passes each, and keeps the fastest for
 * cost = force pass + 0.15 x list build.  Forces in fx, fy, fz are those of x, y, z
 * afterwards.  Synchronising; forces a list rebuild at the next pass.  Results: get_stats
 * [9], [11], [12], [13]. */
int rbmd_force_autotune(rbmd_force_t h, const double* x, const double* y, const double* z,
                        void* fx, void* fy, void* fz, int32_t reps, void* stream);

/* Per-kernel device time (HIP events on the launch stream).  enable != 0
 * starts recording around every launch of the named passes; get returns the
 * mean milliseconds per launch and the launch count for one pass name:
 * "check", "bin", "scan", "scatter", "cellsort", "gather", "sample", "rho",
 * "rho_final", "pair" (the tile pass: Verlet rows on rebuild steps, then the pair
 * forces), "kforce", "ev_final".  In MD runs: "check" is the kick/drift kernel
 * that opens a run segment, "rho" the rho(k) partials + cell-sort kernel, "pair"
 * the rho(k) final reduction + tile pass, "kforce" the k-space force with the
 * closing kick (and the next step's kick/drift). */
int rbmd_force_set_timing(rbmd_force_t h, int32_t enable);
int rbmd_force_get_timing(rbmd_force_t h, const char* pass, double* mean_ms, int64_t* launches);

/* Device time of the pair pass alone: two rbmd_force_compute calls at (x, y, z)
 * (the first may rebuild the rows), then reps launches of the tile pass back to
 * back between one pair of HIP events; *mean_ms = elapsed / reps, which a kernel
 * trace of the same launches reproduces (the per-launch events of
 * rbmd_force_get_timing add their own latency to every launch).  fx, fy, fz end
 * holding the forces of one more rbmd_force_compute.  No counterpart in the
 * reference: a measurement aid.  reps in [1, 10000]; synchronises the stream. */
int rbmd_force_time_pair(rbmd_force_t h, const double* x, const double* y, const double* z,
                         void* fx, void* fy, void* fz, int32_t reps, double* mean_ms,
                         void* stream);

/* ---- the k-space step without a force handle ----
 *
 * Structure factor rho(k) = sum_j q_j exp(i k.r_j) over a given k-list and the
 * k-space force F_i += q_i sum_k c_k k (cos(k.r_i) Im rho_k - sin(k.r_i) Re rho_k),
 * c_k = -4 pi w_k / (V k^2) qqr2e, with the kernels of rbmd_force_compute.  Replaces
 * ExecutionMD::ComputeRBEEleForce with ComputeNewChargeStructureFactorRBE
 * (rbmd/framework/run/ExecutionMD.cpp:389-448) and RunWorklet::ComputeNewRBEForce
 * (rbmd/framework/run/worklet/RunWorklet.cpp:1508-1548, 1616-1677).
 * x, y, z, q: FP64 device arrays (n); m: int32 device (3P, m != 0); w: FP64 device
 * (P) or NULL (the RBE weight S/P); fx, fy, fz: float* (precision 32) or double*
 * (64) device arrays the k-space force is ADDED to; rho: FP64 device (2P: Re, Im
 * per k) or NULL; ev: FP64 device (7: E_k, W_xx, W_yy, W_zz, W_xy, W_xz, W_yz) or
 * NULL.  Error text: rbmd_kspace_last_error(). */
typedef struct {
    double box_lo[3];
    double box_len[3];
    double alpha;       /* exp(-k^2 / 4 alpha) convention, as rbmd_force_params */
    double qqr2e;       /* Coulomb unit factor (1 in LJ units)               */
    int32_t precision;  /* 32: FP32 sines and force sums, 64: FP64          */
} rbmd_kspace_params;

/* Allocates its scratch (device memory) for the call and synchronises the stream. */
int rbmd_kspace_compute(const rbmd_kspace_params* p, int64_t n, const double* x, const double* y,
                        const double* z, const double* q, int32_t P, const int32_t* m,
                        const double* w, void* fx, void* fy, void* fz, double* rho, double* ev,
                        void* stream);

/* Same computation and outputs as rbmd_kspace_compute (bit for bit), with its
 * scratch in a fixed 4 MiB block of device memory per GPU instead of a hipMalloc:
 * it calls no hipMalloc / hipFree and does not synchronise (the work is enqueued on
 * `stream`).  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  P <= 8192 (larger
 * P returns RBMD_ERR_INVALID).  Every call on one GPU shares that block, so calls
 * that overlap on one GPU (other streams, other threads) race: order them on one
 * stream or synchronise between them. */
int rbmd_kspace_compute_noalloc(const rbmd_kspace_params* p, int64_t n, const double* x,
                                const double* y, const double* z, const double* q, int32_t P,
                                const int32_t* m, const double* w, void* fx, void* fy, void* fz,
                                double* rho, double* ev, void* stream);

const char* rbmd_kspace_last_error(void);

/* Plain device copy of nbytes (multiple of 16), used to calibrate HBM
 * traffic counters against a kernel of known traffic. */
int rbmd_util_copy(void* dst, const void* src, int64_t nbytes, void* stream);

/* ---- batched replicas ----
 *
 * Many independent systems of uneven size advanced together: each kernel of
 * the velocity-Verlet step (rho(k) partials + cell sort, rho(k) final + tile
 * pass, k-space force + kicks + drift + k-draw) is ONE launch over all
 * replicas, driven by a table of (replica, workgroup) work items with
 * per-replica arguments in device memory, and 32 steps (and a run's
 * remainder) are captured once in a HIP graph.  Each replica is an
 * ordinary force handle (rbmd_batch_replica) with the ensemble loop of
 * ExecutionNVT::Solve (rbmd/framework/run/ExecutionNVT.cpp:133-159) per replica;
 * a replica advanced in a batch takes bit-identical steps to the same handle
 * advanced alone by rbmd_md_run with the same seed.  The reference runs one
 * system per process and has no batch interface. */
typedef struct rbmd_batch_s* rbmd_batch_t;

/* Device pointers of one replica's MD state (as for rbmd_md_run). */
typedef struct {
    double *x, *y, *z, *vx, *vy, *vz;
    void *fx, *fy, *fz;   /* FP32 or FP64 by the replicas' precision */
    int32_t* img;         /* nullable */
    const double* mass;   /* nullable: unit masses */
} rbmd_replica_state;

/* params[nrep], n_max[nrep] HOST.  All replicas share precision, k-space mode
 * and whether they have several atom types. */
int rbmd_batch_create(int32_t nrep, const rbmd_force_params* params, const int64_t* n_max,
                      rbmd_batch_t* out);
void rbmd_batch_destroy(rbmd_batch_t b);
const char* rbmd_batch_last_error(rbmd_batch_t b);
/* Borrowed handle of replica r (set_atoms, sample_kvectors, set_proposal, compute,
 * get_* work on it as on any handle); NULL if r is out of range. */
rbmd_force_t rbmd_batch_replica(rbmd_batch_t b, int32_t r);
/* nsteps velocity-Verlet steps of every replica.  st[nrep], seeds[nrep] HOST; replica r
 * draws its k-vectors with seeds[r], step numbers step0 .. step0 + nsteps - 1.
 * Synchronising. */
int rbmd_batch_md_run(rbmd_batch_t b, const rbmd_replica_state* st, double dt, int64_t nsteps,
                      const uint64_t* seeds, uint64_t step0, int32_t use_graph, void* stream);
/* The tile-shape autotuner of rbmd_force_autotune over the batched phases: one shape
 * policy for all replicas, timed as batched launches.  Velocities are left unchanged
 * (the tuning passes kick with dt = 0); forces are recomputed.  Synchronising. */
int rbmd_batch_autotune(rbmd_batch_t b, const rbmd_replica_state* st, int32_t reps,
                        void* stream);

/* ---- learned k-space proposal ----
 *
 * Replaces the analytic k-sampler RBEPSAMPLE (rbmd/framework/model/RBEPSample.h:143-175,
 * called from ExecutionNVT::PreForce, rbmd/framework/run/ExecutionNVT.cpp:331-353) by an
 * importance proposal: with probability lambda a draw from a tabulated density on the
 * lattice box |m_d| <= Mg, otherwise a draw from the analytic sampler.  Each sampled k
 * carries w = a(m) / (P q_mix(m)), a(m) = exp(-k^2/4 alpha), in place of S/P, so the RBE
 * force and energy estimators stay unbiased.  The reference has no learned sampler. */
typedef struct rbmd_proposal_s* rbmd_proposal_t;

/* density: HOST array of (2 Mg + 1)^3 values >= 0, x fastest, m_d = -Mg..Mg (the entry
 * at m = 0 is ignored).  lambda in (0, 1).  alpha and box_len[3] must be those of the
 * force handles the proposal is used with.  Returns RBMD_ERR_INVALID for a density
 * that is negative, non-finite or all zero. */
int rbmd_proposal_create(int32_t Mg, const double* density, double lambda, double alpha,
                         const double* box_len, rbmd_proposal_t* out);
void rbmd_proposal_destroy(rbmd_proposal_t p);
/* The proposal's draws on the host (bit-identical m and w to the device sampler of a
 * handle using it): m[3P], w[P] HOST. */
int rbmd_proposal_sample_host(rbmd_proposal_t p, int32_t P, uint64_t seed, uint64_t step,
                              int32_t* m, double* w);
/* Use the proposal for all later sample_kvectors / md_run draws of this handle
 * (p == NULL: back to the analytic sampler).  Copies the tables to the device
 * (synchronising); the proposal may be destroyed afterwards. */
int rbmd_force_set_proposal(rbmd_force_t h, rbmd_proposal_t p);

/* ---- proposal network: 4 -> 64 -> 64 -> 1 MLP, SiLU, softplus output ----
 *
 * No counterpart in the reference.  Features (|kx|, |ky|, |kz|, |k|) / k_scale in FP32,
 * BF16 weights with FP32 accumulation in the forward pass, FP32 master weights updated
 * by Adam.  Parameter vector (rbmd_mlp_nparams() = 4545 FP32 values): W1[64][4], b1[64],
 * W2[64][64], b2[64], w3[64], b3.  Training loss over a batch: sum_s w_s (log h_s - y_s)^2. */
typedef struct rbmd_mlp_s* rbmd_mlp_t;
int rbmd_mlp_nparams(void);
/* Glorot-uniform weights from Philox(seed), zero biases, b3 = log(e - 1). */
int rbmd_mlp_create(uint64_t seed, rbmd_mlp_t* out);
void rbmd_mlp_destroy(rbmd_mlp_t h);
const char* rbmd_mlp_last_error(rbmd_mlp_t h);
/* FP32 master weights, HOST arrays of rbmd_mlp_nparams() values (synchronising).
 * set also resets the Adam moments. */
int rbmd_mlp_get_params(rbmd_mlp_t h, float* params);
int rbmd_mlp_set_params(rbmd_mlp_t h, const float* params);
/* h(f) for feat[4B] (DEVICE) into out[B] (DEVICE). */
int rbmd_mlp_forward(rbmd_mlp_t h, const float* feat, int64_t B, float* out, void* stream);
/* Loss and its gradient w.r.t. the parameters at the BF16 weights (synchronising);
 * feat[4B], y[B], w[B] DEVICE; loss, grad[nparams] HOST (either may be NULL). */
int rbmd_mlp_loss_grad(rbmd_mlp_t h, const float* feat, const float* y, const float* w,
                       int64_t B, double* loss, float* grad, void* stream);
/* iters full-batch Adam steps (beta 0.9/0.999, eps 1e-8) at learning rate lr;
 * final_loss (HOST, nullable, synchronising) = loss after the last step. */
int rbmd_mlp_train(rbmd_mlp_t h, const float* feat, const float* y, const float* w, int64_t B,
                   int32_t iters, double lr, double* final_loss, void* stream);

/* ---- host-side system generators (CPU, deterministic, Philox4x32-10) ---- */

/* Philox4x32-10 block (Salmon et al., SC'11): out[4] for counter ctr[4], key[2]. */
void rbmd_philox4x32(const uint32_t* ctr, const uint32_t* key, uint32_t* out);

/* FCC lattice of nc^3 unit cells (4 atoms each) at number density rho,
 * box [0, L)^3, L = nc (4/rho)^(1/3).  Writes 4 nc^3 FP64 positions (HOST). */
int rbmd_sys_fcc(int32_t nc, double rho, double* x, double* y, double* z, double* L_out);

/* n points uniform in [0,L)^3 with minimum-image separation >= dmin,
 * by seeded random sequential addition.  HOST arrays.  Returns RBMD_OK or
 * RBMD_ERR_INVALID if the packing could not be completed. */
int rbmd_sys_random(int64_t n, double L, double dmin, uint64_t seed, double* x, double* y,
                    double* z);

/* Charges: first half +1, second half -1 of a seeded random permutation. */
int rbmd_sys_charges(int64_t n, uint64_t seed, double* q);

/* Maxwell-Boltzmann velocities at temperature T (k_B = 1), zero total
 * momentum, rescaled so that sum m v^2 / (3n - 3) = T exactly.  mass NULL => 1. */
int rbmd_sys_velocities(int64_t n, double T, const double* mass, uint64_t seed, double* vx,
                        double* vy, double* vz);

/* The device k-sampler, run on the host (same draws, bit-identical m):
 * m[3P] (HOST), returns S in *S_out (nullable). */
int rbmd_sample_kvectors_host(double alpha, const double* L, int32_t P, uint64_t seed,
                              uint64_t step, int32_t* m, double* S_out);

/* ------------------------------------------------------------------ */
/* Input and output formats (host, no GPU)                              */
/* ------------------------------------------------------------------ */

/* LAMMPS data files as the reference reads them with "read_data"
 * (rbmd-gpu src/reader/src/atomic_reader.cpp:24-58, 135-230, 322-350).
 * rbmd_read_data_info reads the header (atom count, types, box) and which of
 * the Masses / Pair Coeffs / Velocities sections are present. */
typedef struct {
    int64_t natoms;
    int32_t ntypes;
    double lo[3], hi[3];
    int32_t has_masses, has_pair_coeffs, has_velocities;
} rbmd_data_info;
int rbmd_read_data_info(const char* path, rbmd_data_info* info);
/* Atoms lines of atom_style "atomic" (id type x y z), "charge" (id type q x y z)
 * or "full" (id mol type q x y z), each with optional image flags: the atom with
 * id k fills index k - 1 of type (0-based), q, x, y, z and img (3n, xyz
 * interleaved); Velocities lines (id vx vy vz) fill vx, vy, vz; Masses and Pair
 * Coeffs (type eps sigma) fill mass, eps, sigma per type.  HOST arrays; any but
 * type/x/y/z may be NULL.  n must equal the file's atom count and every id in
 * 1..n must appear once. */
int rbmd_read_data(const char* path, const char* atom_style, int64_t n, int32_t* type,
                   double* q, double* x, double* y, double* z, int32_t* img, double* vx,
                   double* vy, double* vz, double* mass, double* eps, double* sigma);
/* One trajectory frame in the reference's dump layout (TrajectoryOutput::Execute,
 * rbmd-parallel/framework/output/TrajectoryOutput.cpp:21-44): ITEM: TIMESTEP,
 * NUMBER OF ATOMS, BOX BOUNDS pp pp pp, ATOMS id type x y z (ids and types
 * 1-based), positions to 17 significant digits.  append != 0 adds the frame. */
int rbmd_write_dump(const char* path, int32_t append, int64_t step, int64_t n, const double* lo,
                    const double* hi, const int32_t* type, const double* x, const double* y,
                    const double* z);
/* Message of the last failed I/O call of this thread. */
const char* rbmd_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* RBMD_FORCE_H */
