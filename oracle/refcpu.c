/*
 * oracle/refcpu.c -- CPU oracle for the charged-LJ force evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke test
 * and the cpu_baseline leg of bench.py load it; the product library never
 * links, loads or calls it.
 *
 * It restates, in plain C and FP64, what the reference RBMD CPU path
 * (VTK-m Serial/OpenMP worklets) computes for one force evaluation of a
 * charged Lennard-Jones system.  Every function names the reference
 * file:line it follows (paths relative to the reference checkout).  The
 * reference-faithful FP32 step used as the timed CPU baseline lives in
 * faithful.c.
 *
 * Deliberate, documented extensions of the reference semantics:
 *   - two cutoffs (rc_lj for LJ, rc_coul for the erfc real-space term); the
 *     reference uses one cut_off for both (run/ExecutionNVT.cpp:793-797);
 *   - a Verlet skin (the reference rebuilds at rc every force call);
 *   - closed-form erfc only (the reference's Gnear table is alpha=0.1 only);
 *   - the sampled k-vectors are an input (the reference draws them from
 *     std::random_device, math/Utiles.h:32-39);
 *   - FP64 energy and a 6-component virial.
 *
 * Pair-distance convention shared bit-for-bit with the GPU list builder, in
 * the reference's own order (RunWorklet.cpp:77-79: p_j = position(j) -
 * coord_offset, r_ij = p_j - p_i):
 *   k  = rint((x_j - x_i) / L)   (image index, in {-1,0,1} for wrapped input)
 *   dx = (x_j - k*L) - x_i       (k*L is exact)
 *   d2 = fma(dz,dz, fma(dy,dy, dx*dx))
 * and the neighbour criterion of RunWorklet.cpp:83
 *   (r_list^2 - d2 > 1e-4) && (d2 > 1e-4).
 * Built with -ffp-contract=off so no other contraction changes the bits.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* periodic helpers                                                     */
/* ------------------------------------------------------------------ */

/* Minimum-image displacement x_j - x_i along one axis.  Follows
 * ExecPointLocator::MinDistanceVec (locator/ExecPointLocator.h:225-260): a
 * single shift by the box length when |x_j - x_i| > L/2, written with rint so
 * that it also covers |x_j - x_i| > L; the shift is applied to x_j before the
 * difference, as the reference's list build does (RunWorklet.cpp:77-79). */
static inline double orc_mi(double xj, double xi, double L)
{
    double k = nearbyint((xj - xi) / L);
    return (xj - k * L) - xi;
}

static inline double orc_d2(double dx, double dy, double dz)
{
    return fma(dz, dz, fma(dy, dy, dx * dx));
}

/* Wrap positions into [lo, lo+L).  Follows UpdateFlagOverRangePoint
 * (locator/ExecPointLocator.h:341-357), generalised to several periods:
 * x <- x - L*floor((x-lo)/L), image flag += floor((x-lo)/L). */
void orc_wrap(int64_t n, double* x, double* y, double* z, const double* lo, const double* L,
              int32_t* img /* 3n or NULL */)
{
    double* c[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        for (int64_t i = 0; i < n; ++i) {
            double f = floor((c[d][i] - lo[d]) / L[d]);
            c[d][i] = c[d][i] - f * L[d];
            if (img) img[3 * i + d] += (int32_t)f;
        }
    }
}

/* ------------------------------------------------------------------ */
/* neighbour sets                                                       */
/* ------------------------------------------------------------------ */

static int cmp_i32(const void* a, const void* b)
{
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* All-pairs O(N^2) full neighbour list.  Set semantics of
 * ComputeNeighboursWorklet (run/worklet/RunWorklet.cpp:40-96): j is a
 * neighbour of i iff r_list^2 - d2 > 1e-4 and d2 > 1e-4 under the minimum
 * image.  Output: CSR, each row sorted ascending.  Returns the entry count,
 * or -(needed) when cap is too small. */
int64_t orc_neighbors_brute(int64_t n, const double* x, const double* y, const double* z,
                            const double* L, double r_list, int64_t* offsets, int32_t* idx,
                            int64_t cap)
{
    const double rl2 = r_list * r_list;
    int64_t e = 0;
    offsets[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            double dx = orc_mi(x[j], x[i], L[0]);
            double dy = orc_mi(y[j], y[i], L[1]);
            double dz = orc_mi(z[j], z[i], L[2]);
            double d2 = orc_d2(dx, dy, dz);
            if ((rl2 - d2) > 1e-4 && d2 > 1e-4) {
                if (e < cap) idx[e] = (int32_t)j;
                ++e;
            }
        }
        offsets[i + 1] = e;
    }
    return e <= cap ? e : -e;
}

/* All-pairs neighbour list over periodic images, for boxes narrower than two list
 * radii: j (i itself included) appears once for every shift k L, k in {-1,0,1}^3,
 * whose image satisfies the criterion above with dx = (x_j - k L) - x_i (the
 * reference's PeriodicCoordOffset order, RunWorklet.cpp:77-79).  The 27-cell stencil
 * of a grid with cells at least r_list wide (ContPointLocator.h:44-53) visits exactly
 * these images when the box is at least r_list wide.  Positions must be wrapped. */
int64_t orc_neighbors_images(int64_t n, const double* x, const double* y, const double* z,
                             const double* L, double r_list, int64_t* offsets, int32_t* idx,
                             int64_t cap)
{
    const double rl2 = r_list * r_list;
    int64_t e = 0;
    offsets[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            for (int s = 0; s < 27; ++s) {
                const double kx = (double)(s / 9 - 1), ky = (double)((s / 3) % 3 - 1),
                             kz = (double)(s % 3 - 1);
                double dx = (x[j] - kx * L[0]) - x[i];
                double dy = (y[j] - ky * L[1]) - y[i];
                double dz = (z[j] - kz * L[2]) - z[i];
                double d2 = orc_d2(dx, dy, dz);
                if ((rl2 - d2) > 1e-4 && d2 > 1e-4) {
                    if (e < cap) idx[e] = (int32_t)j;
                    ++e;
                }
            }
        }
        offsets[i + 1] = e;
    }
    return e <= cap ? e : -e;
}

/* Cell grid of the reference locator: dims = trunc(L / r) per axis
 * (ContPointLocator::SetRange, locator/ContPointLocator.h:44-53, with
 * num_cycles = 1), cell of a point = trunc((p - lo)/dxdydz) clamped to the
 * grid (BinPointsWorklet, locator/ContPointLocator.cxx:38-60), points grouped
 * by a sort on cell id with lower/upper bounds (ContPointLocator.cxx:70-92). */
typedef struct {
    int dims[3];
    double cell[3];
    int64_t ncell;
    int64_t* start; /* ncell+1 */
    int32_t* ids;   /* n, grouped by cell, ascending id inside a cell */
} orc_cells;

static void orc_cells_build(orc_cells* c, int64_t n, const double* x, const double* y,
                            const double* z, const double* lo, const double* L, double r)
{
    const double* p[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        c->dims[d] = (int)(L[d] / r);
        if (c->dims[d] < 1) c->dims[d] = 1;
        c->cell[d] = L[d] / c->dims[d];
    }
    c->ncell = (int64_t)c->dims[0] * c->dims[1] * c->dims[2];
    c->start = (int64_t*)calloc((size_t)c->ncell + 1, sizeof(int64_t));
    c->ids = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int64_t* cid = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        int64_t ijk[3];
        for (int d = 0; d < 3; ++d) {
            int64_t v = (int64_t)((p[d][i] - lo[d]) / c->cell[d]);
            if (v < 0) v = 0;
            if (v > c->dims[d] - 1) v = c->dims[d] - 1;
            ijk[d] = v;
        }
        cid[i] = ijk[0] + ijk[1] * c->dims[0] + ijk[2] * (int64_t)c->dims[0] * c->dims[1];
        c->start[cid[i] + 1]++;
    }
    for (int64_t k = 0; k < c->ncell; ++k) c->start[k + 1] += c->start[k];
    int64_t* fill = (int64_t*)malloc((size_t)c->ncell * sizeof(int64_t));
    memcpy(fill, c->start, (size_t)c->ncell * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) c->ids[fill[cid[i]]++] = (int32_t)i; /* stable: ascending i */
    free(fill);
    free(cid);
}

static void orc_cells_free(orc_cells* c)
{
    free(c->start);
    free(c->ids);
}

static inline int64_t orc_cell_of(const orc_cells* c, const double* lo, double px, double py,
                                  double pz, int ijk[3])
{
    double p[3] = {px, py, pz};
    for (int d = 0; d < 3; ++d) {
        int64_t v = (int64_t)((p[d] - lo[d]) / c->cell[d]);
        if (v < 0) v = 0;
        if (v > c->dims[d] - 1) v = c->dims[d] - 1;
        ijk[d] = (int)v;
    }
    return ijk[0] + ijk[1] * (int64_t)c->dims[0] + ijk[2] * (int64_t)c->dims[0] * c->dims[1];
}

/* Cell-stencil full list (27 cells, periodic index offsets as in
 * ExecPointLocator::PeriodicIndexOffset, locator/ExecPointLocator.h:303-319).
 * Positions must be wrapped into the box and every axis must have >= 3
 * cells; returns -1 otherwise.  Same set and output format as the brute
 * force version.  Returns entry count or -(needed). */
int64_t orc_neighbors_cells(int64_t n, const double* x, const double* y, const double* z,
                            const double* lo, const double* L, double r_list, int64_t* offsets,
                            int32_t* idx, int64_t cap)
{
    orc_cells c;
    orc_cells_build(&c, n, x, y, z, lo, L, r_list);
    if (c.dims[0] < 3 || c.dims[1] < 3 || c.dims[2] < 3) {
        orc_cells_free(&c);
        return -1;
    }
    const double rl2 = r_list * r_list;
    int64_t e = 0;
    offsets[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        int ijk[3];
        orc_cell_of(&c, lo, x[i], y[i], z[i], ijk);
        int64_t row0 = e;
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
                for (int g = -1; g <= 1; ++g) {
                    int cx = (ijk[0] + a + c.dims[0]) % c.dims[0];
                    int cy = (ijk[1] + b + c.dims[1]) % c.dims[1];
                    int cz = (ijk[2] + g + c.dims[2]) % c.dims[2];
                    int64_t cc = cx + cy * (int64_t)c.dims[0] + cz * (int64_t)c.dims[0] * c.dims[1];
                    for (int64_t s = c.start[cc]; s < c.start[cc + 1]; ++s) {
                        int32_t j = c.ids[s];
                        double dx = orc_mi(x[j], x[i], L[0]);
                        double dy = orc_mi(y[j], y[i], L[1]);
                        double dz = orc_mi(z[j], z[i], L[2]);
                        double d2 = orc_d2(dx, dy, dz);
                        if ((rl2 - d2) > 1e-4 && d2 > 1e-4) {
                            if (e < cap) idx[e] = j;
                            ++e;
                        }
                    }
                }
        if (e <= cap) qsort(idx + row0, (size_t)(e - row0), sizeof(int32_t), cmp_i32);
        offsets[i + 1] = e;
    }
    orc_cells_free(&c);
    return e <= cap ? e : -e;
}

/* ------------------------------------------------------------------ */
/* pair forces: LJ + erfc real-space Coulomb                            */
/* ------------------------------------------------------------------ */

/* G(alpha, r) of ExecForceFunction::Gnear (forceFunction/ExecForceFunction.h:217-224):
 * erfc(sqrt(a) r)/r^2 + 2 sqrt(a) exp(-a r^2) / (sqrt(pi) r). */
double orc_gnear(double alpha, double r)
{
    double sa = sqrt(alpha);
    return erfc(sa * r) / (r * r) + 2.0 * sa * exp(-alpha * r * r) / (sqrt(ORC_PI) * r);
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of
 * (hi0, lo0) = 0xD2511F53 c0, (hi1, lo1) = 0xCD9E8D57 c2,
 * c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0), key += (0x9E3779B9, 0xBB67AE85).
 * Written here on its own (the product library has its own copy); pinned by the
 * Random123 known-answer vectors. */
void orc_philox4x32(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
        const uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        c0 = h1 ^ c1 ^ k0;
        c1 = l1;
        c2 = h0 ^ c3 ^ k1;
        c3 = l0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Random batch list shell selection of the pair (i, j) (atom ids, i the atom whose
 * list holds j) at list build `epoch`: Philox4x32-10 word 0 of counter
 * (i, j, epoch, 0x52424C53) under key (seed_lo, seed_hi), divisible by pice.  The
 * reference keeps every pice_num-th shell pair of its stencil sweep
 * (ComputeRBLNeighboursOnceWorklet, run/worklet/RunWorklet.cpp:690-697); the
 * counter-based draw replaces that order, which the cell sort's ties decide. */
int orc_rbl_select(uint64_t seed, uint32_t epoch, int32_t i, int32_t j, uint32_t pice)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)j, epoch, 0x52424C53u};
    uint32_t r[4];
    orc_philox4x32(ctr, key, r);
    return r[0] % pice == 0u;
}

/* Per-pair weight of the near force: 1 for the full list; for a random batch
 * list (rbl != NULL) the reference's core / shell criteria
 * (ComputeRBLNeighboursOnceWorklet, RunWorklet.cpp:686-703, and the pice_num factor
 * of ComputeNearForceRBLWorklet, :1015-1022): 1 when r_core^2 - d2 > 1e-4 and
 * d2 > 1e-4, pice for a selected pair with rc^2 - d2 > 1e-4 and d2 - r_core^2 > 1e-4
 * (rc = the larger cutoff), 0 otherwise. */
typedef struct {
    double rs2, rc2;
    uint32_t pice, epoch;
    uint64_t seed;
} orc_rbl;

static double orc_pair_weight(const orc_rbl* rb, double d2, int64_t i, int64_t j)
{
    if (!rb) return 1.0;
    if ((rb->rs2 - d2 > 1e-4) && (d2 > 1e-4)) return 1.0;
    if ((rb->rc2 - d2 > 1e-4) && (d2 - rb->rs2 > 1e-4) &&
        orc_rbl_select(rb->seed, rb->epoch, (int32_t)i, (int32_t)j, rb->pice))
        return (double)rb->pice;
    return 0.0;
}

static void orc_pair_forces_w(int64_t n, const double* x, const double* y, const double* z,
                              const double* q, const int32_t* type, const double* eps,
                              const double* sig, const double* lo, const double* L, double rc_lj,
                              double rc_coul, double alpha, double qqr2e, const orc_rbl* rb,
                              double* fx, double* fy, double* fz, double* ev)
{
    double rmax = rc_lj > rc_coul ? rc_lj : rc_coul;
    orc_cells c;
    orc_cells_build(&c, n, x, y, z, lo, L, rmax);
    int use_cells = c.dims[0] >= 3 && c.dims[1] >= 3 && c.dims[2] >= 3;
    const double rcl2 = rc_lj * rc_lj, rcc2 = rc_coul * rc_coul;
    const double sa = sqrt(alpha), tsp = 2.0 * sqrt(alpha / ORC_PI);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : acc[:8])
    for (int64_t i = 0; i < n; ++i) {
        double fxi = 0, fyi = 0, fzi = 0;
        int ti = type ? type[i] : 0;
        int ijk[3];
        orc_cell_of(&c, lo, x[i], y[i], z[i], ijk);
        int ncand = use_cells ? 27 : 1;
        for (int st = 0; st < ncand; ++st) {
            int64_t s0, s1;
            if (use_cells) {
                int a = st / 9 - 1, b = (st / 3) % 3 - 1, g = st % 3 - 1;
                int cx = (ijk[0] + a + c.dims[0]) % c.dims[0];
                int cy = (ijk[1] + b + c.dims[1]) % c.dims[1];
                int cz = (ijk[2] + g + c.dims[2]) % c.dims[2];
                int64_t cc = cx + cy * (int64_t)c.dims[0] + cz * (int64_t)c.dims[0] * c.dims[1];
                s0 = c.start[cc];
                s1 = c.start[cc + 1];
            } else {
                s0 = 0;
                s1 = n;
            }
            for (int64_t s = s0; s < s1; ++s) {
                int64_t j = use_cells ? c.ids[s] : s;
                double dx = orc_mi(x[j], x[i], L[0]);
                double dy = orc_mi(y[j], y[i], L[1]);
                double dz = orc_mi(z[j], z[i], L[2]);
                double d2 = orc_d2(dx, dy, dz);
                if (!(d2 > 1e-4)) continue;
                const double w = orc_pair_weight(rb, d2, i, j);
                if (w == 0.0) continue;
                double fs = 0.0;
                if (d2 < rcl2) {
                    int tj = type ? type[j] : 0;
                    double e_ij = sqrt(eps[ti] * eps[tj]);
                    double s_ij = 0.5 * (sig[ti] + sig[tj]);
                    double s2 = s_ij * s_ij / d2;
                    double s6 = s2 * s2 * s2;
                    fs += -24.0 * e_ij * (2.0 * s6 - 1.0) * s6 / d2;
                    acc[0] += w * (0.5 * 4.0 * e_ij * (s6 * s6 - s6));
                }
                if (d2 < rcc2) {
                    double d = sqrt(d2);
                    double qq = q[i] * q[j] * qqr2e;
                    double er = erfc(sa * d);
                    double g = er / d2 + tsp * exp(-alpha * d2) / d;
                    fs += -qq * g / d;
                    acc[1] += w * (0.5 * qq * er / d);
                }
                fs *= w;
                fxi += fs * dx;
                fyi += fs * dy;
                fzi += fs * dz;
                double h = -0.5 * fs;
                acc[2] += h * dx * dx;
                acc[3] += h * dy * dy;
                acc[4] += h * dz * dz;
                acc[5] += h * dx * dy;
                acc[6] += h * dx * dz;
                acc[7] += h * dy * dz;
            }
        }
        fx[i] = fxi;
        fy[i] = fyi;
        fz[i] = fzi;
    }
    if (ev)
        for (int k = 0; k < 8; ++k) ev[k] = acc[k];
    orc_cells_free(&c);
}

/* Per-atom pair forces, energy and virial over every j within the cutoffs.
 *   LJ  (ExecForceFunction::ComputeLJForce, forceFunction/ExecForceFunction.h:74-96):
 *       f_i = -24 sqrt(e_i e_j) (2s-1) s / d2 * r_ij, s = (sig_ij^2/d2)^3,
 *       sig_ij = (sig_i+sig_j)/2, r_ij = r_j - r_i, when 1e-4 < d2 < rc_lj^2.
 *   Coulomb (ComputeNearEnergyForce + Gnear, ExecForceFunction.h:125-140,217-224):
 *       f_i = -q_i q_j G(alpha,d) r_ij / d, when 0.01 < d < rc_coul, times qqr2e.
 *   Energies (ComputePotentialEn :296-313, ComputeNearEleEnergy :315-329):
 *       1/2 * 4 eps (s^2 - s) and 1/2 * q_i q_j erfc(sqrt(a) d)/d per directed pair.
 *   Virial, per directed pair, W_ab += -1/2 fs r_a r_b with f_i = fs * r_ij (the
 *       convention of the reference GPU kernel ComputeLJForceViral,
 *       rbmd-gpu src/force/src/ljforce_op/rocm/ljforce_op.hip.cu:555-575).
 * The loop runs over all j of the 27-cell stencil (or all j when the grid has
 * fewer than 3 cells per axis), i.e. the set the full Verlet list of
 * RunWorklet.cpp:170-237 visits.  No same-molecule exclusion: in-built ions
 * have molecule id = atom id (InitCondition/MeshFreeCondition.cpp:120-128).
 * ev (nullable): [E_lj, E_real, W_xx, W_yy, W_zz, W_xy, W_xz, W_yz]. */
void orc_pair_forces(int64_t n, const double* x, const double* y, const double* z,
                     const double* q, const int32_t* type, const double* eps, const double* sig,
                     const double* lo, const double* L, double rc_lj, double rc_coul,
                     double alpha, double qqr2e, double* fx, double* fy, double* fz, double* ev)
{
    orc_pair_forces_w(n, x, y, z, q, type, eps, sig, lo, L, rc_lj, rc_coul, alpha, qqr2e, NULL,
                      fx, fy, fz, ev);
}

/* Random batch list near forces (ExecutionMD::ComputeRBLNearForce,
 * run/ExecutionMD.cpp:460-566): the pair forces of orc_pair_forces over the core
 * pairs (weight 1) and the shell pairs orc_rbl_select keeps at build `epoch`
 * (weight pice), energies and virial weighted alike; then the mean force over all
 * atoms, summed in atom order, is subtracted from every atom (SumRBLCorrForceWorklet,
 * run/worklet/RunWorklet.cpp:1304-1321).  rc of the shell = max(rc_lj, rc_coul). */
void orc_rbl_forces(int64_t n, const double* x, const double* y, const double* z,
                    const double* q, const int32_t* type, const double* eps, const double* sig,
                    const double* lo, const double* L, double rc_lj, double rc_coul,
                    double alpha, double qqr2e, double r_core, int32_t pice, uint64_t seed,
                    uint32_t epoch, double* fx, double* fy, double* fz, double* ev)
{
    const double rc = rc_lj > rc_coul ? rc_lj : rc_coul;
    orc_rbl rb = {r_core * r_core, rc * rc, (uint32_t)pice, epoch, seed};
    orc_pair_forces_w(n, x, y, z, q, type, eps, sig, lo, L, rc_lj, rc_coul, alpha, qqr2e, &rb,
                      fx, fy, fz, ev);
    double m[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < n; ++i) {
        m[0] += fx[i];
        m[1] += fy[i];
        m[2] += fz[i];
    }
    for (int d = 0; d < 3; ++d) m[d] /= (double)n;
    for (int64_t i = 0; i < n; ++i) {
        fx[i] -= m[0];
        fy[i] -= m[1];
        fz[i] -= m[2];
    }
}

/* ------------------------------------------------------------------ */
/* k-space: RBE normaliser, structure factor, forces, Ewald             */
/* ------------------------------------------------------------------ */

/* H_d and S = H_x H_y H_z - 1 of RBEPSAMPLE::Compute_H/Compute_S
 * (model/RBEPSample.h:38-61): H_d = sqrt(alpha L_d^2/pi) sum_{m=-10..10}
 * exp(-alpha L_d^2 m^2) (the Poisson-dual form of sum_m exp(-pi^2 m^2/(alpha L^2))). */
double orc_rbe_S(double alpha, const double* L, double* H /* 3, nullable */)
{
    double prod = 1.0;
    for (int d = 0; d < 3; ++d) {
        double f = -(alpha * L[d] * L[d]);
        double h = 0.0;
        for (int m = -10; m <= 10; ++m) h += exp((double)(m * m) * f);
        h *= sqrt(-f / ORC_PI);
        if (H) H[d] = h;
        prod *= h;
    }
    return prod - 1.0;
}

/* Neumaier-compensated accumulator. */
typedef struct {
    double s, c;
} orc_ksum;
static inline void orc_kadd(orc_ksum* a, double v)
{
    double t = a->s + v;
    if (fabs(a->s) >= fabs(v))
        a->c += (a->s - t) + v;
    else
        a->c += (v - t) + a->s;
    a->s = t;
}

/* rho(k_l) = sum_j q_j exp(i k_l . r_j), k_l = 2 pi m_l / L.
 * ComputePnumberChargeStructureFactorWorklet + ReduceByKey
 * (run/worklet/RunWorklet.cpp:1616-1677, run/ExecutionMD.cpp:412-448), in FP64
 * with compensated sums over atoms. */
void orc_rho(int64_t n, const double* x, const double* y, const double* z, const double* q,
             const double* L, int32_t P, const int32_t* m, double* re, double* im)
{
#pragma omp parallel for schedule(static)
    for (int32_t l = 0; l < P; ++l) {
        double kx = 2.0 * ORC_PI * m[3 * l] / L[0];
        double ky = 2.0 * ORC_PI * m[3 * l + 1] / L[1];
        double kz = 2.0 * ORC_PI * m[3 * l + 2] / L[2];
        orc_ksum a = {0, 0}, b = {0, 0};
        for (int64_t j = 0; j < n; ++j) {
            double ph = kx * x[j] + ky * y[j] + kz * z[j];
            orc_kadd(&a, q[j] * cos(ph));
            orc_kadd(&b, q[j] * sin(ph));
        }
        re[l] = a.s + a.c;
        im[l] = b.s + b.c;
    }
}

/* k-space force and energy/virial from P k-vectors with per-k weights w_l.
 *   F_i = sum_l w_l (-4 pi q_i k_l / (V k_l^2)) (cos(k.r_i) Im rho_l - sin(k.r_i) Re rho_l) qqr2e
 * is ComputeRBEForceSum/ComputeRBEForce (forceFunction/ExecForceFunction.h:247-269,
 * run/worklet/RunWorklet.cpp:1508-1548) with w_l = S/P, and ComputeFarEle
 * (ExecForceFunction.h:226-245) with w_l = exp(-k^2/4 alpha) over the full lattice.
 *   E_k  = sum_l w_l (2 pi / V) |rho_l|^2 / k_l^2 qqr2e     (ThermoOutput.cpp:209-255)
 *   W_ab = sum_l w_l (2 pi / V) |rho_l|^2 / k_l^2 (d_ab - 2 (1/k^2 + 1/(4 alpha)) k_a k_b) qqr2e
 * ev (nullable): [E_k, W_xx, W_yy, W_zz, W_xy, W_xz, W_yz]. fx/fy/fz are ADDED to. */
void orc_kspace_forces(int64_t n, const double* x, const double* y, const double* z,
                       const double* q, const double* L, double alpha, double qqr2e, int32_t P,
                       const int32_t* m, const double* w, const double* re, const double* im,
                       double* fx, double* fy, double* fz, double* ev)
{
    const double V = L[0] * L[1] * L[2];
    double* kx = (double*)malloc(sizeof(double) * (P > 0 ? P : 1) * 4);
    double* ky = kx + P;
    double* kz = ky + P;
    double* cf = kz + P;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int32_t l = 0; l < P; ++l) {
        kx[l] = 2.0 * ORC_PI * m[3 * l] / L[0];
        ky[l] = 2.0 * ORC_PI * m[3 * l + 1] / L[1];
        kz[l] = 2.0 * ORC_PI * m[3 * l + 2] / L[2];
        double k2 = kx[l] * kx[l] + ky[l] * ky[l] + kz[l] * kz[l];
        cf[l] = w[l] * (-4.0 * ORC_PI) / (V * k2) * qqr2e;
        double e = w[l] * (2.0 * ORC_PI / V) * (re[l] * re[l] + im[l] * im[l]) / k2 * qqr2e;
        double b = 2.0 * (1.0 / k2 + 1.0 / (4.0 * alpha));
        acc[0] += e;
        acc[1] += e * (1.0 - b * kx[l] * kx[l]);
        acc[2] += e * (1.0 - b * ky[l] * ky[l]);
        acc[3] += e * (1.0 - b * kz[l] * kz[l]);
        acc[4] += e * (-b * kx[l] * ky[l]);
        acc[5] += e * (-b * kx[l] * kz[l]);
        acc[6] += e * (-b * ky[l] * kz[l]);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double ax = 0, ay = 0, az = 0;
        for (int32_t l = 0; l < P; ++l) {
            double ph = kx[l] * x[i] + ky[l] * y[i] + kz[l] * z[i];
            double s = q[i] * cf[l] * (cos(ph) * im[l] - sin(ph) * re[l]);
            ax += s * kx[l];
            ay += s * ky[l];
            az += s * kz[l];
        }
        fx[i] += ax;
        fy[i] += ay;
        fz[i] += az;
    }
    if (ev)
        for (int k = 0; k < 7; ++k) ev[k] = acc[k];
    free(kx);
}

/* Full Ewald k-space lattice: every m != 0 with |m_d| <= kmax_d, weight
 * exp(-k^2/(4 alpha)) (ExecutionMD::ComputeChargeStructureFactorEwald,
 * run/ExecutionMD.cpp:355-387, with the reference's Kmax cube).  Writes the
 * m-list and weights; returns the count.  Call with m == NULL to size. */
int64_t orc_ewald_kvectors(const double* L, double alpha, const int32_t* kmax, int32_t* m,
                           double* w)
{
    int64_t c = 0;
    for (int32_t a = -kmax[0]; a <= kmax[0]; ++a)
        for (int32_t b = -kmax[1]; b <= kmax[1]; ++b)
            for (int32_t g = -kmax[2]; g <= kmax[2]; ++g) {
                if (a == 0 && b == 0 && g == 0) continue;
                if (m) {
                    double kx = 2.0 * ORC_PI * a / L[0], ky = 2.0 * ORC_PI * b / L[1],
                           kz = 2.0 * ORC_PI * g / L[2];
                    double k2 = kx * kx + ky * ky + kz * kz;
                    m[3 * c] = a;
                    m[3 * c + 1] = b;
                    m[3 * c + 2] = g;
                    w[c] = exp(-k2 / (4.0 * alpha));
                }
                ++c;
            }
    return c;
}

/* Self energy -sqrt(alpha/pi) sum q^2 (output/ThermoOutput.cpp:196-203). */
double orc_self_energy(int64_t n, const double* q, double alpha, double qqr2e)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += q[i] * q[i];
    return -sqrt(alpha / ORC_PI) * s * qqr2e;
}

/* ------------------------------------------------------------------ */
/* velocity Verlet                                                      */
/* ------------------------------------------------------------------ */

/* Half kick v += 1/2 dt F/m * fmt2v (UpdateVelocityWorklet,
 * run/worklet/RunWorklet.cpp:1764-1782). */
void orc_kick(int64_t n, double* vx, double* vy, double* vz, const double* fx, const double* fy,
              const double* fz, const double* mass, double dt, double fmt2v)
{
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.5 * dt * fmt2v / (mass ? mass[i] : 1.0);
        vx[i] += s * fx[i];
        vy[i] += s * fy[i];
        vz[i] += s * fz[i];
    }
}

/* Drift x += v dt, then wrap with image flags (UpdatePositionFlagWorklet,
 * run/worklet/RunWorklet.cpp:1854-1877). */
void orc_drift(int64_t n, double* x, double* y, double* z, const double* vx, const double* vy,
               const double* vz, double dt, const double* lo, const double* L, int32_t* img)
{
    for (int64_t i = 0; i < n; ++i) {
        x[i] += vx[i] * dt;
        y[i] += vy[i] * dt;
        z[i] += vz[i] * dt;
    }
    orc_wrap(n, x, y, z, lo, L, img);
}

/* ------------------------------------------------------------------ */
/* NVT velocity thermostats                                              */
/* ------------------------------------------------------------------ */

/* Kinetic temperature T = mvv2e * sum_i m_i |v_i|^2 / (dof * kB)
 * (ComputerKineticEnergyWorklet, run/worklet/RunWorklet.cpp:1806-1816, reduced and
 * divided in ExecutionNVT::ComputeTempe, run/ExecutionNVT.cpp:749-787: 0.5 * sum /
 * (dof * kB / 2) with dof = 3n - 3 for read-in systems, 3n for the in-built lattice). */
double orc_temperature(int64_t n, const double* vx, const double* vy, const double* vz,
                       const double* mass, double mvv2e, double dof, double kB)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double m = mass ? mass[i] : 1.0;
        s += mvv2e * (m * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]));
    }
    return s / (dof * kB);
}

/* v *= c (UpdateVelocityRescaleWorklet, run/worklet/RunWorklet.cpp:1818-1830). */
void orc_scale_velocities(int64_t n, double* vx, double* vy, double* vz, double c)
{
    for (int64_t i = 0; i < n; ++i) {
        vx[i] = c * vx[i];
        vy[i] = c * vy[i];
        vz[i] = c * vz[i];
    }
}

/* Rescale coefficients of ExecutionNVT::UpdateVelocityByTempConType
 * (run/ExecutionNVT.cpp:272-289): TEMP_RESCALE sqrt(T0 / T), BERENDSEN
 * sqrt(1 + dt / Tdamp * (T0 / T - 1)). */
double orc_rescale_coeff(double T, double T0) { return sqrt(T0 / T); }

double orc_berendsen_coeff(double T, double T0, double dt, double tdamp)
{
    return sqrt(1.0 + dt / tdamp * (T0 / T - 1.0));
}

/* Nose-Hoover half step v += 1/2 dt (F/m - xi v) fmt2v
 * (UpdateVelocityNoseHooverWorklet, run/worklet/RunWorklet.cpp:1784-1804), applied
 * after the closing kick with the friction xi of the step before. */
void orc_nose_hoover_kick(int64_t n, double* vx, double* vy, double* vz, const double* fx,
                          const double* fy, const double* fz, const double* mass, double dt,
                          double fmt2v, double xi)
{
    for (int64_t i = 0; i < n; ++i) {
        double m = mass ? mass[i] : 1.0;
        vx[i] = vx[i] + 0.5 * dt * (fx[i] / m - xi * vx[i]) * fmt2v;
        vy[i] = vy[i] + 0.5 * dt * (fy[i] / m - xi * vy[i]) * fmt2v;
        vz[i] = vz[i] + 0.5 * dt * (fz[i] / m - xi * vz[i]) * fmt2v;
    }
}

/* Friction update xi += 1/2 dt (T / T0 - 1) / tauT with tauT = dt / 10
 * (run/ExecutionNVT.cpp:269-270). */
double orc_nose_hoover_xi(double xi, double T, double T0, double dt)
{
    double tauT = 0.1 * dt;
    return xi + 0.5 * dt * (T / T0 - 1.0) / tauT;
}

/* Underdamped Langevin force F += -m gamma v + sqrt(2 m gamma kBT / dt) g with one
 * Gaussian 3-vector g per step shared by every atom (UnderdampedLangevinWorklet,
 * run/worklet/RunWorklet.cpp:1700-1728; g drawn in ExecutionNVT::ComputeAllForce,
 * run/ExecutionNVT.cpp:347-352; applied after the force by TempConTypeForce,
 * :735-747).  v is the velocity at the force evaluation (after the opening kick).
 * m is the atom's own mass (as the library's Langevin term); the reference passes a
 * unit-mass array (ExecutionNVT.cpp:737-738), the same thing for unit masses, its only
 * case.  With other masses this restates the library's variant, not the reference. */
void orc_langevin_force(int64_t n, double* fx, double* fy, double* fz, const double* vx,
                        const double* vy, const double* vz, const double* mass, double gamma,
                        double kbt, double dt, const double* g)
{
    for (int64_t i = 0; i < n; ++i) {
        double m = mass ? mass[i] : 1.0;
        double amp = sqrt(2.0 * m * gamma * kbt / dt);
        fx[i] = fx[i] - m * gamma * vx[i] + amp * g[0];
        fy[i] = fy[i] - m * gamma * vy[i] + amp * g[1];
        fz[i] = fz[i] - m * gamma * vz[i] + amp * g[2];
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
