/*
 * oracle/faithful.c -- reference-faithful FP32 MD step, the timed CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg and by
 * tests/, never by the product library.  It mirrors the reference RBMD CPU
 * path's per-step algorithm (FP32 Real, AoS, list rebuilt at rc every step,
 * N*P structure-factor arrays), so timing it on the host cores stands in for
 * timing the reference binary, which cannot be built here (it needs VTK-m).
 * Built separately from refcpu.c so it may use FMA contraction like an
 * optimised reference build would.
 *
 * The file is compiled twice: with REAL = float as orc_faithful_run (the
 * reference's default FP32 `Real`, rbmd/framework/Types.h:48) and with
 * REAL = double as orc_faithful_run_f64 (a double-precision build of the same
 * path, the analogue for the FP64 LJ workload).  <tgmath.h> picks the float or
 * double maths functions from the argument types.
 */
#include <stdint.h>
#include <tgmath.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

#ifndef REAL
#define REAL float
#endif
#ifndef FAITHFUL_NAME
#define FAITHFUL_NAME orc_faithful_run
#endif
#define R(v) ((REAL)(v))

double orc_rbe_S(double alpha, const double* L, double* H);

/* ------------------------------------------------------------------ */
/* reference-faithful step (timed CPU baseline)                          */
/* ------------------------------------------------------------------ */

/* One or more MD steps in the reference's own shape, `Real` = REAL
 * (Types.h:48), AoS Vec3f arrays, for timing the CPU path:
 *   kick (RunWorklet.cpp:1764-1782); drift + single-shift wrap
 *   (:1854-1877); locator rebuild = bin + SortByKey + bounds
 *   (ContPointLocator.cxx:70-92; the sort by cell key, ties by atom id, is a
 *   counting sort here: the same order as a stable SortByKey, in O(n));
 *   full list at rc with NO skin into
 *   N x (rho ceil(4/3 pi rc^3)+1) slots with Vec3f offsets
 *   (ExecutionMD.cpp:670-705, RunWorklet.cpp:40-96); near force over the list
 *   (RunWorklet.cpp:170-237); RBE rho through materialised N*P Re/Im arrays
 *   and a reduce by key (RunWorklet.cpp:1616-1677, ExecutionMD.cpp:412-448);
 *   RBE force loop (RunWorklet.cpp:1508-1548) with S/P; F = far + near
 *   (RunWorklet.cpp:1730-1744); kick.  ExecutionNVT::Solve order
 *   (run/ExecutionNVT.cpp:133-159).  The k-list of every step is an input
 *   (m: nsteps x P x 3).  q NULL: no charges (the LJ-only force,
 *   ComputeLJForceVerletWorklet, RunWorklet.cpp:239-303).  nthreads > 1
 *   parallelises the per-atom loops with OpenMP (the VTK-m OpenMP device).
 *   Returns 0, or -1 on list overflow. */
int FAITHFUL_NAME(int64_t n, REAL* pos, REAL* vel, REAL* force, const REAL* q,
                  const double* lo_d, const double* L_d, double rc_lj_d, double rc_coul_d,
                  double alpha_d, int32_t P, const int32_t* m, int32_t nsteps, double dt_d,
                  int32_t nthreads)
{
#ifdef _OPENMP
    /* the thread count of this call only: the caller's setting is restored below */
    const int threads_before = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const REAL lo[3] = {R(lo_d[0]), R(lo_d[1]), R(lo_d[2])};
    const REAL L[3] = {R(L_d[0]), R(L_d[1]), R(L_d[2])};
    const REAL hi[3] = {lo[0] + L[0], lo[1] + L[1], lo[2] + L[2]};
    const REAL rc = R(rc_coul_d > rc_lj_d ? rc_coul_d : rc_lj_d);
    const REAL rcl = R(rc_lj_d), rcc = R(rc_coul_d);
    const REAL alpha = R(alpha_d), dt = R(dt_d);
    const REAL V = L[0] * L[1] * L[2];
    const REAL rho_sys = R((double)n / ((double)L[0] * L[1] * L[2]));
    const REAL eps4 = R(0.0001);
    /* The reference sizes the fixed-stride rows as rho*ceil(4/3 pi rc^3) + 1
     * (ExecutionMD.cpp:676-691) without an overflow check; at these densities
     * that is below the longest rows, so the baseline doubles it to do the
     * full work (overflow is still reported). */
    const int64_t max_j =
        2 * ((int64_t)(rho_sys * ceil(R(4.0 / 3.0) * R(ORC_PI) * rc * rc * rc)) + 1);
    int dims[3];
    REAL dxyz[3];
    for (int d = 0; d < 3; ++d) {
        dims[d] = (int)(L[d] / rc);
        if (dims[d] < 1) dims[d] = 1;
        dxyz[d] = L[d] / R(dims[d]);
    }
    const int64_t ncell = (int64_t)dims[0] * dims[1] * dims[2];
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* lower = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    int64_t* upper = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    int32_t* pid = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* vid = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * max_j));
    REAL* voff = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(n * max_j));
    int64_t* vnum = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    REAL* near = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)n);
    REAL* far = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)n);
    REAL* rre = (REAL*)malloc(sizeof(REAL) * (size_t)(n * P > 0 ? n * P : 1));
    REAL* rim = (REAL*)malloc(sizeof(REAL) * (size_t)(n * P > 0 ? n * P : 1));
    REAL* rho = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)(P > 0 ? P : 1));
    REAL* kl = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(P > 0 ? P : 1));
    int overflow = 0;
    const REAL sa = sqrt(alpha), spi = sqrt(R(ORC_PI));
    double H[3];
    const REAL S = P > 0 ? R(orc_rbe_S(alpha_d, L_d, H)) : R(0);

    for (int32_t step = 0; step < nsteps; ++step) {
        /* stage 1: kick; stage 2: drift + wrap */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            for (int d = 0; d < 3; ++d) {
                vel[3 * i + d] += R(0.5) * force[3 * i + d] / R(1) * dt * R(1);
                pos[3 * i + d] += vel[3 * i + d] * dt;
                if (pos[3 * i + d] < lo[d])
                    pos[3 * i + d] += hi[d] - lo[d];
                else if (pos[3 * i + d] > hi[d])
                    pos[3 * i + d] -= hi[d] - lo[d];
            }
        }
        /* locator: bin, sort by key (ties by id), lower/upper bounds */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            int64_t ijk[3];
            for (int d = 0; d < 3; ++d) {
                int64_t v = (int64_t)((pos[3 * i + d] - lo[d]) / dxyz[d]);
                if (v < 0) v = 0;
                if (v > dims[d] - 1) v = dims[d] - 1;
                ijk[d] = v;
            }
            key[i] = ijk[0] + ijk[1] * dims[0] + ijk[2] * (int64_t)dims[0] * dims[1];
        }
        for (int64_t c = 0; c < ncell; ++c) upper[c] = 0;
        for (int64_t i = 0; i < n; ++i) ++upper[key[i]];
        {
            int64_t s = 0;
            for (int64_t c = 0; c < ncell; ++c) {
                lower[c] = s;
                s += upper[c];
                upper[c] = lower[c];
            }
        }
        for (int64_t i = 0; i < n; ++i) pid[upper[key[i]]++] = (int32_t)i;
        /* full Verlet list at rc, fixed-stride slots with image offsets */
#pragma omp parallel for schedule(dynamic, 64) reduction(| : overflow)
        for (int64_t i = 0; i < n; ++i) {
            const REAL* pi = pos + 3 * i;
            int64_t ijk[3];
            for (int d = 0; d < 3; ++d) ijk[d] = (int64_t)((pi[d] - lo[d]) / dxyz[d]);
            int64_t cnt = 0;
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b)
                    for (int g = -1; g <= 1; ++g) {
                        int64_t nb[3] = {ijk[0] + a, ijk[1] + b, ijk[2] + g};
                        int64_t w[3];
                        REAL off[3];
                        for (int d = 0; d < 3; ++d) {
                            w[d] = (nb[d] + dims[d]) % dims[d];
                            off[d] = dxyz[d] * R(w[d] - nb[d]);
                        }
                        int64_t cc = w[0] + w[1] * dims[0] + w[2] * (int64_t)dims[0] * dims[1];
                        for (int64_t p = lower[cc]; p < upper[cc]; ++p) {
                            int32_t j = pid[p];
                            REAL rx = (pos[3 * j] - off[0]) - pi[0];
                            REAL ry = (pos[3 * j + 1] - off[1]) - pi[1];
                            REAL rz = (pos[3 * j + 2] - off[2]) - pi[2];
                            REAL d2 = rx * rx + ry * ry + rz * rz;
                            if (rc * rc - d2 > eps4 && d2 > eps4) {
                                if (cnt < max_j) {
                                    vid[i * max_j + cnt] = j;
                                    voff[3 * (i * max_j + cnt)] = off[0];
                                    voff[3 * (i * max_j + cnt) + 1] = off[1];
                                    voff[3 * (i * max_j + cnt) + 2] = off[2];
                                } else {
                                    overflow |= 1;
                                }
                                ++cnt;
                            }
                        }
                    }
            vnum[i] = cnt < max_j ? cnt : max_j;
        }
        /* near force over the list */
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i) {
            REAL f[3] = {0, 0, 0}, fe[3] = {0, 0, 0};
            const REAL* pi = pos + 3 * i;
            for (int64_t p = 0; p < vnum[i]; ++p) {
                int32_t j = vid[i * max_j + p];
                REAL r[3];
                for (int d = 0; d < 3; ++d) r[d] = (pos[3 * j + d] - voff[3 * (i * max_j + p) + d]) - pi[d];
                REAL d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                REAL dis = sqrt(d2);
                if (q && dis < rcc && dis > R(0.01)) {
                    REAL g = (R(1) - erf(sa * dis)) / (dis * dis) +
                             R(2) * sa * exp(-alpha * dis * dis) / (spi * dis);
                    for (int d = 0; d < 3; ++d) fe[d] += -q[i] * q[j] * g * r[d] / dis;
                }
                if (d2 < rcl * rcl && d2 > eps4) {
                    REAL s6 = R(1) / (d2 * d2 * d2);
                    REAL c = R(-24) * ((R(2) * s6 - R(1)) * s6) / d2;
                    for (int d = 0; d < 3; ++d) f[d] += c * r[d];
                }
            }
            for (int d = 0; d < 3; ++d) near[3 * i + d] = f[d] + fe[d];
        }
        /* RBE: N*P materialised structure-factor terms, reduce by key */
        if (P > 0 && q) {
            const int32_t* ms = m + (int64_t)step * P * 3;
            for (int32_t l = 0; l < P; ++l)
                for (int d = 0; d < 3; ++d) kl[3 * l + d] = R(2) * R(ORC_PI) * R(ms[3 * l + d]) / L[d];
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                for (int32_t l = 0; l < P; ++l) {
                    REAL ph = kl[3 * l] * pos[3 * i] + kl[3 * l + 1] * pos[3 * i + 1] + kl[3 * l + 2] * pos[3 * i + 2];
                    rre[i + l * n] = q[i] * cos(ph);
                    rim[i + l * n] = q[i] * sin(ph);
                }
            }
#pragma omp parallel for schedule(static)
            for (int32_t l = 0; l < P; ++l) {
                REAL a = 0, b = 0;
                for (int64_t i = 0; i < n; ++i) {
                    a += rre[i + l * n];
                    b += rim[i + l * n];
                }
                rho[2 * l] = a;
                rho[2 * l + 1] = b;
            }
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                REAL f[3] = {0, 0, 0};
                for (int32_t l = 0; l < P; ++l) {
                    const REAL* k = kl + 3 * l;
                    REAL k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                    REAL ph = k[0] * pos[3 * i] + k[1] * pos[3 * i + 1] + k[2] * pos[3 * i + 2];
                    REAL bc = cos(ph) * rho[2 * l + 1];
                    REAL cs = sin(ph) * rho[2 * l];
                    for (int d = 0; d < 3; ++d) f[d] += (R(-4) * R(ORC_PI) * q[i] * k[d] / (V * k2)) * (bc - cs);
                }
                for (int d = 0; d < 3; ++d) far[3 * i + d] = f[d] * S / R(P);
            }
        } else {
            memset(far, 0, sizeof(REAL) * 3 * (size_t)n);
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                force[3 * i + d] = far[3 * i + d] + near[3 * i + d];
                vel[3 * i + d] += R(0.5) * force[3 * i + d] * dt;
            }
    }
    free(key);
    free(lower);
    free(upper);
    free(pid);
    free(vid);
    free(voff);
    free(vnum);
    free(near);
    free(far);
    free(rre);
    free(rim);
    free(rho);
    free(kl);
#ifdef _OPENMP
    omp_set_num_threads(threads_before);
#endif
    return overflow ? -1 : 0;
}
