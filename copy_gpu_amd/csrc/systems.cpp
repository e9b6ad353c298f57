// systems.cpp -- host-side deterministic inputs: Philox4x32-10, FCC lattice, random
// ions with a minimum separation, +/-1 charges, Maxwell-Boltzmann velocities, and
// the discrete-Gaussian k-sampler tables shared with the device sampler.
//
// The reference has no seeded generators for these configurations (its in-built
// lattice is a simple-cubic cell-centroid set, InitCondition/MeshFreeCondition.cpp:58-110,
// and its k-sampler draws from std::random_device, math/Utiles.h:32-39), so these
// are this project's own; the k-sampler targets the same distribution as
// RBEPSAMPLE (model/RBEPSample.h:87-93): p(m) ~ exp(-(2 pi m / L)^2 / (4 alpha)), m != 0.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/rbmd_force.h"
#include "engine.h"

namespace rbmd {

// Per-axis discrete Gaussian CDF over m in [-M, M], M = ceil(9.5 sigma) + 1 with
// sigma^2 = alpha L^2 / (2 pi^2) (pmf below 1e-19 of the peak beyond M).
void build_kcdf(double alpha, const double* L, std::vector<double>& cdf, int M[3])
{
    cdf.clear();
    for (int d = 0; d < 3; ++d) {
        const double sigma = sqrt(alpha / 2.0) * L[d] / M_PI;
        M[d] = (int)ceil(9.5 * sigma) + 1;
        const double c = M_PI * M_PI / (alpha * L[d] * L[d]);
        std::vector<double> p((size_t)(2 * M[d] + 1));
        double tot = 0.0;
        for (int m = -M[d]; m <= M[d]; ++m) {
            p[(size_t)(m + M[d])] = exp(-c * (double)m * (double)m);
        }
        // normaliser summed small-to-large for accuracy
        for (int k = M[d]; k >= 1; --k) tot += p[(size_t)(M[d] + k)] + p[(size_t)(M[d] - k)];
        tot += p[(size_t)M[d]];
        double run = 0.0;
        for (size_t i = 0; i < p.size(); ++i) {
            run += p[i];
            cdf.push_back(run / tot);
        }
        cdf.back() = 1.0;
    }
}

// S = H_x H_y H_z - 1, H_d = sqrt(alpha L_d^2 / pi) sum_{m=-10..10} exp(-alpha L_d^2 m^2)
// (RBEPSAMPLE::Compute_H / Compute_S, model/RBEPSample.h:38-61).
double rbe_S(double alpha, const double* L)
{
    double prod = 1.0;
    for (int d = 0; d < 3; ++d) {
        const double f = -(alpha * L[d] * L[d]);
        double h = 0.0;
        for (int m = -10; m <= 10; ++m) h += exp((double)(m * m) * f);
        h *= sqrt(-f / M_PI);
        prod *= h;
    }
    return prod - 1.0;
}

namespace {

struct Rng {
    uint32_t key[2];
    uint32_t tag;
    uint64_t ctr = 0;
    Rng(uint64_t seed, uint32_t t) : tag(t)
    {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
    }
    // two 53-bit uniforms per Philox block
    void uniform2(double* a, double* b)
    {
        uint32_t c[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), tag, 0x5EEDu}, r[4];
        philox4x32_10(c, key, r);
        ++ctr;
        *a = u53(r[0], r[1]);
        *b = u53(r[2], r[3]);
    }
    double uniform()
    {
        double a, b;
        uniform2(&a, &b);
        return a;
    }
};

}  // namespace

}  // namespace rbmd

using namespace rbmd;

extern "C" {

void rbmd_philox4x32(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    philox4x32_10(ctr, key, out);
}

int rbmd_sys_fcc(int32_t nc, double rho, double* x, double* y, double* z, double* L_out)
{
    if (nc < 1 || !(rho > 0) || !x || !y || !z) return RBMD_ERR_INVALID;
    const double a = cbrt(4.0 / rho);
    const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
    int64_t k = 0;
    for (int iz = 0; iz < nc; ++iz)
        for (int iy = 0; iy < nc; ++iy)
            for (int ix = 0; ix < nc; ++ix)
                for (int b = 0; b < 4; ++b) {
                    x[k] = (ix + basis[b][0]) * a;
                    y[k] = (iy + basis[b][1]) * a;
                    z[k] = (iz + basis[b][2]) * a;
                    ++k;
                }
    if (L_out) *L_out = nc * a;
    return RBMD_OK;
}

int rbmd_sys_random(int64_t n, double L, double dmin, uint64_t seed, double* x, double* y,
                    double* z)
{
    if (n < 0 || !(L > 0) || !(dmin >= 0) || (n > 0 && (!x || !y || !z))) return RBMD_ERR_INVALID;
    Rng rng(seed, 0x504F5331u);
    int dims = dmin > 0 ? (int)(L / dmin) : 1;
    if (dims < 1) dims = 1;
    if (dims > 1024) dims = 1024;
    const double cell = L / dims;
    const int64_t ncell = (int64_t)dims * dims * dims;
    std::vector<int32_t> head((size_t)ncell, -1), next((size_t)std::max<int64_t>(n, 1), -1);
    const double d2min = dmin * dmin;
    const int64_t max_attempts = 1000 * std::max<int64_t>(n, 1) + 1000;
    int64_t placed = 0;
    for (int64_t att = 0; placed < n && att < max_attempts; ++att) {
        double u0, u1, u2, u3;
        rng.uniform2(&u0, &u1);
        rng.uniform2(&u2, &u3);
        const double p[3] = {u0 * L, u1 * L, u2 * L};
        int c[3];
        for (int d = 0; d < 3; ++d) {
            c[d] = (int)(p[d] / cell);
            if (c[d] >= dims) c[d] = dims - 1;
        }
        bool ok = true;
        const int span = dims >= 3 ? 1 : 0;
        for (int a = -span; a <= span && ok; ++a)
            for (int b = -span; b <= span && ok; ++b)
                for (int g = -span; g <= span && ok; ++g) {
                    const int cx = (c[0] + a + dims) % dims, cy = (c[1] + b + dims) % dims,
                              cz = (c[2] + g + dims) % dims;
                    const int64_t cc = cx + (int64_t)cy * dims + (int64_t)cz * dims * dims;
                    for (int32_t j = head[(size_t)cc]; j >= 0 && ok; j = next[(size_t)j]) {
                        double dx = x[j] - p[0], dy = y[j] - p[1], dz = z[j] - p[2];
                        dx -= L * rint(dx / L);
                        dy -= L * rint(dy / L);
                        dz -= L * rint(dz / L);
                        if (dx * dx + dy * dy + dz * dz < d2min) ok = false;
                    }
                }
        if (!ok) continue;
        x[placed] = p[0];
        y[placed] = p[1];
        z[placed] = p[2];
        const int64_t cc = c[0] + (int64_t)c[1] * dims + (int64_t)c[2] * dims * dims;
        next[(size_t)placed] = head[(size_t)cc];
        head[(size_t)cc] = (int32_t)placed;
        ++placed;
    }
    return placed == n ? RBMD_OK : RBMD_ERR_INVALID;
}

int rbmd_sys_charges(int64_t n, uint64_t seed, double* q)
{
    if (n < 0 || (n > 0 && !q)) return RBMD_ERR_INVALID;
    for (int64_t i = 0; i < n; ++i) q[i] = i < n / 2 ? 1.0 : -1.0;
    Rng rng(seed, 0x43484731u);
    for (int64_t i = n - 1; i > 0; --i) {
        int64_t j = (int64_t)(rng.uniform() * (double)(i + 1));
        if (j > i) j = i;
        std::swap(q[i], q[j]);
    }
    return RBMD_OK;
}

int rbmd_sys_velocities(int64_t n, double T, const double* mass, uint64_t seed, double* vx,
                        double* vy, double* vz)
{
    if (n < 2 || !(T >= 0) || !vx || !vy || !vz) return RBMD_ERR_INVALID;
    Rng rng(seed, 0x56454C31u);
    double* v[3] = {vx, vy, vz};
    for (int64_t i = 0; i < n; ++i) {
        double g[4];
        for (int k = 0; k < 2; ++k) {
            double u1, u2;
            rng.uniform2(&u1, &u2);
            const double r = sqrt(-2.0 * log(1.0 - u1));
            g[2 * k] = r * cos(2.0 * M_PI * u2);
            g[2 * k + 1] = r * sin(2.0 * M_PI * u2);
        }
        const double sd = sqrt(1.0 / (mass ? mass[i] : 1.0));
        for (int d = 0; d < 3; ++d) v[d][i] = g[d] * sd;
    }
    double mt = 0.0, p[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        const double m = mass ? mass[i] : 1.0;
        mt += m;
        for (int d = 0; d < 3; ++d) p[d] += m * v[d][i];
    }
    double ke2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double m = mass ? mass[i] : 1.0;
        for (int d = 0; d < 3; ++d) {
            v[d][i] -= p[d] / mt;
            ke2 += m * v[d][i] * v[d][i];
        }
    }
    const double Tnow = ke2 / (3.0 * (double)n - 3.0);
    const double sc = Tnow > 0 ? sqrt(T / Tnow) : 0.0;
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) v[d][i] *= sc;
    return RBMD_OK;
}

int rbmd_sample_kvectors_host(double alpha, const double* L, int32_t P, uint64_t seed,
                              uint64_t step, int32_t* m, double* S_out)
{
    if (!(alpha > 0) || !L || P < 0 || (P > 0 && !m)) return RBMD_ERR_INVALID;
    std::vector<double> cdf;
    int M[3];
    build_kcdf(alpha, L, cdf, M);
    for (int32_t l = 0; l < P; ++l) sample_kvector(seed, step, (uint32_t)l, cdf.data(), M, m + 3 * l);
    if (S_out) *S_out = rbe_S(alpha, L);
    return RBMD_OK;
}

}  // extern "C"
