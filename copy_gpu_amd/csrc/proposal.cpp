// proposal.cpp -- host tables of the learned k-space proposal (mixture of a
// tabulated density on the lattice box |m_d| <= Mg with the analytic sampler)
// and its host-side sampler, which draws exactly what the device sampler draws.
//
// The analytic component targets RBEPSAMPLE's distribution
// (rbmd/framework/model/RBEPSample.h:87-93); the reference has no learned
// proposal.
#include <math.h>

#include <vector>

#include "../../include/rbmd_force.h"
#include "engine.h"
#include "proposal.h"

namespace rbmd {
void build_kcdf(double alpha, const double* L, std::vector<double>& cdf, int M[3]);
double rbe_S(double alpha, const double* L);
}  // namespace rbmd

using namespace rbmd;

extern "C" {

int rbmd_proposal_create(int32_t Mg, const double* density, double lambda, double alpha,
                         const double* box_len, rbmd_proposal_t* out)
{
    if (!out) return RBMD_ERR_INVALID;
    *out = nullptr;
    if (Mg < 1 || Mg > 256 || !density || !box_len || !(lambda > 0.0 && lambda < 1.0) ||
        !(alpha > 0.0))
        return RBMD_ERR_INVALID;
    for (int d = 0; d < 3; ++d)
        if (!(box_len[d] > 0.0)) return RBMD_ERR_INVALID;
    const int side = 2 * Mg + 1;
    const size_t G = (size_t)side * side * side;
    const size_t zero = (size_t)Mg + side * ((size_t)Mg + side * (size_t)Mg);
    double tot = 0.0;
    for (size_t i = 0; i < G; ++i) {
        const double v = density[i];
        if (!(v >= 0.0) || !isfinite(v)) return RBMD_ERR_INVALID;
        if (i != zero) tot += v;
    }
    if (!(tot > 0.0) || !isfinite(tot)) return RBMD_ERR_INVALID;

    auto* p = new rbmd_proposal_s();
    p->Mg = Mg;
    p->lambda = lambda;
    p->alpha = alpha;
    for (int d = 0; d < 3; ++d) p->L[d] = box_len[d];
    build_kcdf(alpha, box_len, p->acdf, p->M);
    p->S = rbe_S(alpha, box_len);
    p->cdf.resize(G);
    p->ratio.resize(G);
    double run = 0.0;
    for (size_t i = 0; i < G; ++i) {
        const double pg = i == zero ? 0.0 : density[i] / tot;
        run += pg;
        p->cdf[i] = run;
        if (i == zero) {
            p->ratio[i] = 0.0;
            continue;
        }
        const int mx = (int)(i % side) - Mg, my = (int)((i / side) % side) - Mg,
                  mz = (int)(i / ((size_t)side * side)) - Mg;
        const double kx = 2.0 * M_PI * mx / box_len[0], ky = 2.0 * M_PI * my / box_len[1],
                     kz = 2.0 * M_PI * mz / box_len[2];
        const double a = exp(-(kx * kx + ky * ky + kz * kz) / (4.0 * alpha));
        p->ratio[i] = a / (lambda * pg + (1.0 - lambda) * a / p->S);
    }
    // the table closes at exactly 1 from the last entry of positive mass on
    // (a rounding-level shift of that entry's mass), so every u < 1 lands on it
    size_t last = G - 1;
    while (last > 0 && (last == zero || density[last] == 0.0)) --last;
    for (size_t j = last; j < G; ++j) p->cdf[j] = 1.0;
    // guide table: the smallest power of two >= G entries (exact u * K)
    p->guide_bits = 1;
    while (((size_t)1 << p->guide_bits) < G) ++p->guide_bits;
    const size_t K = (size_t)1 << p->guide_bits;
    p->guide.resize(K);
    size_t i = 0;
    for (size_t j = 0; j < K; ++j) {
        const double t = (double)j / (double)K;
        while (!(t < p->cdf[i])) ++i;  // cdf[last] = 1 > t ends the walk
        p->guide[j] = (int32_t)i;
    }
    *out = p;
    return RBMD_OK;
}

void rbmd_proposal_destroy(rbmd_proposal_t p) { delete p; }

int rbmd_proposal_sample_host(rbmd_proposal_t p, int32_t P, uint64_t seed, uint64_t step,
                              int32_t* m, double* w)
{
    if (!p || P < 1 || !m || !w) return RBMD_ERR_INVALID;
    GridProposal g;
    g.cdf = p->cdf.data();
    g.ratio = p->ratio.data();
    g.guide = p->guide.data();
    g.guide_bits = p->guide_bits;
    g.Mg = p->Mg;
    g.lambda = p->lambda;
    const double w_out = p->S / ((1.0 - p->lambda) * (double)P);
    for (int32_t l = 0; l < P; ++l) {
        int mm[3];
        w[l] = draw_kvector(seed, step, (uint32_t)l, p->acdf.data(), p->M, g, P, 0.0, w_out, mm);
        m[3 * l] = mm[0];
        m[3 * l + 1] = mm[1];
        m[3 * l + 2] = mm[2];
    }
    return RBMD_OK;
}

}  // extern "C"
