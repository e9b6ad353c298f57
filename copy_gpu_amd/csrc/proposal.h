// proposal.h -- host-side state of a learned k-space proposal (tables shared
// by the host sampler and, after rbmd_force_set_proposal, the device sampler).
#pragma once
#include <stdint.h>

#include <vector>

struct rbmd_proposal_s {
    int Mg = 0;
    double lambda = 0, alpha = 0, L[3] = {0, 0, 0};
    double S = 0;               // analytic normaliser sum_{m != 0} a(m)
    std::vector<double> cdf;    // inclusive running sum of the grid pmf, closes at 1
    std::vector<double> ratio;  // a(m) / q_mix(m) per grid entry
    std::vector<int32_t> guide; // guide table of the inverse-CDF lookup
    int guide_bits = 0;
    std::vector<double> acdf;   // analytic per-axis CDF tables
    int M[3] = {0, 0, 0};
};
