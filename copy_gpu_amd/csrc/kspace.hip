// kspace.hip -- Random Batch Ewald (and full Ewald) k-space launches of the
// force pass: device k-sampler, rho(k) partials and final reduction, and the
// per-atom k-space force (bodies in kspace.h).
//
// Reference (paths relative to the reference checkout):
//   sampler target / normaliser S   rbmd/framework/model/RBEPSample.h:38-61, 87-93, 143-175
//   rho(k)                          rbmd/framework/run/worklet/RunWorklet.cpp:1616-1677,
//                                   rbmd/framework/run/ExecutionMD.cpp:412-448
//   RBE force                       RunWorklet.cpp:1508-1548, ExecForceFunction.h:247-269
//   Ewald force / energy            ExecForceFunction.h:226-245, output/ThermoOutput.cpp:209-255
#include "kspace.h"
#include "state.h"

namespace rbmd {

void rho_geometry(const Engine& e, int* atoms_per_block, int* nbx, int* nkt)
{
    const int n = (int)e.n;
    *nkt = (e.P + 63) / 64;
    int apb;
    if (e.rho_apb > 0) {
        apb = e.rho_apb;
    } else {
        int want = 2048 / *nkt;
        if (want < 1) want = 1;
        apb = (n + want - 1) / want;
    }
    // the partials [nbx][P][2] must fit the kpart buffer
    const int64_t max_nbx = e.P > 0 ? e.kpart_cap / (2 * (int64_t)e.P) : 1;
    if (max_nbx >= 1 && (int64_t)apb * max_nbx < (int64_t)n)
        apb = (int)(((int64_t)n + max_nbx - 1) / max_nbx);
    apb = ((apb + 63) / 64) * 64;
    if (apb < 64) apb = 64;
    *atoms_per_block = apb;
    *nbx = (n + apb - 1) / apb;
    if (*nbx < 1) *nbx = 1;
}

namespace {

// ---------------------------------------------------------------- sampler
struct Sample {
    static constexpr int kThreads = 256;
    using Args = SampleArgs;
    __device__ static void run(const SampleArgs& sa, int blk, int)
    {
        const int l = blk * 256 + threadIdx.x;
        if (l >= sa.P) return;
        const uint64_t step = sa.step_dev ? (uint64_t)(*sa.step_dev) : sa.step;
        const int Ms[3] = {sa.M.x, sa.M.y, sa.M.z};
        int m[3];
        const double w =
            draw_kvector(sa.seed, step, (uint32_t)l, sa.cdf, Ms, sa.gp, sa.P, sa.w, sa.w_out, m);
        sa.km[3 * l] = m[0];
        sa.km[3 * l + 1] = m[1];
        sa.km[3 * l + 2] = m[2];
        sa.kw[l] = w;
    }
};

__global__ void k_fill(double* __restrict__ a, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

template <typename T>
hipError_t kspace_prepare_t(Batch& b)
{
    const size_t R = b.es.size();
    std::vector<typename RhoPartial<T>::Args> ra(R);
    std::vector<int> nr(R);
    for (size_t r = 0; r < R; ++r) {
        const RepState& st = b.st[r];
        ra[r] = rho_args<T>(*b.es[r], st.x, st.y, st.z, &nr[r]);
    }
    return phase_upload(b.ph[BP_RHO], ra, nr, 0);
}

}  // namespace

// ================================================================= launchers

SampleArgs sample_args(const Engine& e, int P, uint64_t seed, const uint64_t* step_host)
{
    SampleArgs sa;
    sa.P = P;
    sa.seed = seed;
    sa.step = step_host ? *step_host : 0ull;
    sa.step_dev = step_host ? nullptr : e.d_step;
    sa.cdf = e.kcdf;
    sa.M = make_int3(e.kM[0], e.kM[1], e.kM[2]);
    sa.w = e.S / (double)P;
    sa.gp = e.gp;
    sa.w_out = e.gp.cdf ? e.S / ((1.0 - e.gp.lambda) * (double)P) : 0.0;
    sa.km = e.km;
    sa.kw = e.kw;
    return sa;
}

void launch_sample(Engine& e, int P, uint64_t seed, const uint64_t* step_host, hipStream_t s)
{
    ScopedTimer t(e, P_SAMPLE, s);
    k_one<Sample><<<(P + 255) / 256, 256, 0, s>>>(sample_args(e, P, seed, step_host));
}

void launch_kweights_default(Engine& e, hipStream_t s)
{
    k_fill<<<(e.P + 255) / 256, 256, 0, s>>>(e.kw, e.P, e.S / (double)e.P);
}

void launch_rho(Engine& e, const double* x, const double* y, const double* z, hipStream_t s)
{
    int blocks;
    unsigned long long* step_dev = e.capturing ? e.d_step : nullptr;
    {
        ScopedTimer t(e, P_RHO, s);
        if (e.precision == 32) {
            const auto a = rho_args<float>(e, x, y, z, &blocks);
            k_one<RhoPartial<float>><<<blocks, kRhoBlock, 0, s>>>(a);
        } else {
            const auto a = rho_args<double>(e, x, y, z, &blocks);
            k_one<RhoPartial<double>><<<blocks, kRhoBlock, 0, s>>>(a);
        }
    }
    {
        ScopedTimer t(e, P_RHO_FINAL, s);
        if (e.precision == 32)
            k_one<RhoFinal<float>><<<e.P, RhoFinal<float>::kThreads, 0, s>>>(
                rho_final_args<float>(e, step_dev));
        else
            k_one<RhoFinal<double>><<<e.P, RhoFinal<double>::kThreads, 0, s>>>(
                rho_final_args<double>(e, step_dev));
    }
}

void launch_kforce(Engine& e, const double* x, const double* y, const double* z, void* fx,
                   void* fy, void* fz, hipStream_t s)
{
    ScopedTimer t(e, P_KFORCE, s);
    if (e.precision == 32)
        k_one<KForce<float, false>><<<kforce_blocks(e), kKfThreads, 0, s>>>(
            kforce_args<float>(e, x, y, z, fx, fy, fz, nullptr, nullptr, nullptr, nullptr, 0.0));
    else
        k_one<KForce<double, false>><<<kforce_blocks(e), kKfThreads, 0, s>>>(
            kforce_args<double>(e, x, y, z, fx, fy, fz, nullptr, nullptr, nullptr, nullptr, 0.0));
}

hipError_t kspace_batch_prepare(Batch& b)
{
    return b.precision == 32 ? kspace_prepare_t<float>(b) : kspace_prepare_t<double>(b);
}

void kspace_batch_launch(Batch& b, int phase, hipStream_t s)
{
    const bool f32 = b.precision == 32;
    switch (phase) {
    case BP_RHO:
        if (f32) phase_launch<RhoPartial<float>>(b.ph[phase], s);
        else phase_launch<RhoPartial<double>>(b.ph[phase], s);
        break;
    default: break;
    }
}

}  // namespace rbmd
