// md.hip -- launches of the velocity-Verlet MD-step kernels (bodies in md.h and,
// for the tile pass, pair.hip) for one handle and for replica batches.
//
// Reference behaviour followed (paths relative to the reference checkout):
//   velocity Verlet step          rbmd/framework/run/ExecutionNVT.cpp:133-161, 224-250,
//                                 rbmd/framework/run/worklet/RunWorklet.cpp:1764-1782, 1854-1877
#include "md.h"
#include "state.h"

namespace rbmd {

namespace {

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Thermostat constants and device state for the closing kick (type none when the
// handle has no thermostat).
ThermoArgs thermo_args(const Engine& e, double dt)
{
    ThermoArgs t;
    if (e.thermo_type == kThermoNone) return t;
    t.type = e.thermo_type;
    t.nblk = std::max(kforce_blocks(e), 1);
    t.part = e.thermo_part;
    t.cnt = e.thermo_cnt;
    t.st = e.thermo_st;
    t.noise = e.thermo_noise;
    t.t0 = e.thermo_t0;
    t.dt_over_damp = e.thermo_tdamp > 0 ? dt / e.thermo_tdamp : 0.0;
    t.inv_dof_kb = 1.0 / (thermo_dof(e) * e.thermo_kb);
    t.gamma = e.thermo_gamma;
    t.kbt = e.thermo_kbt;
    t.seed = e.thermo_seed;
    t.log = e.thermo_log;
    t.log_len = e.thermo_log_len;
    return t;
}

template <typename T>
KickDriftArgs<T> kickdrift_args(Engine& e, const RepState& st, double dt, const SampleArgs& sa)
{
    const double h = 0.5 * e.skin;
    KickDriftArgs<T> a{st.x,   st.y,   st.z,   st.vx,         st.vy,
                       st.vz,  (const T*)st.fx, (const T*)st.fy, (const T*)st.fz,
                       st.img, st.mass, dt,    (int)e.n,      e.box,
                       e.xr,   e.yr,   e.zr,   h * h,         e.flags,
                       e.q,    e.atom_rank /* inverse perm */, e.xs, sa};
    if (e.thermo_type == kThermoBerendsen || e.thermo_type == kThermoRescale)
        a.vscale = e.thermo_st;
    return a;
}

CellSortMDArgs cellsort_md_args(Engine& e, const RepState& st)
{
    return CellSortMDArgs{st.x,    st.y,    st.z,     e.q,         e.type,
                          (int)e.n, e.ncell, e.box,   md_bins(e),  e.cell_start,
                          e.perm,  e.atom_rank /* inverse perm */, e.scell,
                          e.ts,    e.xr,    e.yr,     e.zr,        e.xs_ref,
                          e.xs,    e.flags, e.stats};
}

template <typename T>
typename SortRho<T>::Args sortrho_args(Engine& e, const RepState& st, int* blocks)
{
    typename SortRho<T>::Args a;
    int nr = 0;
    if (has_kspace(e)) a.rho = rho_args<T>(e, st.x, st.y, st.z, &nr);
    else memset(&a.rho, 0, sizeof a.rho);
    a.nrho = nr;
    a.cs = cellsort_md_args(e, st);
    a.nsort = sortrho_cell_blocks(e.ncell);
    a.dims = make_int3(e.dims[0], e.dims[1], e.dims[2]);
    a.split = e.split;
    a.ncu = tile_order_ncu(e.ncell, e.split, e.num_cu);
    a.tile_order = e.tile_order;
    *blocks = nr + a.nsort + 1;
    return a;
}

template <typename T>
KForceArgs<T> close_args(Engine& e, const RepState& st, double dt)
{
    KForceArgs<T> a = kforce_args<T>(e, st.x, st.y, st.z, st.fx, st.fy, st.fz, st.vx, st.vy,
                                     st.vz, st.mass, dt, has_kspace(e));
    a.th = thermo_args(e, dt);
    return a;
}

template <typename T>
typename KForceDrift<T>::Args close_open_args(Engine& e, const RepState& st, double dt,
                                              const SampleArgs& sa_next)
{
    return typename KForceDrift<T>::Args{close_args<T>(e, st, dt),
                                         kickdrift_args<T>(e, st, dt, sa_next)};
}

// sampler workgroups first, then the atom workgroups
int kickdrift_blocks(const Engine& e, const SampleArgs& sa)
{
    return ceil_div(sa.P, kMdBlock) + std::max(ceil_div(e.n, kMdBlock), 1);
}

int close_open_blocks(const Engine& e, const SampleArgs& sa)
{
    return ceil_div(sa.P, kKfThreads) + std::max(kforce_blocks(e), 1);
}

template <typename T>
void md_open_t(Engine& e, const RepState& st, double dt, const SampleArgs& sa, bool fuse_thermo,
               hipStream_t s)
{
    KickDriftArgs<T> a = kickdrift_args<T>(e, st, dt, sa);
    if (fuse_thermo && e.thermo_type != kThermoNone) {
        a.th = thermo_args(e, dt);
        a.vscale = nullptr;
    }
    k_one<KickDrift<T>><<<kickdrift_blocks(e, sa), kMdBlock, 0, s>>>(a);
}

template <typename T>
void md_sort_rho_t(Engine& e, const RepState& st, hipStream_t s)
{
    int blocks;
    const auto a = sortrho_args<T>(e, st, &blocks);
    k_one<SortRho<T>><<<blocks, kRhoBlock, 0, s>>>(a);
}

template <typename T>
void md_close_t(Engine& e, const RepState& st, double dt, hipStream_t s)
{
    k_one<KForce<T, true>><<<std::max(kforce_blocks(e), 1), kKfThreads, 0, s>>>(
        close_args<T>(e, st, dt));
}

template <typename T>
void md_close_open_t(Engine& e, const RepState& st, double dt, const SampleArgs& sa,
                     hipStream_t s)
{
    k_one<KForceDrift<T>><<<close_open_blocks(e, sa), kKfThreads, 0, s>>>(
        close_open_args<T>(e, st, dt, sa));
}

template <typename T>
hipError_t md_prepare_t(Batch& b)
{
    const size_t R = b.es.size();
    std::vector<typename KickDrift<T>::Args> kd(R);
    std::vector<typename SortRho<T>::Args> sr(R);
    std::vector<KForceArgs<T>> kc(R);
    std::vector<typename KForceDrift<T>::Args> ko(R);
    std::vector<int> nkd(R), nsr(R), nkc(R), nko(R);
    for (size_t r = 0; r < R; ++r) {
        Engine& e = *b.es[r];
        const RepState& st = b.st[r];
        kd[r] = kickdrift_args<T>(e, st, b.dt, b.sa[r]);
        sr[r] = sortrho_args<T>(e, st, &nsr[r]);
        kc[r] = close_args<T>(e, st, b.dt);
        ko[r] = close_open_args<T>(e, st, b.dt, b.sa[r]);
        nkd[r] = kickdrift_blocks(e, b.sa[r]);
        nkc[r] = std::max(kforce_blocks(e), 1);
        nko[r] = close_open_blocks(e, b.sa[r]);
    }
    hipError_t err;
    if ((err = phase_upload(b.ph[BP_KICKDRIFT], kd, nkd, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_SORT_RHO], sr, nsr, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_KFORCE], kc, nkc, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_KFORCE_DRIFT], ko, nko, 0)) != hipSuccess) return err;
    // thermostats: one workgroup per thermostatted replica for the state reset and
    // the per-step reduction, the atoms of rescaling ones for the final scale
    std::vector<ThermoBegin::Args> tb(R);
    std::vector<ThermoReduce::Args> tr(R);
    std::vector<VScale::Args> te(R);
    std::vector<int> ntb(R, 0), ntr(R, 0), nte(R, 0);
    for (size_t r = 0; r < R; ++r) {
        Engine& e = *b.es[r];
        const RepState& st = b.st[r];
        tb[r] = ThermoBegin::Args{e.thermo_st, e.thermo_cnt};
        tr[r] = ThermoReduce::Args{thermo_args(e, b.dt), b.dt};
        te[r] = VScale::Args{st.vx, st.vy, st.vz, e.thermo_st, (int)e.n};
        if (e.thermo_type == kThermoNone) continue;
        ntb[r] = ntr[r] = 1;
        if (e.thermo_type == kThermoBerendsen || e.thermo_type == kThermoRescale)
            nte[r] = std::max(ceil_div(e.n, kMdBlock), 1);
    }
    if ((err = phase_upload(b.ph[BP_THERMO_BEGIN], tb, ntb, 0)) != hipSuccess) return err;
    if ((err = phase_upload(b.ph[BP_THERMO_REDUCE], tr, ntr, 0)) != hipSuccess) return err;
    return phase_upload(b.ph[BP_THERMO_END], te, nte, 0);
}

}  // namespace

// ================================================================= launchers

void launch_md_open(Engine& e, const RepState& st, double dt, const SampleArgs& sa,
                    hipStream_t s, bool fuse_thermo)
{
    ScopedTimer t(e, P_CHECK, s);
    if (e.precision == 32) md_open_t<float>(e, st, dt, sa, fuse_thermo, s);
    else md_open_t<double>(e, st, dt, sa, fuse_thermo, s);
}

void launch_md_sort_rho(Engine& e, const RepState& st, hipStream_t s)
{
    ScopedTimer t(e, P_RHO, s);
    if (e.precision == 32) md_sort_rho_t<float>(e, st, s);
    else md_sort_rho_t<double>(e, st, s);
}

void launch_md_close(Engine& e, const RepState& st, double dt, hipStream_t s)
{
    ScopedTimer t(e, P_KFORCE, s);
    if (e.precision == 32) md_close_t<float>(e, st, dt, s);
    else md_close_t<double>(e, st, dt, s);
}

void launch_md_close_open(Engine& e, const RepState& st, double dt, const SampleArgs& sa,
                          hipStream_t s)
{
    ScopedTimer t(e, P_KFORCE, s);
    if (e.precision == 32) md_close_open_t<float>(e, st, dt, sa, s);
    else md_close_open_t<double>(e, st, dt, sa, s);
}

void launch_thermo_begin(Engine& e, hipStream_t s)
{
    k_one<ThermoBegin><<<1, ThermoBegin::kThreads, 0, s>>>(
        ThermoBegin::Args{e.thermo_st, e.thermo_cnt});
}

void launch_thermo_reduce(Engine& e, double dt, hipStream_t s)
{
    k_one<ThermoReduce><<<1, ThermoReduce::kThreads, 0, s>>>(
        ThermoReduce::Args{thermo_args(e, dt), dt});
}

void launch_thermo_end(Engine& e, const RepState& st, hipStream_t s)
{
    if (e.thermo_type != kThermoBerendsen && e.thermo_type != kThermoRescale) return;
    k_one<VScale><<<std::max(ceil_div(e.n, kMdBlock), 1), kMdBlock, 0, s>>>(
        VScale::Args{st.vx, st.vy, st.vz, e.thermo_st, (int)e.n});
}

hipError_t md_batch_prepare(Batch& b)
{
    return b.precision == 32 ? md_prepare_t<float>(b) : md_prepare_t<double>(b);
}

void md_batch_launch(Batch& b, int phase, hipStream_t s)
{
    const bool f32 = b.precision == 32;
#define RBMD_MB(B)                                                                            \
    do {                                                                                      \
        if (f32) phase_launch<B<float>>(b.ph[phase], s);                                      \
        else phase_launch<B<double>>(b.ph[phase], s);                                         \
    } while (0)
    switch (phase) {
    case BP_KICKDRIFT: RBMD_MB(KickDrift); break;
    case BP_SORT_RHO: RBMD_MB(SortRho); break;
    case BP_KFORCE_DRIFT: RBMD_MB(KForceDrift); break;
    case BP_KFORCE:
        if (f32) phase_launch<KForce<float, true>>(b.ph[phase], s);
        else phase_launch<KForce<double, true>>(b.ph[phase], s);
        break;
    case BP_THERMO_BEGIN: phase_launch<ThermoBegin>(b.ph[phase], s); break;
    case BP_THERMO_REDUCE: phase_launch<ThermoReduce>(b.ph[phase], s); break;
    case BP_THERMO_END: phase_launch<VScale>(b.ph[phase], s); break;
    default: break;
    }
#undef RBMD_MB
}

}  // namespace rbmd
