// kcompute.hip -- the RBE / Ewald k-space step of a given k-list without a force
// handle: structure factor rho(k) = sum_j q_j exp(i k.r_j) (per-workgroup partials,
// then a fixed-order FP64 reduction per k with the force coefficients and per-k
// energy / virial terms) and the k-space force scatter, with the kernel bodies of
// the force pass (kspace.h: RhoPartial, RhoFinal, KForce).
//
// Reference (paths relative to the reference checkout):
//   ExecutionMD::ComputeRBEEleForce      rbmd/framework/run/ExecutionMD.cpp:389-410
//   ComputeNewChargeStructureFactorRBE   rbmd/framework/run/ExecutionMD.cpp:412-448
//   rho(k) worklet                       rbmd/framework/run/worklet/RunWorklet.cpp:1616-1677
//   RBE force worklet                    RunWorklet.cpp:1508-1548, ExecForceFunction.h:247-269
//
// rbmd_kspace_compute allocates its scratch (partials, rho, per-k coefficients and
// terms, weights) for the call.  rbmd_kspace_compute_noalloc runs the same kernels
// on a fixed block of device memory (`scratch` below, 4 MiB per device) and
// allocates nothing, for callers that cannot call hipMalloc.

#include <mutex>

#include "kspace.h"
#include "state.h"

namespace rbmd {

double rbe_S(double alpha, const double* L);

namespace {

// Scratch of one call, carved from one block of device memory: weights [P],
// partial structure factors [nbx][P][2], rho [P][2], coefficients KCoef<T>[P],
// per-k energy / virial terms [P][7].  Offsets are multiples of 256 bytes.
struct KsLayout {
    size_t kw, kpart, rho, coef, kev, total;
};

inline size_t up256(size_t v) { return (v + 255) / 256 * 256; }

// The force pass's rho(k) geometry for n atoms and P k-vectors (kspace.hip
// rho_geometry with the default workgroup size): about 2048 workgroups of 64 k's
// by atom tiles, so the partials hold at most 2048 x 64 x 2 doubles.
void ks_geometry(int64_t n, int P, int* apb, int* nbx, int* nkt)
{
    *nkt = (P + 63) / 64;
    int want = 2048 / *nkt;
    if (want < 1) want = 1;
    int a = (int)((n + want - 1) / want);
    a = ((a + 63) / 64) * 64;
    if (a < 64) a = 64;
    *apb = a;
    *nbx = (int)((n + a - 1) / a);
    if (*nbx < 1) *nbx = 1;
}

KsLayout ks_layout(int64_t n, int P)
{
    int apb, nbx, nkt;
    ks_geometry(n, P, &apb, &nbx, &nkt);
    KsLayout l;
    size_t o = 0;
    l.kw = o;
    o += up256((size_t)P * 8);
    l.kpart = o;
    o += up256((size_t)nbx * P * 2 * 8);
    l.rho = o;
    o += up256((size_t)P * 2 * 8);
    l.coef = o;
    o += up256((size_t)P * sizeof(KCoef<double>));
    l.kev = o;
    o += up256((size_t)P * 7 * 8);
    l.total = o;
    return l;
}

}  // namespace

namespace {

// The k-space energy and virial: per-k terms [P][7] summed in a fixed order.
__global__ __launch_bounds__(256) void k_kev_sum(const double* __restrict__ kev, int P,
                                                 double* __restrict__ out)
{
    __shared__ double red[256];
    const int t = threadIdx.x;
    for (int c = 0; c < 7; ++c) {
        double v = 0.0;
        for (int k = t; k < P; k += 256) v += kev[(size_t)k * 7 + c];
        red[t] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) red[t] += red[t + o];
            __syncthreads();
        }
        if (t == 0) out[c] = red[0];
        __syncthreads();
    }
}

__global__ void k_fill_w(double* __restrict__ a, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_copy_rho(const double* __restrict__ src, int n, double* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

struct KsCall {
    double lo[3], L[3], alpha, qqr2e;
    int precision;
    int64_t n;
    const double *x, *y, *z, *q;
    int P;
    const int32_t* m;
    const double* w;
    void *fx, *fy, *fz;
    double* rho_out;
    double* ev_out;
};

template <typename T>
void ks_launch(const KsCall& c, char* buf, hipStream_t s)
{
    const KsLayout l = ks_layout(c.n, c.P);
    double* kw = (double*)(buf + l.kw);
    double* kpart = (double*)(buf + l.kpart);
    double* rho = (double*)(buf + l.rho);
    KCoef<T>* coef = (KCoef<T>*)(buf + l.coef);
    double* kev = (double*)(buf + l.kev);
    Box box;
    for (int d = 0; d < 3; ++d) {
        box.lo[d] = c.lo[d];
        box.L[d] = c.L[d];
        box.invL[d] = 1.0 / c.L[d];
    }
    const double* w = c.w;
    if (!w) {
        k_fill_w<<<(c.P + 255) / 256, 256, 0, s>>>(kw, c.P, rbe_S(c.alpha, c.L) / (double)c.P);
        w = kw;
    }
    int apb, nbx, nkt;
    ks_geometry(c.n, c.P, &apb, &nbx, &nkt);
    k_one<RhoPartial<T>><<<nbx * nkt, kRhoBlock, 0, s>>>(typename RhoPartial<T>::Args{
        c.x, c.y, c.z, c.q, (int)c.n, apb, nbx, box, c.m, c.P, kpart});
    k_one<RhoFinal<T>><<<c.P, RhoFinal<T>::kThreads, 0, s>>>(RhoFinalArgs<T>{
        kpart, nbx, c.P, c.m, w, box, c.alpha, c.qqr2e, rho, coef, kev, nullptr});
    KForceArgs<T> ka{};
    ka.x = c.x;
    ka.y = c.y;
    ka.z = c.z;
    ka.q = c.q;
    ka.n = (int)c.n;
    ka.box = box;
    ka.coef = coef;
    ka.P = c.P;
    ka.fx = (T*)c.fx;
    ka.fy = (T*)c.fy;
    ka.fz = (T*)c.fz;
    ka.slices = kf_slices(c.n, c.P);
    if (c.n > 0) k_one<KForce<T, false>><<<kf_blocks(c.n, ka.slices), kKfThreads, 0, s>>>(ka);
    if (c.rho_out) k_copy_rho<<<(2 * c.P + 255) / 256, 256, 0, s>>>(rho, 2 * c.P, c.rho_out);
    if (c.ev_out) k_kev_sum<<<1, 256, 0, s>>>(kev, c.P, c.ev_out);
}

thread_local std::string g_ks_err;

int ks_fail(int code, const std::string& msg)
{
    g_ks_err = msg;
    return code;
}

int ks_check(const KsCall& c)
{
    if (c.precision != 32 && c.precision != 64)
        return ks_fail(RBMD_ERR_INVALID, "precision must be 32 or 64");
    if (c.n < 0 || c.n > (int64_t)INT32_MAX / 2 || c.P < 1)
        return ks_fail(RBMD_ERR_INVALID, "n must be in [0, 2^30) and P >= 1");
    if (!(c.alpha > 0)) return ks_fail(RBMD_ERR_INVALID, "alpha must be positive");
    for (int d = 0; d < 3; ++d)
        if (!(c.L[d] > 0)) return ks_fail(RBMD_ERR_INVALID, "box lengths must be positive");
    if (!c.m) return ks_fail(RBMD_ERR_INVALID, "null k-vector array");
    if (c.n > 0 && (!c.x || !c.y || !c.z || !c.q || !c.fx || !c.fy || !c.fz))
        return ks_fail(RBMD_ERR_INVALID, "null atom or force array");
    return RBMD_OK;
}

KsCall ks_call(const rbmd_kspace_params* p, int64_t n, const double* x, const double* y,
               const double* z, const double* q, int32_t P, const int32_t* m, const double* w,
               void* fx, void* fy, void* fz, double* rho, double* ev)
{
    KsCall c{};
    for (int d = 0; d < 3; ++d) {
        c.lo[d] = p->box_lo[d];
        c.L[d] = p->box_len[d];
    }
    c.alpha = p->alpha;
    c.qqr2e = p->qqr2e;
    c.precision = p->precision;
    c.n = n;
    c.x = x;
    c.y = y;
    c.z = z;
    c.q = q;
    c.P = P;
    c.m = m;
    c.w = w;
    c.fx = fx;
    c.fy = fy;
    c.fz = fz;
    c.rho_out = rho;
    c.ev_out = ev;
    return c;
}

void ks_dispatch(const KsCall& c, char* buf, hipStream_t s)
{
    if (c.precision == 32) ks_launch<float>(c, buf, s);
    else ks_launch<double>(c, buf, s);
}

}  // namespace
}  // namespace rbmd

// Fixed scratch of rbmd_kspace_compute_noalloc, one copy per device, held until the
// process ends.  4 MiB covers every P <= 8192 at any n: the partials hold at most
// 2048 x 64 x 2 doubles (2 MiB, ks_geometry) and the per-k arrays 120 bytes per k.
// In host code `scratch` is only the symbol's host-side stand-in; kernels take the
// device address hipGetSymbolAddress returns (ks_scratch_ptr).
constexpr size_t kKsScratchBytes = (size_t)4 << 20;
constexpr int kKsScratchMaxP = 8192;
static __device__ __align__(256) unsigned char scratch[kKsScratchBytes];

namespace rbmd {
namespace {

constexpr int kKsMaxDevices = 64;
std::mutex g_ks_scratch_mu;
char* g_ks_scratch[kKsMaxDevices] = {};

// Device address of the current device's `scratch`, looked up once per device.
int ks_scratch_ptr(char** out)
{
    int dev = -1;
    hipError_t st = hipGetDevice(&dev);
    if (st != hipSuccess)
        return ks_fail(RBMD_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(st));
    if (dev < 0 || dev >= kKsMaxDevices) return ks_fail(RBMD_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lk(g_ks_scratch_mu);
    if (!g_ks_scratch[dev]) {
        void* p = nullptr;
        st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess || !p)
            return ks_fail(RBMD_ERR_HIP,
                           std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st));
        g_ks_scratch[dev] = (char*)p;
    }
    *out = g_ks_scratch[dev];
    return RBMD_OK;
}

}  // namespace
}  // namespace rbmd

using namespace rbmd;

extern "C" {

const char* rbmd_kspace_last_error(void) { return g_ks_err.c_str(); }

int rbmd_kspace_compute(const rbmd_kspace_params* p, int64_t n, const double* x, const double* y,
                        const double* z, const double* q, int32_t P, const int32_t* m,
                        const double* w, void* fx, void* fy, void* fz, double* rho, double* ev,
                        void* stream)
{
    g_ks_err.clear();
    if (!p) return ks_fail(RBMD_ERR_INVALID, "null parameters");
    const KsCall c = ks_call(p, n, x, y, z, q, P, m, w, fx, fy, fz, rho, ev);
    int rc = ks_check(c);
    if (rc != RBMD_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    char* buf = nullptr;
    hipError_t st = hipMalloc((void**)&buf, ks_layout(n, P).total);
    if (st != hipSuccess) return ks_fail(RBMD_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(st));
    ks_dispatch(c, buf, s);
    st = hipGetLastError();
    hipError_t st2 = hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (st != hipSuccess) return ks_fail(RBMD_ERR_HIP, std::string("launch: ") + hipGetErrorString(st));
    if (st2 != hipSuccess)
        return ks_fail(RBMD_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(st2));
    return RBMD_OK;
}

int rbmd_kspace_compute_noalloc(const rbmd_kspace_params* p, int64_t n, const double* x,
                                const double* y, const double* z, const double* q, int32_t P,
                                const int32_t* m, const double* w, void* fx, void* fy, void* fz,
                                double* rho, double* ev, void* stream)
{
    g_ks_err.clear();
    if (!p) return ks_fail(RBMD_ERR_INVALID, "null parameters");
    const KsCall c = ks_call(p, n, x, y, z, q, P, m, w, fx, fy, fz, rho, ev);
    int rc = ks_check(c);
    if (rc != RBMD_OK) return rc;
    if (P > kKsScratchMaxP || ks_layout(n, P).total > kKsScratchBytes)
        return ks_fail(RBMD_ERR_INVALID,
                       "rbmd_kspace_compute_noalloc: P = " + std::to_string(P) +
                           " exceeds the fixed scratch (P <= " +
                           std::to_string(kKsScratchMaxP) + ")");
    char* buf = nullptr;
    rc = ks_scratch_ptr(&buf);
    if (rc != RBMD_OK) return rc;
    ks_dispatch(c, buf, (hipStream_t)stream);
    const hipError_t st = hipGetLastError();
    if (st != hipSuccess) return ks_fail(RBMD_ERR_HIP, std::string("launch: ") + hipGetErrorString(st));
    return RBMD_OK;
}

}  // extern "C"
