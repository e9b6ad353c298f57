// mlp.hip -- learned k-space importance proposal: a 4 -> 64 -> 64 -> 1 MLP
// (SiLU, softplus output, BF16 weights, FP32 accumulation) with hand-written
// forward, backward and Adam kernels.
//
// The reference has no learned component; the proposal replaces its analytic
// Gaussian sampler (rbmd/framework/model/RBEPSample.h:143-175) as the source of
// the RBE k-vectors, with importance weights keeping the force estimator
// unbiased (weights are computed by the force engine's sampler, kspace.hip).
//
// Parameter vector (4545 FP32 values, master copy; a BF16 copy feeds the
// forward pass): W1[64][4], b1[64], W2[64][64], b2[64], w3[64], b3.
// Input features: (|kx|, |ky|, |kz|, |k|) / k_scale.  Output h = softplus(z).
// Training loss over a batch: sum_s w_s (log h_s - y_s)^2.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/rbmd_force.h"
#include "engine.h"

namespace rbmd {

namespace {

constexpr int kH = 64;
constexpr int kNP = kH * 4 + kH + kH * kH + kH + kH + 1;  // 4545
constexpr int oW1 = 0, oB1 = 256, oW2 = 320, oB2 = 4416, oW3 = 4480, oB3 = 4544;
constexpr int kGradBlocks = 256;
constexpr int kSub = 128;  // samples per sub-batch of the loss/gradient kernel
constexpr int kRow = kH + 1;  // padded LDS row (bank-conflict free column writes)

__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ inline float silu(float x) { return x / (1.0f + expf(-x)); }
__device__ inline float dsilu(float x)
{
    const float s = 1.0f / (1.0f + expf(-x));
    return s * (1.0f + x * (1.0f - s));
}
__device__ inline float softplus(float z) { return z > 20.0f ? z : log1pf(expf(z)); }

// Stage the BF16 parameters as floats in LDS.
__device__ inline void stage_params(const uint16_t* __restrict__ p16, float* __restrict__ W)
{
    for (int e = threadIdx.x; e < kNP; e += blockDim.x) W[e] = bf2f(p16[e]);
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_mlp_forward(const uint16_t* __restrict__ p16,
                                                     const float* __restrict__ feat, int64_t B,
                                                     float* __restrict__ out)
{
    __shared__ float W[kNP];
    stage_params(p16, W);
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < B;
         s += (int64_t)gridDim.x * blockDim.x) {
        const float f0 = feat[4 * s], f1 = feat[4 * s + 1], f2 = feat[4 * s + 2],
                    f3 = feat[4 * s + 3];
        float a1[kH];
#pragma unroll
        for (int j = 0; j < kH; ++j) {
            const float* w = W + oW1 + 4 * j;
            a1[j] = silu(W[oB1 + j] + w[0] * f0 + w[1] * f1 + w[2] * f2 + w[3] * f3);
        }
        float z = W[oB3];
        for (int j = 0; j < kH; ++j) {
            float acc = W[oB2 + j];
            const float* w = W + oW2 + kH * j;
#pragma unroll
            for (int i = 0; i < kH; ++i) acc += w[i] * a1[i];
            z += W[oW3 + j] * silu(acc);
        }
        out[s] = softplus(z);
    }
}

// Forward + backward of the weighted log-MSE loss.  Each block loops over
// sub-batches of 128 samples: per-sample work by threads 0..127 (activations
// and back-propagated signals into LDS), then all threads accumulate the
// parameter gradients they own (entries tid, tid + 256, ...) over the
// sub-batch.  Block partials go to gpart[block][4545], loss to lossp[block].
__global__ __launch_bounds__(256) void k_mlp_loss_grad(const uint16_t* __restrict__ p16,
                                                       const float* __restrict__ feat,
                                                       const float* __restrict__ y,
                                                       const float* __restrict__ wt, int64_t B,
                                                       float* __restrict__ gpart,
                                                       double* __restrict__ lossp)
{
    extern __shared__ float sm[];
    float* W = sm;                     // kNP
    float* A1 = W + kNP;               // kSub x kRow   activations of layer 1
    float* A2 = A1 + kSub * kRow;      // kSub x kRow   activations of layer 2
    float* D1 = A2 + kSub * kRow;      // kSub x kRow   dL/dpre1
    float* D2 = D1 + kSub * kRow;      // kSub x kRow   dL/dpre2
    float* F = D2 + kSub * kRow;       // kSub x 4      features
    float* DZ = F + kSub * 4;          // kSub          dL/dz
    __shared__ double lred[256];
    stage_params(p16, W);
    const int tid = threadIdx.x;
    constexpr int kOwn = (kNP + 255) / 256;
    float g[kOwn];
#pragma unroll
    for (int r = 0; r < kOwn; ++r) g[r] = 0.0f;
    double loss = 0.0;
    const int64_t nsub = (B + kSub - 1) / kSub;
    for (int64_t sb = blockIdx.x; sb < nsub; sb += gridDim.x) {
        if (tid < kSub) {
            const int64_t s = sb * kSub + tid;
            const bool ok = s < B;
            float f[4] = {0, 0, 0, 0};
            if (ok)
                for (int i = 0; i < 4; ++i) f[i] = feat[4 * s + i];
            for (int i = 0; i < 4; ++i) F[tid * 4 + i] = f[i];
            float a1[kH];
#pragma unroll
            for (int j = 0; j < kH; ++j) {
                const float* w = W + oW1 + 4 * j;
                const float pre = W[oB1 + j] + w[0] * f[0] + w[1] * f[1] + w[2] * f[2] + w[3] * f[3];
                a1[j] = silu(pre);
                A1[tid * kRow + j] = a1[j];
                D1[tid * kRow + j] = dsilu(pre);  // d silu / d pre1, scaled below
            }
            float z = W[oB3];
            for (int j = 0; j < kH; ++j) {
                float acc = W[oB2 + j];
                const float* w = W + oW2 + kH * j;
#pragma unroll
                for (int i = 0; i < kH; ++i) acc += w[i] * a1[i];
                const float a2 = silu(acc);
                A2[tid * kRow + j] = a2;
                D2[tid * kRow + j] = dsilu(acc);
                z += W[oW3 + j] * a2;
            }
            float dz = 0.0f;
            if (ok) {
                // log softplus(z) and its derivative sigmoid(z) / softplus(z); below
                // z = -15 both equal z and 1 to FP32 precision (and softplus(z)
                // would underflow further down)
                const float h = softplus(z);
                const bool tiny = z < -15.0f;
                const float t = (tiny ? z : logf(h)) - y[s];
                loss += (double)wt[s] * (double)t * (double)t;
                const float dlog = tiny ? 1.0f : 1.0f / ((1.0f + expf(-z)) * h);
                dz = 2.0f * wt[s] * t * dlog;
            }
            DZ[tid] = dz;
            float d2[kH];
#pragma unroll
            for (int j = 0; j < kH; ++j) {
                d2[j] = dz * W[oW3 + j] * D2[tid * kRow + j];
                D2[tid * kRow + j] = d2[j];
            }
            for (int i = 0; i < kH; ++i) {
                float acc = 0.0f;
#pragma unroll
                for (int j = 0; j < kH; ++j) acc += W[oW2 + kH * j + i] * d2[j];
                D1[tid * kRow + i] *= acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kOwn; ++r) {
            const int e = tid + 256 * r;
            if (e >= kNP) break;
            float acc = 0.0f;
            if (e < oB1) {
                const int j = e >> 2, i = e & 3;
                for (int s = 0; s < kSub; ++s) acc += D1[s * kRow + j] * F[s * 4 + i];
            } else if (e < oW2) {
                const int j = e - oB1;
                for (int s = 0; s < kSub; ++s) acc += D1[s * kRow + j];
            } else if (e < oB2) {
                const int j = (e - oW2) / kH, i = (e - oW2) % kH;
                for (int s = 0; s < kSub; ++s) acc += D2[s * kRow + j] * A1[s * kRow + i];
            } else if (e < oW3) {
                const int j = e - oB2;
                for (int s = 0; s < kSub; ++s) acc += D2[s * kRow + j];
            } else if (e < oB3) {
                const int j = e - oW3;
                for (int s = 0; s < kSub; ++s) acc += DZ[s] * A2[s * kRow + j];
            } else {
                for (int s = 0; s < kSub; ++s) acc += DZ[s];
            }
            g[r] += acc;
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kOwn; ++r) {
        const int e = tid + 256 * r;
        if (e < kNP) gpart[(size_t)blockIdx.x * kNP + e] = g[r];
    }
    lred[tid] = loss;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) lred[tid] += lred[tid + o];
        __syncthreads();
    }
    if (tid == 0) lossp[blockIdx.x] = lred[0];
}

// Deterministic sum of the block partials (fixed order).
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ gpart, int nb,
                                                    const double* __restrict__ lossp,
                                                    float* __restrict__ grad,
                                                    double* __restrict__ loss)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kNP) {
        float acc = 0.0f;
        for (int b = 0; b < nb; ++b) acc += gpart[(size_t)b * kNP + e];
        grad[e] = acc;
    }
    if (e == 0) {
        double l = 0.0;
        for (int b = 0; b < nb; ++b) l += lossp[b];
        *loss = l;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; refreshes the BF16 copy.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ p, uint16_t* __restrict__ p16,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ g, float lr, float b1,
                                                  float b2, float eps, float c1, float c2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kNP) return;
    const float ge = g[e];
    const float mm = b1 * m[e] + (1.0f - b1) * ge;
    const float vv = b2 * v[e] + (1.0f - b2) * ge * ge;
    m[e] = mm;
    v[e] = vv;
    const float pe = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    p[e] = pe;
    p16[e] = f2bf(pe);
}

__global__ void k_mlp_cast(const float* __restrict__ p, uint16_t* __restrict__ p16)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kNP) p16[e] = f2bf(p[e]);
}

constexpr size_t kGradLds =
    sizeof(float) * ((size_t)kNP + 4 * (size_t)kSub * kRow + (size_t)kSub * 4 + kSub);

}  // namespace

}  // namespace rbmd

using namespace rbmd;

struct rbmd_mlp_s {
    float* p = nullptr;
    uint16_t* p16 = nullptr;
    float *m = nullptr, *v = nullptr, *grad = nullptr, *gpart = nullptr;
    double *lossp = nullptr, *loss = nullptr;
    int64_t step = 0;
    std::string err;
};

static int mlp_fail(rbmd_mlp_t h, hipError_t e, const char* what)
{
    h->err = std::string(what) + ": " + hipGetErrorString(e);
    return RBMD_ERR_HIP;
}

#define MLP_TRY(expr)                                  \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return mlp_fail(h, _e, #expr); \
    } while (0)

extern "C" {

int rbmd_mlp_nparams(void) { return kNP; }

int rbmd_mlp_create(uint64_t seed, rbmd_mlp_t* out)
{
    if (!out) return RBMD_ERR_INVALID;
    rbmd_mlp_t h = new rbmd_mlp_s();
    *out = nullptr;
    // Glorot-uniform weights from Philox, zero biases, b3 = log(e - 1) so h = 1 at start
    std::vector<float> p(kNP, 0.0f);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    auto uni = [&](uint32_t idx) {
        uint32_t c[4] = {idx, 0x4D4C5031u, 0u, 0u}, r[4];
        philox4x32_10(c, key, r);
        return (float)u53(r[0], r[1]) * 2.0f - 1.0f;
    };
    const float s1 = sqrtf(6.0f / (4 + kH)), s2 = sqrtf(6.0f / (kH + kH)), s3 = sqrtf(6.0f / (kH + 1));
    for (int e = 0; e < kH * 4; ++e) p[oW1 + e] = s1 * uni(oW1 + e);
    for (int e = 0; e < kH * kH; ++e) p[oW2 + e] = s2 * uni(oW2 + e);
    for (int e = 0; e < kH; ++e) p[oW3 + e] = s3 * uni(oW3 + e);
    p[oB3] = logf(expf(1.0f) - 1.0f);
    int rc = RBMD_OK;
    hipError_t e;
    if ((e = hipMalloc((void**)&h->p, kNP * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&h->p16, kNP * 2)) != hipSuccess ||
        (e = hipMalloc((void**)&h->m, kNP * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&h->v, kNP * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&h->grad, kNP * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&h->gpart, (size_t)kGradBlocks * kNP * 4)) != hipSuccess ||
        (e = hipMalloc((void**)&h->lossp, kGradBlocks * 8)) != hipSuccess ||
        (e = hipMalloc((void**)&h->loss, 8)) != hipSuccess ||
        (e = hipMemcpy(h->p, p.data(), kNP * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(h->m, 0, kNP * 4)) != hipSuccess ||
        (e = hipMemset(h->v, 0, kNP * 4)) != hipSuccess) {
        rc = RBMD_ERR_HIP;
    }
    if (rc == RBMD_OK &&
        (e = hipFuncSetAttribute((const void*)k_mlp_loss_grad,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGradLds)) !=
            hipSuccess) {
        h->err = std::string("dynamic LDS of the gradient kernel: ") + hipGetErrorString(e);
        rc = RBMD_ERR_HIP;
    }
    if (rc == RBMD_OK) {
        k_mlp_cast<<<(kNP + 255) / 256, 256>>>(h->p, h->p16);
        if (hipDeviceSynchronize() != hipSuccess) rc = RBMD_ERR_HIP;
    }
    if (rc != RBMD_OK) {
        rbmd_mlp_destroy(h);
        return rc;
    }
    *out = h;
    return RBMD_OK;
}

void rbmd_mlp_destroy(rbmd_mlp_t h)
{
    if (!h) return;
    void* ptrs[] = {h->p, h->p16, h->m, h->v, h->grad, h->gpart, h->lossp, h->loss};
    for (void* q : ptrs)
        if (q) hipFree(q);
    delete h;
}

const char* rbmd_mlp_last_error(rbmd_mlp_t h) { return h ? h->err.c_str() : "null handle"; }

int rbmd_mlp_get_params(rbmd_mlp_t h, float* params)
{
    if (!h || !params) return RBMD_ERR_INVALID;
    MLP_TRY(hipDeviceSynchronize());
    MLP_TRY(hipMemcpy(params, h->p, kNP * 4, hipMemcpyDeviceToHost));
    return RBMD_OK;
}

int rbmd_mlp_set_params(rbmd_mlp_t h, const float* params)
{
    if (!h || !params) return RBMD_ERR_INVALID;
    MLP_TRY(hipMemcpy(h->p, params, kNP * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemset(h->m, 0, kNP * 4));
    MLP_TRY(hipMemset(h->v, 0, kNP * 4));
    h->step = 0;
    k_mlp_cast<<<(kNP + 255) / 256, 256>>>(h->p, h->p16);
    MLP_TRY(hipGetLastError());
    return RBMD_OK;
}

int rbmd_mlp_forward(rbmd_mlp_t h, const float* feat, int64_t B, float* out, void* stream)
{
    if (!h || B < 0 || (B > 0 && (!feat || !out))) return RBMD_ERR_INVALID;
    if (B == 0) return RBMD_OK;
    int blocks = (int)std::min<int64_t>((B + 255) / 256, 4096);
    k_mlp_forward<<<blocks, 256, 0, (hipStream_t)stream>>>(h->p16, feat, B, out);
    MLP_TRY(hipGetLastError());
    return RBMD_OK;
}

static int mlp_grad(rbmd_mlp_t h, const float* feat, const float* y, const float* w, int64_t B,
                    hipStream_t s)
{
    k_mlp_loss_grad<<<kGradBlocks, 256, kGradLds, s>>>(h->p16, feat, y, w, B, h->gpart, h->lossp);
    k_mlp_reduce<<<(kNP + 255) / 256, 256, 0, s>>>(h->gpart, kGradBlocks, h->lossp, h->grad,
                                                    h->loss);
    MLP_TRY(hipGetLastError());
    return RBMD_OK;
}

int rbmd_mlp_loss_grad(rbmd_mlp_t h, const float* feat, const float* y, const float* w,
                       int64_t B, double* loss, float* grad, void* stream)
{
    if (!h || B < 1 || !feat || !y || !w) return RBMD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int rc = mlp_grad(h, feat, y, w, B, s);
    if (rc) return rc;
    MLP_TRY(hipStreamSynchronize(s));
    if (loss) MLP_TRY(hipMemcpy(loss, h->loss, 8, hipMemcpyDeviceToHost));
    if (grad) MLP_TRY(hipMemcpy(grad, h->grad, kNP * 4, hipMemcpyDeviceToHost));
    return RBMD_OK;
}

int rbmd_mlp_train(rbmd_mlp_t h, const float* feat, const float* y, const float* w, int64_t B,
                   int32_t iters, double lr, double* final_loss, void* stream)
{
    if (!h || B < 1 || iters < 0 || !feat || !y || !w || !(lr > 0)) return RBMD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int32_t it = 0; it < iters; ++it) {
        int rc = mlp_grad(h, feat, y, w, B, s);
        if (rc) return rc;
        h->step += 1;
        const float c1 = 1.0f - powf(b1, (float)h->step), c2 = 1.0f - powf(b2, (float)h->step);
        k_mlp_adam<<<(kNP + 255) / 256, 256, 0, s>>>(h->p, h->p16, h->m, h->v, h->grad,
                                                      (float)lr, b1, b2, eps, c1, c2);
    }
    MLP_TRY(hipGetLastError());
    if (final_loss) {
        int rc = mlp_grad(h, feat, y, w, B, s);
        if (rc) return rc;
        MLP_TRY(hipStreamSynchronize(s));
        MLP_TRY(hipMemcpy(final_loss, h->loss, 8, hipMemcpyDeviceToHost));
    }
    return RBMD_OK;
}

}  // extern "C"
