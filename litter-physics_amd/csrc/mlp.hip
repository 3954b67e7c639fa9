// Forward, backward and training kernels of the return-map MLP surrogate, and
// the mpm_mlp_* entry points of the C ABI.
//
// Forward and backward run on bf16 MFMA (v_mfma_f32_16x16x16_bf16) with f32
// accumulation, one wavefront per 16-sample tile.  Training keeps f32 master
// weights and Adam moments on the device; after every update the weights are
// repacked into bf16 fragments.  The gradient of a batch is the fixed-order sum
// of per-workgroup partials, so it is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpm_engine.h"
#include "engine_internal.h"
#include "mlp.h"

using namespace mpm;

namespace {

constexpr int kGradWgs = 512;  // workgroups (one wavefront each) of the gradient kernel
// Largest batch of mpm_mlp_loss_grad_noalloc: sample indices stay in int.
constexpr size_t kNoAllocMaxSamples = (size_t)INT32_MAX / 16;

__global__ void mlp_pack_kernel(const float* __restrict__ p, uint16_t* fwd, uint16_t* bwd, float* bias) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = kFwdFrags * kFragElems, nb = kBwdFrags * kFragElems;
    if (i < nf)
        fwd[i] = mlp_fwd_elem(p, i);
    else if (i < nf + nb)
        bwd[i - nf] = mlp_bwd_elem(p, i - nf);
    else if (i < nf + nb + kBiasPacked)
        bias[i - nf - nb] = mlp_bias_elem(p, i - nf - nb);
}

// Layer-1 B operand of sample tile st: features 4 (lane/16) .. + 3 of sample lane % 16.
__device__ inline f32x4 load_inputs(const float* X, const int* idx, int n, int st, int lane) {
    const int s = st * 16 + (lane & 15), k0 = 4 * (lane >> 4);
    f32x4 x = {0.0f, 0.0f, 0.0f, 0.0f};
    if (s < n) {
        const size_t src = (size_t)(idx ? idx[s] : s);
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (k0 + e < kMlpIn) x[e] = X[src * kMlpIn + k0 + e];
    }
    return x;
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const uint16_t* __restrict__ fwd,
                                                          const float* __restrict__ bias, const float* __restrict__ X,
                                                          float* __restrict__ Y, int n) {
    __shared__ uint16_t sf[kFwdFrags * kFragElems];
    __shared__ float sb[kBiasPacked];
    for (int i = threadIdx.x; i < kFwdFrags * kFragElems / 8; i += blockDim.x)
        reinterpret_cast<uint4*>(sf)[i] = reinterpret_cast<const uint4*>(fwd)[i];
    for (int i = threadIdx.x; i < kBiasPacked; i += blockDim.x) sb[i] = bias[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntiles = (n + 15) / 16;
    // The tile index is uniform across the wavefront, so every MFMA runs with all lanes.
    for (int st = blockIdx.x * 4 + wave; st < ntiles; st += gridDim.x * 4) {
        MlpTile t;
        mlp_tile_forward<false>(sf, sb, pack4(load_inputs(X, nullptr, n, st, lane)), lane, t);
        const int s = st * 16 + (lane & 15), u0 = 4 * (lane >> 4);
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (s < n && u0 + e < kMlpOut) Y[(size_t)s * kMlpOut + u0 + e] = t.y[e];
    }
}

// 16 x 16 tile transpose through LDS: v in D layout (row 4 (lane/16) + e, column
// lane % 16) comes back as row lane % 16, columns 4 (lane/16) + e.
__device__ inline f32x4 transpose16(float (*tile)[17], f32x4 v, int lane) {
    const int col = lane & 15, grp = lane >> 4;
#pragma unroll
    for (int e = 0; e < 4; ++e) tile[4 * grp + e][col] = v[e];
    __syncthreads();
    f32x4 r;
#pragma unroll
    for (int e = 0; e < 4; ++e) r[e] = tile[col][4 * grp + e];
    __syncthreads();
    return r;
}

__device__ inline f32x4 unpack4(bf16x4 b) {
    return f32x4{bf16_value((uint16_t)b[0]), bf16_value((uint16_t)b[1]), bf16_value((uint16_t)b[2]),
                 bf16_value((uint16_t)b[3])};
}

// Sum over the four lanes sharing lane % 16.
__device__ inline float sum_groups(float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// Loss sum_s sum_u w_u (y_su - t_su)^2 / n and its gradient over a batch, one
// wavefront per workgroup, per-workgroup partials in the master layout.
__global__ __launch_bounds__(64) void mlp_grad_kernel(const uint16_t* __restrict__ fwd,
                                                      const uint16_t* __restrict__ bwd,
                                                      const float* __restrict__ bias, const float* __restrict__ X,
                                                      const float* __restrict__ T, const int* __restrict__ idx, int n,
                                                      float inv_n, const float* __restrict__ lw,
                                                      float* __restrict__ gpart, float* __restrict__ lpart) {
    __shared__ float acc[kMlpParams];
    __shared__ float tile[16][17];
    const int lane = threadIdx.x, col = lane & 15, grp = lane >> 4;
    for (int i = lane; i < kMlpParams; i += 64) acc[i] = 0.0f;
    __syncthreads();
    float loss = 0.0f;
    const int ntiles = (n + 15) / 16;
    for (int st = blockIdx.x; st < ntiles; st += gridDim.x) {
        const f32x4 xin = load_inputs(X, idx, n, st, lane);
        MlpTile t;
        mlp_tile_forward<true>(fwd, bias, pack4(xin), lane, t);
        // Output error (D layout: unit 4 grp + e of sample col).
        const int s = st * 16 + col;
        f32x4 dy = {0.0f, 0.0f, 0.0f, 0.0f};
        if (s < n) {
            const size_t src = (size_t)(idx ? idx[s] : s);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int u = 4 * grp + e;
                if (u < kMlpOut) {
                    float r = t.y[e] - T[src * kMlpOut + u];
                    const float wu = lw ? lw[u] : 1.0f;
                    loss += wu * r * r;
                    dy[e] = 2.0f * wu * r * inv_n;
                }
            }
        }
        // Deltas of the hidden layers: dZ_l = (W_{l+1} dZ_{l+1}) * silu'(Z_l).
        f32x4 d[4][4];
        d[3][0] = dy;
        const bf16x4 dyb = pack4(dy);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            f32x4 a = {0.0f, 0.0f, 0.0f, 0.0f};
            a = mfma(frag(bwd, kBwdFrag4 + rt, lane), dyb, a);
#pragma unroll
            for (int e = 0; e < 4; ++e) a[e] *= silu_grad(t.z[2][rt][e]);
            d[2][rt] = a;
        }
#pragma unroll
        for (int l = 1; l >= 0; --l) {
            const int fb = l == 1 ? kBwdFrag3 : kBwdFrag2;
            bf16x4 nb[4];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) nb[ct] = pack4(d[l + 1][ct]);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) {
                f32x4 a = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) a = mfma(frag(bwd, fb + rt * 4 + ct, lane), nb[ct], a);
#pragma unroll
                for (int e = 0; e < 4; ++e) a[e] *= silu_grad(t.z[l][rt][e]);
                d[l][rt] = a;
            }
        }
        // Weight gradients dW_l^T = sum_s dZ_l^T[:, s] H_{l-1}[s, :] (A = dZ^T, B = H),
        // bias gradients sum_s dZ_l[s, :].
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const int wbase = l == 0 ? kMlpW1 : (l == 1 ? kMlpW2 : (l == 2 ? kMlpW3 : kMlpW4));
            const int bbase = l == 0 ? kMlpB1 : (l == 1 ? kMlpB2 : (l == 2 ? kMlpB3 : kMlpB4));
            const int nin = l == 0 ? kMlpIn : kMlpHidden, nout = l == 3 ? kMlpOut : kMlpHidden;
            const int ntn = l == 3 ? 1 : 4, ntk = l == 0 ? 1 : 4;
            bf16x4 hb[4];
            if (l == 0) {
                // B operand X[s][k]: samples 4 grp + e, feature col.
                f32x4 xb = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int ss = st * 16 + 4 * grp + e;
                    if (ss < n && col < kMlpIn) xb[e] = X[(size_t)(idx ? idx[ss] : ss) * kMlpIn + col];
                }
                hb[0] = pack4(xb);
            } else {
                for (int kt = 0; kt < 4; ++kt) hb[kt] = pack4(transpose16(tile, unpack4(t.h[l - 1][kt]), lane));
            }
            for (int nt = 0; nt < ntn; ++nt) {
                const f32x4 dt = transpose16(tile, d[l][nt], lane);  // dZ^T row col, samples 4 grp + e
                const float bsum = sum_groups(dt[0] + dt[1] + dt[2] + dt[3]);
                if (grp == 0 && nt * 16 + col < nout) acc[bbase + nt * 16 + col] += bsum;
                const bf16x4 ab = pack4(dt);
                for (int kt = 0; kt < ntk; ++kt) {
                    f32x4 g = {0.0f, 0.0f, 0.0f, 0.0f};
                    g = mfma(ab, hb[kt], g);
                    const int k = kt * 16 + col;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int un = nt * 16 + 4 * grp + e;
                        if (k < nin && un < nout) acc[wbase + k * nout + un] += g[e];
                    }
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    __syncthreads();
    for (int i = lane; i < kMlpParams; i += 64) gpart[(size_t)blockIdx.x * kMlpParams + i] = acc[i];
    if (lane == 0) lpart[blockIdx.x] = loss;
}

// Fixed-order sum of the partial gradients, then (lr > 0) one Adam step.
__global__ void mlp_adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                const float* __restrict__ gpart, int nparts, float lr, float bc1, float bc2,
                                float* __restrict__ gout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    float g = 0.0f;
    for (int w = 0; w < nparts; ++w) g += gpart[(size_t)w * kMlpParams + i];
    if (gout) gout[i] = g;
    if (lr > 0.0f) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        float mi = b1 * m[i] + (1.0f - b1) * g;
        float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

// Loss of a batch from its per-workgroup partials, in a fixed order: sum / n.
__global__ void mlp_loss_sum_kernel(const float* __restrict__ lpart, int nparts, double inv_n, double* loss) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double l = 0.0;
    for (int w = 0; w < nparts; ++w) l += lpart[w];
    *loss = l * inv_n;
}

// Scratch of mpm_mlp_loss_grad_noalloc: kNoAllocWgs partial gradients of
// kMlpParams floats, then kNoAllocWgs partial losses (2.46 MB).  Each GPU
// holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns on that GPU (scratch_of).
constexpr int kNoAllocWgs = 64;
constexpr size_t kNoAllocFloats = (size_t)kNoAllocWgs * kMlpParams + kNoAllocWgs;
static __device__ __align__(256) float scratch[kNoAllocFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// resolved once per device.
hipError_t scratch_of(int device, float** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<float*>(ptr);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

uint64_t splitmix(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double unit(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / (double)(1ull << 53)); }

}  // namespace

struct mpm_mlp {
    float* p = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* fwd = nullptr;
    uint16_t* bwd = nullptr;
    float* bias = nullptr;
    float* gpart = nullptr;
    float* lpart = nullptr;
    float* gsum = nullptr;
    float* lw = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    float* noalloc_scratch = nullptr;  // this device's `scratch` (scratch_of)
    uint64_t steps = 0;
    std::string err;

    int dev_fail(const char* what, hipError_t e) {
        err = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
        return MPM_ERR_DEVICE;
    }
    ~mpm_mlp() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* q : {(void*)p, (void*)m, (void*)v, (void*)fwd, (void*)bwd, (void*)bias, (void*)gpart, (void*)lpart,
                        (void*)gsum, (void*)lw})
            (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

#define MLP_OK(expr)                                            \
    do {                                                        \
        hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) return mlp->dev_fail(#expr, e_);  \
    } while (0)

namespace {

int pack(mpm_mlp* mlp) {
    const int total = (kFwdFrags + kBwdFrags) * kFragElems + kBiasPacked;
    mlp_pack_kernel<<<(total + 255) / 256, 256, 0, mlp->stream>>>(mlp->p, mlp->fwd, mlp->bwd, mlp->bias);
    MLP_OK(hipGetLastError());
    return MPM_OK;
}

template <class T>
struct Dev {
    T* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)); }
    ~Dev() { (void)hipFree(p); }
};

int grad_batch(mpm_mlp* mlp, const float* X, const float* T, const int* idx, int n, float lr, float* gout) {
    const int tiles = (n + 15) / 16;
    const int wgs = std::max(1, std::min(kGradWgs, tiles));
    mlp_grad_kernel<<<wgs, 64, 0, mlp->stream>>>(mlp->fwd, mlp->bwd, mlp->bias, X, T, idx, n, 1.0f / (float)n, mlp->lw,
                                                 mlp->gpart, mlp->lpart);
    float bc1 = 1.0f, bc2 = 1.0f;
    if (lr > 0.0f) {
        mlp->steps += 1;
        bc1 = 1.0f - std::pow(0.9f, (float)mlp->steps);
        bc2 = 1.0f - std::pow(0.999f, (float)mlp->steps);
    }
    mlp_adam_kernel<<<(kMlpParams + 255) / 256, 256, 0, mlp->stream>>>(mlp->p, mlp->m, mlp->v, mlp->gpart, wgs, lr,
                                                                       bc1, bc2, gout);
    MLP_OK(hipGetLastError());
    if (lr > 0.0f) return pack(mlp);
    return MPM_OK;
}

int set_loss_weights(mpm_mlp* mlp, const float* w7) {
    float w[kMlpOut];
    for (int u = 0; u < kMlpOut; ++u) w[u] = w7 ? w7[u] : 1.0f;
    MLP_OK(hipMemcpyAsync(mlp->lw, w, sizeof(w), hipMemcpyHostToDevice, mlp->stream));
    MLP_OK(hipStreamSynchronize(mlp->stream));
    return MPM_OK;
}

}  // namespace

namespace mpm {
const uint16_t* mlp_device_fragments(const mpm_mlp* mlp) { return mlp ? mlp->fwd : nullptr; }
const float* mlp_device_biases(const mpm_mlp* mlp) { return mlp ? mlp->bias : nullptr; }
}  // namespace mpm

extern "C" {

size_t mpm_mlp_param_count(void) { return (size_t)kMlpParams; }

int mpm_mlp_create(const float* params, uint64_t seed, mpm_mlp** out) {
    if (!out) return MPM_ERR_INVALID;
    *out = nullptr;
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) return MPM_ERR_DEVICE;
    mpm_mlp* mlp = new mpm_mlp();
    auto fail = [&](hipError_t e) {
        delete mlp;
        return e == hipSuccess ? MPM_OK : MPM_ERR_DEVICE;
    };
    hipError_t e = hipGetDevice(&mlp->device);
    // The no-allocation gradient's scratch on this device, resolved at set-up.
    if (e == hipSuccess) e = scratch_of(mlp->device, &mlp->noalloc_scratch);
    if (e == hipSuccess) e = hipStreamCreate(&mlp->stream);
    if (e == hipSuccess) e = hipMalloc(&mlp->p, sizeof(float) * kMlpParams);
    if (e == hipSuccess) e = hipMalloc(&mlp->m, sizeof(float) * kMlpParams);
    if (e == hipSuccess) e = hipMalloc(&mlp->v, sizeof(float) * kMlpParams);
    if (e == hipSuccess) e = hipMalloc(&mlp->fwd, sizeof(uint16_t) * kFwdFrags * kFragElems);
    if (e == hipSuccess) e = hipMalloc(&mlp->bwd, sizeof(uint16_t) * kBwdFrags * kFragElems);
    if (e == hipSuccess) e = hipMalloc(&mlp->bias, sizeof(float) * kBiasPacked);
    if (e == hipSuccess) e = hipMalloc(&mlp->gpart, sizeof(float) * kMlpParams * kGradWgs);
    if (e == hipSuccess) e = hipMalloc(&mlp->lpart, sizeof(float) * kGradWgs);
    if (e == hipSuccess) e = hipMalloc(&mlp->gsum, sizeof(float) * kMlpParams);
    if (e == hipSuccess) e = hipMalloc(&mlp->lw, sizeof(float) * kMlpOut);
    if (e == hipSuccess) e = hipMemset(mlp->m, 0, sizeof(float) * kMlpParams);
    if (e == hipSuccess) e = hipMemset(mlp->v, 0, sizeof(float) * kMlpParams);
    if (e != hipSuccess) return fail(e);
    std::vector<float> p(kMlpParams, 0.0f);
    if (params) {
        std::memcpy(p.data(), params, sizeof(float) * kMlpParams);
    } else {
        // Xavier-uniform weights from SplitMix64, zero biases.
        uint64_t s = seed;
        const int wb[4] = {kMlpW1, kMlpW2, kMlpW3, kMlpW4}, nin[4] = {12, 64, 64, 64}, nout[4] = {64, 64, 64, 7};
        for (int l = 0; l < 4; ++l) {
            double a = std::sqrt(6.0 / (nin[l] + nout[l]));
            for (int i = 0; i < nin[l] * nout[l]; ++i) p[wb[l] + i] = (float)((2.0 * unit(s) - 1.0) * a);
        }
    }
    e = hipMemcpy(mlp->p, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(e);
    if (pack(mlp) || set_loss_weights(mlp, nullptr) || hipStreamSynchronize(mlp->stream) != hipSuccess) {
        delete mlp;
        return MPM_ERR_DEVICE;
    }
    *out = mlp;
    return MPM_OK;
}

void mpm_mlp_destroy(mpm_mlp* mlp) { delete mlp; }

const char* mpm_mlp_last_error(const mpm_mlp* mlp) { return mlp ? mlp->err.c_str() : "null network"; }

int mpm_mlp_get_params(mpm_mlp* mlp, float* out) {
    if (!mlp || !out) return MPM_ERR_INVALID;
    MLP_OK(hipMemcpyAsync(out, mlp->p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, mlp->stream));
    MLP_OK(hipStreamSynchronize(mlp->stream));
    return MPM_OK;
}

int mpm_mlp_set_params(mpm_mlp* mlp, const float* in) {
    if (!mlp || !in) return MPM_ERR_INVALID;
    MLP_OK(hipMemcpyAsync(mlp->p, in, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, mlp->stream));
    if (int rc = pack(mlp)) return rc;
    MLP_OK(hipStreamSynchronize(mlp->stream));
    return MPM_OK;
}

int mpm_mlp_forward(mpm_mlp* mlp, size_t n, const float* x, float* y) {
    if (!mlp || !x || !y || n > (size_t)INT32_MAX / 16) return MPM_ERR_INVALID;
    if (n == 0) return MPM_OK;
    Dev<float> dx, dy;
    MLP_OK(dx.alloc(n * kMlpIn));
    MLP_OK(dy.alloc(n * kMlpOut));
    MLP_OK(hipMemcpyAsync(dx.p, x, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, mlp->stream));
    const int tiles = (int)((n + 15) / 16);
    mlp_forward_kernel<<<std::min(2048, (tiles + 3) / 4), 256, 0, mlp->stream>>>(mlp->fwd, mlp->bias, dx.p, dy.p,
                                                                                 (int)n);
    MLP_OK(hipGetLastError());
    MLP_OK(hipMemcpyAsync(y, dy.p, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost, mlp->stream));
    MLP_OK(hipStreamSynchronize(mlp->stream));
    return MPM_OK;
}

int mpm_mlp_loss_grad(mpm_mlp* mlp, size_t n, const float* x, const float* t, const float* loss_weights,
                      double* loss, float* grad) {
    if (!mlp || !x || !t || n == 0 || n > (size_t)INT32_MAX / 16) return MPM_ERR_INVALID;
    if (int rc = set_loss_weights(mlp, loss_weights)) return rc;
    Dev<float> dx, dt;
    MLP_OK(dx.alloc(n * kMlpIn));
    MLP_OK(dt.alloc(n * kMlpOut));
    MLP_OK(hipMemcpyAsync(dx.p, x, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, mlp->stream));
    MLP_OK(hipMemcpyAsync(dt.p, t, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice, mlp->stream));
    if (int rc = grad_batch(mlp, dx.p, dt.p, nullptr, (int)n, 0.0f, mlp->gsum)) return rc;
    const int wgs = std::max(1, std::min(kGradWgs, (int)((n + 15) / 16)));
    std::vector<float> lp(wgs);
    MLP_OK(hipMemcpyAsync(lp.data(), mlp->lpart, sizeof(float) * wgs, hipMemcpyDeviceToHost, mlp->stream));
    if (grad) MLP_OK(hipMemcpyAsync(grad, mlp->gsum, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, mlp->stream));
    MLP_OK(hipStreamSynchronize(mlp->stream));
    double l = 0.0;
    for (float v : lp) l += v;
    if (loss) *loss = l / (double)n;
    return MPM_OK;
}

int mpm_mlp_loss_grad_noalloc(mpm_mlp* mlp, size_t n, const float* x, const float* t, const float* loss_weights,
                              double* loss, float* grad) {
    if (!mlp || !x || !t || n == 0) {
        if (mlp) mlp->err = "mpm_mlp_loss_grad_noalloc: null network, inputs or empty batch";
        return MPM_ERR_INVALID;
    }
    if (n > kNoAllocMaxSamples) {
        mlp->err = "mpm_mlp_loss_grad_noalloc: batch of " + std::to_string(n) + " samples exceeds the maximum of " +
                   std::to_string(kNoAllocMaxSamples);
        return MPM_ERR_CAPACITY;
    }
    int dev = -1;
    MLP_OK(hipGetDevice(&dev));
    if (dev != mlp->device || !mlp->noalloc_scratch) {
        mlp->err = "mpm_mlp_loss_grad_noalloc: the network's device is not the current device";
        return MPM_ERR_INVALID;
    }
    float* gpart = mlp->noalloc_scratch;
    float* lpart = gpart + (size_t)kNoAllocWgs * kMlpParams;
    const int tiles = (int)((n + 15) / 16);
    const int wgs = std::max(1, std::min(kNoAllocWgs, tiles));
    mlp_grad_kernel<<<wgs, 64, 0, mlp->stream>>>(mlp->fwd, mlp->bwd, mlp->bias, x, t, nullptr, (int)n,
                                                 1.0f / (float)n, loss_weights, gpart, lpart);
    if (grad)
        mlp_adam_kernel<<<(kMlpParams + 255) / 256, 256, 0, mlp->stream>>>(nullptr, nullptr, nullptr, gpart, wgs, 0.0f,
                                                                           1.0f, 1.0f, grad);
    if (loss) mlp_loss_sum_kernel<<<1, 64, 0, mlp->stream>>>(lpart, wgs, 1.0 / (double)n, loss);
    MLP_OK(hipGetLastError());
    MLP_OK(hipStreamSynchronize(mlp->stream));
    return MPM_OK;
}

int mpm_mlp_train(mpm_mlp* mlp, size_t n, const float* x, const float* t, const float* loss_weights, int epochs,
                  int batch, double lr, uint64_t seed, double* losses) {
    if (!mlp || !x || !t || n == 0 || epochs < 1 || batch < 16 || n > (size_t)INT32_MAX / 16 || !(lr > 0.0))
        return MPM_ERR_INVALID;
    if (int rc = set_loss_weights(mlp, loss_weights)) return rc;
    Dev<float> dx, dt;
    Dev<int> didx;
    MLP_OK(dx.alloc(n * kMlpIn));
    MLP_OK(dt.alloc(n * kMlpOut));
    MLP_OK(didx.alloc(n));
    MLP_OK(hipMemcpyAsync(dx.p, x, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, mlp->stream));
    MLP_OK(hipMemcpyAsync(dt.p, t, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice, mlp->stream));
    std::vector<int> perm(n);
    for (size_t i = 0; i < n; ++i) perm[i] = (int)i;
    uint64_t s = seed;
    const size_t per_epoch = (n + batch - 1) / batch;
    const double total = (double)per_epoch * epochs;
    size_t step = 0;
    std::vector<float> lp(kGradWgs);
    for (int ep = 0; ep < epochs; ++ep) {
        for (size_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[(size_t)(unit(s) * (double)(i + 1)) % (i + 1)]);
        MLP_OK(hipMemcpyAsync(didx.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, mlp->stream));
        double ep_loss = 0.0;
        for (size_t b0 = 0; b0 < n; b0 += batch, ++step) {
            const int nb = (int)std::min((size_t)batch, n - b0);
            // Cosine decay from lr to lr / 100.
            const double frac = (double)step / total;
            const float rate = (float)(lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(M_PI * frac))));
            if (int rc = grad_batch(mlp, dx.p, dt.p, didx.p + b0, nb, rate, nullptr)) return rc;
            if (losses) {
                const int wgs = std::max(1, std::min(kGradWgs, (nb + 15) / 16));
                MLP_OK(hipMemcpyAsync(lp.data(), mlp->lpart, sizeof(float) * wgs, hipMemcpyDeviceToHost, mlp->stream));
                MLP_OK(hipStreamSynchronize(mlp->stream));
                for (int w = 0; w < wgs; ++w) ep_loss += lp[w];
            }
        }
        // The index buffer is rewritten next epoch: wait for this epoch's batches.
        MLP_OK(hipStreamSynchronize(mlp->stream));
        if (losses) losses[ep] = ep_loss / (double)n;
    }
    return MPM_OK;
}

}  // extern "C"
