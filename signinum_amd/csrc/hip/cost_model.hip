// Code-block decode-cost model: a 5-32-32-1 ReLU MLP over code-block header
// features, trained on decode times that the lane kernel records, used to
// order code-blocks so that the 64 lanes of a wave carry similar work.
//
// Hand-written training and inference kernels for gfx950.  Nothing here is
// large enough for MFMA (1281 parameters, tens of thousands of samples);
// every reduction runs in a fixed order, so gradients are deterministic.
//
// Parameter layout (floats): W1[32][5] | b1[32] | W2[32][32] | b2[32] |
// W3[32] | b3.  Loss: mean over samples of 0.5 * (prediction - target)^2.
#include <hip/hip_runtime.h>

#include "../host/plan_types.h"

namespace sgn {

constexpr int kMlpIn = 5, kMlpH = 32;
constexpr int kW1 = 0, kB1 = kW1 + kMlpH * kMlpIn, kW2 = kB1 + kMlpH, kB2 = kW2 + kMlpH * kMlpH,
              kW3 = kB2 + kMlpH, kB3 = kW3 + kMlpH, kMlpParams = kB3 + 1;
static_assert(kMlpParams == 1281, "MLP parameter count");

constexpr int kGradWg = 64;                                   // samples per work-group
constexpr int kRec = kMlpIn + 4 * kMlpH + 1;                  // per-sample record in LDS

// Forward pass of one sample; h1/h2 are the post-ReLU activations.
__device__ __forceinline__ float mlp_forward(const float *p, const float *x, float *h1, float *h2) {
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) {
        float a = p[kB1 + i];
#pragma unroll
        for (int j = 0; j < kMlpIn; ++j) a = fmaf(p[kW1 + i * kMlpIn + j], x[j], a);
        h1[i] = a > 0.0f ? a : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) {
        float a = p[kB2 + i];
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) a = fmaf(p[kW2 + i * kMlpH + j], h1[j], a);
        h2[i] = a > 0.0f ? a : 0.0f;
    }
    float y = p[kB3];
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) y = fmaf(p[kW3 + j], h2[j], y);
    return y;
}

// Per-work-group gradient sums: each thread runs forward and backward for
// one sample and leaves (x, h1, d1, h2, d2, e) in LDS; then every thread
// sums a strided set of parameters over the 64 samples in sample order.
extern "C" __global__ void __launch_bounds__(kGradWg)
sgn_mlp_grad_partial(const float *__restrict__ params, const float *__restrict__ X,
                     const float *__restrict__ Y, uint32_t n, float *__restrict__ partial) {
    __shared__ float s_p[kMlpParams];
    __shared__ float s_rec[kGradWg][kRec + 1];  // +1: odd stride across samples
    for (int k = threadIdx.x; k < kMlpParams; k += kGradWg) s_p[k] = params[k];
    __syncthreads();
    const uint32_t smp = blockIdx.x * kGradWg + threadIdx.x;
    float *r = s_rec[threadIdx.x];
    float x[kMlpIn], h1[kMlpH], h2[kMlpH];
    float e = 0.0f;
    if (smp < n) {
#pragma unroll
        for (int j = 0; j < kMlpIn; ++j) x[j] = X[smp * kMlpIn + j];
        e = mlp_forward(s_p, x, h1, h2) - Y[smp];
    } else {
#pragma unroll
        for (int j = 0; j < kMlpIn; ++j) x[j] = 0.0f;
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) h1[j] = h2[j] = 0.0f;
    }
    // backward: d2 = e * W3 (where h2 > 0), d1 = W2^T d2 (where h1 > 0)
    float d2[kMlpH];
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) d2[i] = h2[i] > 0.0f ? e * s_p[kW3 + i] : 0.0f;
#pragma unroll
    for (int j = 0; j < kMlpIn; ++j) r[j] = x[j];
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) {
        float a = 0.0f;
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) a = fmaf(s_p[kW2 + k * kMlpH + i], d2[k], a);
        r[kMlpIn + i] = h1[i];
        r[kMlpIn + kMlpH + i] = h1[i] > 0.0f ? a : 0.0f;
        r[kMlpIn + 2 * kMlpH + i] = h2[i];
        r[kMlpIn + 3 * kMlpH + i] = d2[i];
    }
    r[kMlpIn + 4 * kMlpH] = e;
    __syncthreads();
    constexpr int oX = 0, oH1 = kMlpIn, oD1 = kMlpIn + kMlpH, oH2 = kMlpIn + 2 * kMlpH,
                  oD2 = kMlpIn + 3 * kMlpH, oE = kMlpIn + 4 * kMlpH;
    for (int k = threadIdx.x; k < kMlpParams; k += kGradWg) {
        int ga, gb;  // record offsets of the two factors (gb < 0: factor 1)
        if (k < kB1) {
            ga = oD1 + k / kMlpIn;
            gb = oX + k % kMlpIn;
        } else if (k < kW2) {
            ga = oD1 + (k - kB1);
            gb = -1;
        } else if (k < kB2) {
            ga = oD2 + (k - kW2) / kMlpH;
            gb = oH1 + (k - kW2) % kMlpH;
        } else if (k < kW3) {
            ga = oD2 + (k - kB2);
            gb = -1;
        } else if (k < kB3) {
            ga = oE;
            gb = oH2 + (k - kW3);
        } else {
            ga = oE;
            gb = -1;
        }
        float acc = 0.0f;
        for (int s = 0; s < kGradWg; ++s)
            acc = fmaf(s_rec[s][ga], gb >= 0 ? s_rec[s][gb] : 1.0f, acc);
        partial[(size_t)blockIdx.x * kMlpParams + k] = acc;
    }
}

// grad[k] = scale * sum over work-groups (in order) of partial[w][k].
extern "C" __global__ void sgn_mlp_reduce(const float *__restrict__ partial, uint32_t groups,
                                          float scale, float *__restrict__ grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kMlpParams) return;
    float acc = 0.0f;
    for (uint32_t w = 0; w < groups; ++w) acc += partial[(size_t)w * kMlpParams + k];
    grad[k] = acc * scale;
}

// One Adam step (bias-corrected) on every parameter.
extern "C" __global__ void sgn_mlp_adam(float *__restrict__ params, const float *__restrict__ grad,
                                        float *__restrict__ m, float *__restrict__ v, float lr,
                                        float beta1, float beta2, float eps, float corr1,
                                        float corr2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kMlpParams) return;
    const float g = grad[k];
    const float mk = beta1 * m[k] + (1.0f - beta1) * g;
    const float vk = beta2 * v[k] + (1.0f - beta2) * g * g;
    m[k] = mk;
    v[k] = vk;
    params[k] -= lr * (mk / corr1) / (sqrtf(vk / corr2) + eps);
}

// Predictions for rows of X.
extern "C" __global__ void sgn_mlp_predict(const float *__restrict__ params,
                                           const float *__restrict__ X, uint32_t n,
                                           float *__restrict__ out) {
    __shared__ float s_p[kMlpParams];
    for (int k = threadIdx.x; k < kMlpParams; k += blockDim.x) s_p[k] = params[k];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[kMlpIn], h1[kMlpH], h2[kMlpH];
#pragma unroll
    for (int j = 0; j < kMlpIn; ++j) x[j] = X[i * kMlpIn + j];
    out[i] = mlp_forward(s_p, x, h1, h2);
}

// Predicted cost of every code-block of a plan, features taken from the
// code-block table itself.  Blocks that decode to zero cost nothing.
extern "C" __global__ void sgn_mlp_predict_blocks(const float *__restrict__ params,
                                                  const CbJob *__restrict__ jobs, uint32_t n,
                                                  float *__restrict__ out) {
    __shared__ float s_p[kMlpParams];
    for (int k = threadIdx.x; k < kMlpParams; k += blockDim.x) s_p[k] = params[k];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const CbJob j = jobs[i];
    if (j.flags & kCbZero) {
        out[i] = -1e30f;
        return;
    }
    const CbFeatures f = cb_features(j);
    float h1[kMlpH], h2[kMlpH];
    out[i] = mlp_forward(s_p, f.f, h1, h2);
}

}  // namespace sgn
