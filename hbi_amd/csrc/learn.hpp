// learn.hpp -- the step-size MLP shared by its training kernels (learn.hip)
// and the RK graph (rk.hip): 12 -> 128 -> 128 -> 1, GELU, BF16 weights with
// FP32 biases, activations and accumulation.
#pragma once
#include "common.hpp"

namespace hbi {

constexpr int kIn = 12, kHid = 128;
// parameter vector (FP32 master copy): W1 (kHid x kIn, row-major), b1, W2
// (kHid x kHid, row-major), b2, W3 (kHid), b3
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid, kNParam = kOffB3 + 1;

// device view used for inference
struct MlpDev {
    const uint16_t *w;    // BF16 copy of the parameter vector (weights used from it)
    const float *p;       // FP32 parameters (biases used from it)
    const float *mu, *sd; // feature normalisation: (x - mu) / sd
};

// exact GELU, x Phi(x)
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// One sample on one workgroup of kHid threads.  x: kIn raw features (LDS),
// h1, red: kHid floats of LDS scratch.  Every thread returns the output.
__device__ __forceinline__ float mlp_infer_block(const MlpDev &M, const float *x, float *h1, float *red)
{
    const int j = threadIdx.x;
    float z = M.p[kOffB1 + j];
    for (int i = 0; i < kIn; i++) z += bf16_to_f32(M.w[kOffW1 + j * kIn + i]) * ((x[i] - M.mu[i]) / M.sd[i]);
    h1[j] = gelu(z);
    __syncthreads();
    float z2 = M.p[kOffB2 + j];
    for (int k = 0; k < kHid; k++) z2 += bf16_to_f32(M.w[kOffW2 + j * kHid + k]) * h1[k];
    red[j] = bf16_to_f32(M.w[kOffW3 + j]) * gelu(z2);
    __syncthreads();
    for (int s = kHid / 2; s > 0; s >>= 1) {
        if (j < s) red[j] += red[j + s];
        __syncthreads();
    }
    return red[0] + M.p[kOffB3];
}

}  // namespace hbi

struct hbi_mlp {
    hipStream_t stream = nullptr;
    hbi::DevBuf<float> p, m1, m2, mu, sd;   // parameters, Adam moments, normalisation
    hbi::DevBuf<uint16_t> w;                // BF16 copy of p
    long long adam_t = 0;
    hbi::MlpDev dev() const { return hbi::MlpDev{w.p, p.p, mu.p, sd.p}; }
};
