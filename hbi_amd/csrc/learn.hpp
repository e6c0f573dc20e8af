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

// ---- step features (shared by the single-run and the ensemble RK graphs)
// histogram of log10 |V| over the elements after an accepted step
constexpr int kBins = 1024;
constexpr double kBinLo = -20.0, kBinHi = 4.0;

__device__ __forceinline__ int vbin(double absv)
{
    const double lv = absv > 0.0 ? log10(absv) : kBinLo;
    return min(kBins - 1, max(0, (int)((lv - kBinLo) * (kBins / (kBinHi - kBinLo)))));
}

// memory of the previous accepted step
struct StepMem {
    double prev_q[3], prev_h, prev_e;
    int have_prev;
};

// Block of kHid threads; bins: the histogram of n elements.  Thread 0 fills
// q (max, median, p99 of log10 |V|, bin centres) and the 12 features x:
// q, their changes since the previous step, log10 h, log10 errmax, the
// rejection flag, log10 h and errmax of the previous step, and the
// error-optimal step log10(h errmax^(-1/5)).  part: kHid + 1 ints of LDS,
// hb: kBins ints of LDS.
__device__ __forceinline__ void step_features(const int *bins, int n, double hdid, double errmax, bool rejected,
                                              const StepMem &m, int *part, int *hb, float *x, double *q)
{
    const int j = threadIdx.x;
    constexpr int per = kBins / kHid;
    int cnt = 0;
    for (int b = 0; b < per; b++) {
        const int k = bins[b * kHid + j];      // coalesced; hb in bin order
        hb[b * kHid + j] = k;
    }
    __syncthreads();
    for (int b = 0; b < per; b++) cnt += hb[j * per + b];
    part[j + 1] = cnt;
    __syncthreads();
    if (j == 0) {
        part[0] = 0;
        for (int t = 0; t < kHid; t++) part[t + 1] += part[t];   // part[t] = elements in chunks < t
        const double w = (kBinHi - kBinLo) / kBins;
        // smallest bin whose cumulative count reaches target (1-based rank)
        auto rank_bin = [&](long long target) {
            int t = 0;
            while (t < kHid - 1 && part[t + 1] < target) t++;
            long long cum = part[t];
            for (int b = t * per; b < (t + 1) * per; b++) {
                cum += hb[b];
                if (cum >= target) return b;
            }
            return (t + 1) * per - 1;
        };
        const long long n50 = (n + 1) / 2, n99 = (long long)ceil(0.99 * n);
        const int b50 = rank_bin(n50), b99 = rank_bin(n99), bmax = rank_bin(n);
        q[0] = kBinLo + (bmax + 0.5) * w;
        q[1] = kBinLo + (b50 + 0.5) * w;
        q[2] = kBinLo + (b99 + 0.5) * w;
        const double lh = log10(hdid), le = log10(errmax);
        const double ph = m.have_prev ? log10(m.prev_h) : lh, pe = m.have_prev ? log10(m.prev_e) : le;
        for (int k = 0; k < 3; k++) {
            x[k] = (float)q[k];
            x[3 + k] = m.have_prev ? (float)(q[k] - m.prev_q[k]) : 0.0f;
        }
        x[6] = (float)lh;
        x[7] = (float)le;
        x[8] = rejected ? 1.0f : 0.0f;
        x[9] = (float)ph;
        x[10] = (float)pe;
        x[11] = (float)(lh - 0.2 * le);
    }
    __syncthreads();
}

// the learned proposal, within [lo, hi] x the rkqs proposal h_pi and dtmax
__device__ __forceinline__ double learned_h(float x11, float out, double h_pi, double safety, double lo, double hi,
                                            double dtmax)
{
    double hn = safety * pow(10.0, (double)x11 + (double)out);
    if (!(hn > 0.0) || !isfinite(hn)) hn = h_pi;
    hn = fmin(fmax(hn, lo * h_pi), hi * h_pi);
    return fmin(hn, dtmax);
}

}  // namespace hbi

struct hbi_mlp {
    hipStream_t stream = nullptr;
    hbi::DevBuf<float> p, m1, m2, mu, sd;   // parameters, Adam moments, normalisation
    hbi::DevBuf<uint16_t> w;                // BF16 copy of p
    long long adam_t = 0;
    hbi::MlpDev dev() const { return hbi::MlpDev{w.p, p.p, mu.p, sd.p}; }
};
