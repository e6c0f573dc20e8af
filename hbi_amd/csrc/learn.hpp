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

// The weights of hidden unit j (thread j of a kHid-thread workgroup) in
// registers: W2 row j as 16-byte groups of eight BF16 values, W1 row j as
// 8-byte groups of four, the biases and W3[j].  Loaded together at the top of
// a kernel, their latency overlaps the work before the inference.
struct MlpRow {
    uint4 w2[kHid / 8];
    uint2 w1[kIn / 4];
    float b1, b2, w3, b3;
};
static_assert(kOffW2 % 8 == 0 && kIn % 4 == 0 && kHid % 8 == 0, "MLP row loads need 16- and 8-byte alignment");

__device__ __forceinline__ MlpRow mlp_load_row(const MlpDev &M)
{
    const int j = threadIdx.x;
    MlpRow R;
    const uint4 *w2 = (const uint4 *)(M.w + kOffW2 + (size_t)j * kHid);
#pragma unroll
    for (int q = 0; q < kHid / 8; q++) R.w2[q] = w2[q];
    const uint2 *w1 = (const uint2 *)(M.w + kOffW1 + (size_t)j * kIn);
#pragma unroll
    for (int q = 0; q < kIn / 4; q++) R.w1[q] = w1[q];
    R.b1 = M.p[kOffB1 + j];
    R.b2 = M.p[kOffB2 + j];
    R.w3 = bf16_to_f32(M.w[kOffW3 + j]);
    R.b3 = M.p[kOffB3];
    return R;
}

// BF16 value k of a group of packed pairs (low half first)
__device__ __forceinline__ float bf_at(uint32_t pair, int hi) { return __uint_as_float(hi ? pair & 0xffff0000u : pair << 16); }

// One sample on one workgroup of kHid threads.  x: kIn raw features (LDS),
// h1, red: kHid floats of LDS scratch.  Every thread returns the output.
// Layer sums run in the order of the parameter vector (b + w_0 a_0 + w_1 a_1 ...).
__device__ __forceinline__ float mlp_infer_block(const MlpDev &M, const MlpRow &R, const float *x, float *h1, float *red)
{
    const int j = threadIdx.x;
    float z = R.b1;
#pragma unroll
    for (int i = 0; i < kIn; i++) {
        const uint2 g = R.w1[i / 4];
        const uint32_t pr = (i % 4) < 2 ? g.x : g.y;
        z += bf_at(pr, i % 2) * ((x[i] - M.mu[i]) / M.sd[i]);
    }
    h1[j] = gelu(z);
    __syncthreads();
    float z2 = R.b2;
#pragma unroll
    for (int k = 0; k < kHid; k++) {
        const uint4 g = R.w2[k / 8];
        const int e = k % 8;
        const uint32_t pr = e < 2 ? g.x : e < 4 ? g.y : e < 6 ? g.z : g.w;
        z2 += bf_at(pr, e % 2) * h1[k];
    }
    red[j] = R.w3 * gelu(z2);
    __syncthreads();
    for (int s = kHid / 2; s > 0; s >>= 1) {
        if (j < s) red[j] += red[j + s];
        __syncthreads();
    }
    return red[0] + R.b3;
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

// Block of kHid threads; bins: the histogram of n elements.  Fills q (max,
// median, p99 of log10 |V|, bin centres) and the 12 features x: q, their
// changes since the previous step, log10 h, log10 errmax, the rejection flag,
// log10 h and errmax of the previous step, and the error-optimal step
// log10(h errmax^(-1/5)).  Thread j owns bins per j .. per (j + 1) - 1: the
// counts of the chunks are scanned across the workgroup (shuffles within each
// wavefront, then the first wavefront's total added to the second's), and
// the one thread whose chunk holds a rank walks its bins to it.  part: kHid +
// 1 ints of LDS, hb: kBins ints of LDS.
static_assert(kHid == 128, "the chunk scan spans two wavefronts");
__device__ __forceinline__ void step_features(const int *bins, int n, double hdid, double errmax, bool rejected,
                                              const StepMem &m, int *part, int *hb, float *x, double *q)
{
    const int j = threadIdx.x, lane = j % 64;
    constexpr int per = kBins / kHid;
    constexpr double w = (kBinHi - kBinLo) / kBins;
    for (int b = 0; b < per; b++) hb[b * kHid + j] = bins[b * kHid + j];   // coalesced; hb in bin order
    if (j < 3) q[j] = kBinLo + (kBins - 0.5) * w;   // a rank past the histogram's total: the last bin
    __syncthreads();
    int s = 0;
    for (int b = 0; b < per; b++) s += hb[j * per + b];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(s, o, 64);
        if (lane >= o) s += t;
    }
    if (j == 63) part[0] = s;
    __syncthreads();
    if (j >= 64) s += part[0];
    __syncthreads();
    part[j + 1] = s;   // elements in chunks <= j
    if (j == 0) part[0] = 0;
    __syncthreads();
    // the smallest bin whose cumulative count reaches each 1-based rank
    const long long rank[3] = {(long long)n, (long long)(n + 1) / 2, (long long)ceil(0.99 * n)};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const long long target = rank[k];
        if (part[j] < target && target <= part[j + 1]) {
            long long cum = part[j];
            int bin = (j + 1) * per - 1;
            for (int b = j * per; b < (j + 1) * per; b++) {
                cum += hb[b];
                if (cum >= target) {
                    bin = b;
                    break;
                }
            }
            q[k] = kBinLo + (bin + 0.5) * w;
        }
    }
    __syncthreads();
    if (j == 0) {
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
