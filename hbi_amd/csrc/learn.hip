// learn.hip -- training and evaluation of the step-size MLP on the device.
//
// Batches are processed in groups of kSB samples per workgroup (kHid
// threads, thread j = hidden unit j).  One kernel runs the forward pass and,
// for training, the backward pass, writing each workgroup's gradient into
// its own slot; the slots are summed in a fixed order, then Adam updates the
// FP32 master parameters and refreshes their BF16 copy.  Forward and
// backward use the BF16 weights (the gradient is taken with respect to them
// and applied to the master copy).  Loss: mean squared error.
#include "learn.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

namespace hbi {

constexpr int kSB = 16;

__global__ __launch_bounds__(kHid) void mlp_fb_kernel(const float *__restrict__ X, const float *__restrict__ T,
                                                      const int *__restrict__ perm, int off, int nb, MlpDev M,
                                                      float *__restrict__ Y, float *__restrict__ losspart,
                                                      float *__restrict__ gpart)
{
    __shared__ float xs[kSB][kIn];
    __shared__ float z1[kSB][kHid], a1[kSB][kHid], z2[kSB][kHid], a2[kSB][kHid], red[kSB][kHid];
    __shared__ float err[kSB];
    const int j = threadIdx.x;
    const int s0 = blockIdx.x * kSB;
    const int ns = min(kSB, nb - s0);
    auto sample = [&](int s) { return perm ? perm[off + s0 + s] : off + s0 + s; };
    for (int q = j; q < kSB * kIn; q += kHid) {
        const int s = q / kIn, i = q % kIn;
        xs[s][i] = s < ns ? (X[(size_t)sample(s) * kIn + i] - M.mu[i]) / M.sd[i] : 0.0f;
    }
    __syncthreads();
    {
        float w1[kIn];
#pragma unroll
        for (int i = 0; i < kIn; i++) w1[i] = bf16_to_f32(M.w[kOffW1 + j * kIn + i]);
        const float b = M.p[kOffB1 + j];
#pragma unroll
        for (int s = 0; s < kSB; s++) {
            float z = b;
#pragma unroll
            for (int i = 0; i < kIn; i++) z += w1[i] * xs[s][i];
            z1[s][j] = z;
            a1[s][j] = gelu(z);
        }
    }
    __syncthreads();
    {
        float acc[kSB];
        const float b = M.p[kOffB2 + j];
#pragma unroll
        for (int s = 0; s < kSB; s++) acc[s] = b;
        for (int k = 0; k < kHid; k++) {
            const float w = bf16_to_f32(M.w[kOffW2 + j * kHid + k]);
#pragma unroll
            for (int s = 0; s < kSB; s++) acc[s] += w * a1[s][k];
        }
#pragma unroll
        for (int s = 0; s < kSB; s++) {
            z2[s][j] = acc[s];
            a2[s][j] = gelu(acc[s]);
        }
    }
    const float w3 = bf16_to_f32(M.w[kOffW3 + j]);
#pragma unroll
    for (int s = 0; s < kSB; s++) red[s][j] = w3 * a2[s][j];
    __syncthreads();
    for (int h = kHid / 2; h > 0; h >>= 1) {
        if (j < h)
#pragma unroll
            for (int s = 0; s < kSB; s++) red[s][j] += red[s][j + h];
        __syncthreads();
    }
    if (j < kSB) {
        const float o = red[j][0] + M.p[kOffB3];
        float e = 0.0f;
        if (j < ns) {
            if (Y) Y[sample(j)] = o;
            if (T) e = o - T[sample(j)];
        }
        err[j] = e;
    }
    __syncthreads();
    if (losspart && j == 0) {
        float L = 0.0f;
        for (int s = 0; s < ns; s++) L += err[s] * err[s];
        losspart[blockIdx.x] = L;
    }
    if (!gpart) return;
    // d loss / d output = 2 e / nb
    float go[kSB];
#pragma unroll
    for (int s = 0; s < kSB; s++) go[s] = 2.0f * err[s] / (float)nb;
    float *G = gpart + (size_t)blockIdx.x * kNParam;
    {
        float g = 0.0f;
#pragma unroll
        for (int s = 0; s < kSB; s++) g += go[s] * a2[s][j];
        G[kOffW3 + j] = g;
        if (j == 0) {
            float gb = 0.0f;
#pragma unroll
            for (int s = 0; s < kSB; s++) gb += go[s];
            G[kOffB3] = gb;
        }
    }
    // dz2 into red (all reads of red by the output reduction are complete)
#pragma unroll
    for (int s = 0; s < kSB; s++) red[s][j] = go[s] * w3 * gelu_grad(z2[s][j]);
    __syncthreads();
    {
        float gb = 0.0f;
#pragma unroll
        for (int s = 0; s < kSB; s++) gb += red[s][j];
        G[kOffB2 + j] = gb;
        for (int k = 0; k < kHid; k++) {
            float g = 0.0f;
#pragma unroll
            for (int s = 0; s < kSB; s++) g += red[s][j] * a1[s][k];
            G[kOffW2 + j * kHid + k] = g;
        }
    }
    // hidden unit j of layer 1: da1 = sum over layer-2 units jj of W2[jj][j] dz2[jj]
    float da[kSB];
#pragma unroll
    for (int s = 0; s < kSB; s++) da[s] = 0.0f;
    for (int jj = 0; jj < kHid; jj++) {
        const float w = bf16_to_f32(M.w[kOffW2 + jj * kHid + j]);
#pragma unroll
        for (int s = 0; s < kSB; s++) da[s] += w * red[s][jj];
    }
    float gb1 = 0.0f, gw1[kIn];
#pragma unroll
    for (int i = 0; i < kIn; i++) gw1[i] = 0.0f;
#pragma unroll
    for (int s = 0; s < kSB; s++) {
        const float d = da[s] * gelu_grad(z1[s][j]);
        gb1 += d;
#pragma unroll
        for (int i = 0; i < kIn; i++) gw1[i] += d * xs[s][i];
    }
    G[kOffB1 + j] = gb1;
#pragma unroll
    for (int i = 0; i < kIn; i++) G[kOffW1 + j * kIn + i] = gw1[i];
}

// g[p] = sum over workgroup slots in order
__global__ void mlp_gsum_kernel(const float *__restrict__ gpart, int nslot, float *__restrict__ g)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParam) return;
    float s = 0.0f;
    for (int q = 0; q < nslot; q++) s += gpart[(size_t)q * kNParam + p];
    g[p] = s;
}

__global__ void mlp_adam_kernel(float *__restrict__ p, float *__restrict__ m1, float *__restrict__ m2,
                                uint16_t *__restrict__ w, const float *__restrict__ g, float lr, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNParam) return;
    const float gi = g[i];
    const float m = 0.9f * m1[i] + 0.1f * gi;
    const float v = 0.999f * m2[i] + 0.001f * gi * gi;
    m1[i] = m;
    m2[i] = v;
    const float pi = p[i] - lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
    p[i] = pi;
    w[i] = f32_to_bf16_rne(pi);
}

__global__ void mlp_round_kernel(const float *__restrict__ p, uint16_t *__restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kNParam) w[i] = f32_to_bf16_rne(p[i]);
}

static void refresh_bf16(hbi_mlp *m)
{
    mlp_round_kernel<<<(kNParam + 255) / 256, 256, 0, m->stream>>>(m->p.p, m->w.p);
    HBI_HIP(hipGetLastError());
}

// forward (+ backward) over samples [off, off + nb) of X (through perm when
// given); returns the summed squared error, gradient into g when not NULL
static double run_batch(hbi_mlp *m, const float *dX, const float *dT, const int *dperm, int off, int nb, float *dY,
                        DevBuf<float> &lossp, DevBuf<float> &gpart, float *dg)
{
    const int nwg = (nb + kSB - 1) / kSB;
    if ((int)lossp.n < nwg) lossp.alloc(nwg);
    if (dg && gpart.n < (size_t)nwg * kNParam) gpart.alloc((size_t)nwg * kNParam);
    mlp_fb_kernel<<<nwg, kHid, 0, m->stream>>>(dX, dT, dperm, off, nb, m->dev(), dY, lossp.p, dg ? gpart.p : nullptr);
    HBI_HIP(hipGetLastError());
    if (dg) mlp_gsum_kernel<<<(kNParam + 255) / 256, 256, 0, m->stream>>>(gpart.p, nwg, dg);
    std::vector<float> lp(nwg);
    lossp.download(lp.data(), nwg, m->stream);
    HBI_HIP(hipStreamSynchronize(m->stream));
    double L = 0.0;
    for (float v : lp) L += v;
    return L;
}

}  // namespace hbi

extern "C" {

int hbi_mlp_create(unsigned long long seed, hbi_mlp **out)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(out, "NULL argument");
        *out = nullptr;
        hbi_mlp *m = new hbi_mlp();
        try {
            HBI_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
            // He-uniform weights for the GELU layers, small output layer, zero biases
            std::vector<float> p(hbi::kNParam, 0.0f);
            std::mt19937_64 rng(seed);
            auto fill = [&](int off, int count, int fan_in, float gain) {
                const float lim = gain * std::sqrt(6.0f / (float)fan_in);
                std::uniform_real_distribution<float> U(-lim, lim);
                for (int i = 0; i < count; i++) p[off + i] = U(rng);
            };
            fill(hbi::kOffW1, hbi::kHid * hbi::kIn, hbi::kIn, 1.0f);
            fill(hbi::kOffW2, hbi::kHid * hbi::kHid, hbi::kHid, 1.0f);
            fill(hbi::kOffW3, hbi::kHid, hbi::kHid, 0.1f);
            m->p.upload(p, m->stream);
            std::vector<float> z(hbi::kNParam, 0.0f), mu(hbi::kIn, 0.0f), sd(hbi::kIn, 1.0f);
            m->m1.upload(z, m->stream);
            m->m2.upload(z, m->stream);
            m->mu.upload(mu, m->stream);
            m->sd.upload(sd, m->stream);
            m->w.alloc(hbi::kNParam);
            hbi::refresh_bf16(m);
            HBI_HIP(hipStreamSynchronize(m->stream));
        } catch (...) {
            hbi_mlp_destroy(m);
            throw;
        }
        *out = m;
    });
}

int hbi_mlp_destroy(hbi_mlp *m)
{
    if (!m) return HBI_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    hipStream_t s = m->stream;
    delete m;
    if (s) (void)hipStreamDestroy(s);
    return HBI_OK;
}

int hbi_mlp_nparams(void) { return hbi::kNParam; }

int hbi_mlp_set_params(hbi_mlp *m, const float *params, const float *mu, const float *sd)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(m, "NULL argument");
        if (params) {
            m->p.upload(params, hbi::kNParam, m->stream);
            std::vector<float> z(hbi::kNParam, 0.0f);
            m->m1.upload(z, m->stream);
            m->m2.upload(z, m->stream);
            m->adam_t = 0;
            hbi::refresh_bf16(m);
        }
        if (mu) m->mu.upload(mu, hbi::kIn, m->stream);
        if (sd) {
            for (int i = 0; i < hbi::kIn; i++) HBI_REQUIRE(sd[i] > 0.0f, "feature scales must be > 0");
            m->sd.upload(sd, hbi::kIn, m->stream);
        }
        HBI_HIP(hipStreamSynchronize(m->stream));
    });
}

int hbi_mlp_get_params(const hbi_mlp *m, float *params, float *mu, float *sd)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(m, "NULL argument");
        if (params) m->p.download(params, hbi::kNParam, m->stream);
        if (mu) m->mu.download(mu, hbi::kIn, m->stream);
        if (sd) m->sd.download(sd, hbi::kIn, m->stream);
        HBI_HIP(hipStreamSynchronize(m->stream));
    });
}

int hbi_mlp_forward(hbi_mlp *m, const float *X, int n, float *Y)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(m && X && Y && n > 0, "bad arguments");
        hbi::DevBuf<float> dX, dY(n), lp, gp;
        dX.upload(X, (size_t)n * hbi::kIn, m->stream);
        hbi::run_batch(m, dX.p, nullptr, nullptr, 0, n, dY.p, lp, gp, nullptr);
        dY.download(Y, n, m->stream);
        HBI_HIP(hipStreamSynchronize(m->stream));
    });
}

int hbi_mlp_loss_grad(hbi_mlp *m, const float *X, const float *T, int n, double *loss, float *grad)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(m && X && T && n > 0, "bad arguments");
        hbi::DevBuf<float> dX, dT, lp, gp, dg(hbi::kNParam);
        dX.upload(X, (size_t)n * hbi::kIn, m->stream);
        dT.upload(T, n, m->stream);
        const double L = hbi::run_batch(m, dX.p, dT.p, nullptr, 0, n, nullptr, lp, gp, grad ? dg.p : nullptr);
        if (loss) *loss = L / n;
        if (grad) {
            dg.download(grad, hbi::kNParam, m->stream);
            HBI_HIP(hipStreamSynchronize(m->stream));
        }
    });
}

int hbi_mlp_train(hbi_mlp *m, const float *X, const float *T, int n, int epochs, int batch, float lr,
                  unsigned long long seed, double *epoch_loss)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(m && X && T && n > 0 && epochs >= 0 && batch > 0 && lr > 0, "bad arguments");
        hbi::DevBuf<float> dX, dT, lp, gp, dg(hbi::kNParam);
        hbi::DevBuf<int> dperm;
        dX.upload(X, (size_t)n * hbi::kIn, m->stream);
        dT.upload(T, n, m->stream);
        std::vector<int> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::mt19937_64 rng(seed);
        for (int e = 0; e < epochs; e++) {
            std::shuffle(perm.begin(), perm.end(), rng);
            dperm.upload(perm, m->stream);
            double L = 0.0;
            for (int off = 0; off < n; off += batch) {
                const int nb = std::min(batch, n - off);
                L += hbi::run_batch(m, dX.p, dT.p, dperm.p, off, nb, nullptr, lp, gp, dg.p);
                m->adam_t++;
                const float c1 = (float)(1.0 - std::pow(0.9, (double)m->adam_t));
                const float c2 = (float)(1.0 - std::pow(0.999, (double)m->adam_t));
                hbi::mlp_adam_kernel<<<(hbi::kNParam + 255) / 256, 256, 0, m->stream>>>(m->p.p, m->m1.p, m->m2.p,
                                                                                         m->w.p, dg.p, lr, c1, c2);
                HBI_HIP(hipGetLastError());
            }
            if (epoch_loss) epoch_loss[e] = L / n;
        }
        HBI_HIP(hipStreamSynchronize(m->stream));
    });
}

}  // extern "C"
