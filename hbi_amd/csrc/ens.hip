// ens.hip -- ensembles of fault realisations on one H-matrix.
//
// R <= 64 realisations (different friction parameters, same fault mesh) are
// advanced in lockstep: every attempt evaluates the six stages of all
// realisations with one 64-column H product per stage (multirhs.hip), while
// each realisation keeps its own step size, time, embedded-error decision and
// step count (the reference's rkqs, main_LH.f90:2270-2370, applied per
// column).  A realisation that reached its step target is frozen: its columns
// are still computed, its state no longer changes.  One attempt is captured
// once as a HIP graph.
//
// Device layouts put the realisation index innermost: y[(3 i + c) * 64 + r],
// parameters p[i * 64 + r], so one wavefront covers the 64 realisations of
// an element with unit-stride accesses.
#include "rkcore.hpp"
#include "learn.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hbi {

constexpr int kR = 64;
constexpr int kHistCols = 5;      // (t, dt, errmax, rejections before the step, log10 Vmax)
constexpr int kEnsProbes = 16;    // probe elements at most; each adds (V, slip) to a history row

struct EnsPar {
    int n, nreal, law, backslip, limitsigma, tr;
    double vref, minsig, maxsig, dtmax, hdiv, eps, velmin, velmax;
    Tab T;
    int learn, use_mlp;                   // step features on; MLP proposals on
    double safety, ratio_lo, ratio_hi;    // learned proposal: factor and bounds relative to rkqs
    int nprobe;                           // history rows carry (V, slip) of these elements
    int probe[kEnsProbes];
    // replayed step grid (or null): accepted step k of realisation r has size
    // grid[r * gcap + k]; every attempt is accepted
    const double *grid;
    long long gcap;
};

struct EnsCtl {           // per realisation
    double h[kR], t[kR], hdid[kR], errmax[kR], vmax[kR];
    long long accepted[kR], rejected[kR], target[kR], hist_base[kR];
    long long hist_cap;
    int acc[kR], frozen[kR], status[kR];
    int done;
    // learned step size: previous accepted step, rejections since it, trace rows written
    StepMem mem[kR];
    int nrjct[kR], trace_n[kR];
};

// friction / loading arrays: per element and realisation (i * 64 + r), or,
// when all realisations share the array, once per element (i); bit k of
// `shared` is set for array k in the order a, b, dc, f0, vpl, taudot, sigdot
enum { kPa, kPb, kPdc, kPf0, kPvpl, kPtaudot, kPsigdot, kNPar };
struct EnsArrays {
    const double *a, *b, *dc, *f0, *vpl, *taudot, *sigdot;
    int shared;
    double *y, *yout, *k, *vel, *x, *slip, *velc, *errpart, *monpart;
    float *x32;           // FP32 copy of x for an FP32-core product (written instead of x), else null
    const double *ypart;
    const int *part0;
    int *bins;            // kR x kBins histograms of log10 |V| after an accepted step
    double *trace;        // kR x trace_cap rows of 12 features + (h, errmax)
    int trace_cap;
    MlpDev mlp;
};

__device__ __forceinline__ double prm(const EnsArrays &A, const double *p, int k, int i, int r)
{
    return (A.shared >> k & 1) ? p[i] : p[(size_t)i * kR + r];
}

// thread (i, r): r = threadIdx.x % 64, i = blockIdx.x * 4 + threadIdx.x / 64
__global__ void ens_stage0_kernel(EnsPar P, const EnsCtl *__restrict__ c, EnsArrays A)
{
#pragma clang fp contract(off)
    if (c->done) return;
    const int r = threadIdx.x % kR, i = blockIdx.x * 4 + threadIdx.x / kR;
    if (i >= P.n) return;
    const size_t q = (size_t)i * kR + r;
    if (r >= P.nreal) {
        if (A.x32) A.x32[q] = 0.f;
        else A.x[q] = 0.0;
        return;
    }
    const double *y = A.y;
    const double v = slip_rate(y[(3 * (size_t)i) * kR + r], y[(3 * (size_t)i + 1) * kR + r],
                                 y[(3 * (size_t)i + 2) * kR + r], prm(A, A.a, kPa, i, r), P.vref);
    A.vel[q] = v;
    const double xv = P.backslip ? v - prm(A, A.vpl, kPvpl, i, r) : v;
    if (A.x32) A.x32[q] = (float)xv;
    else A.x[q] = xv;
}

// State of Cash-Karp stage t (1..5) from y and the slopes kv[0 .. t-1]
// (main_LH.f90:2501-2536); the stage that forms it and the stage that uses
// it evaluate this same expression, so both see the same bits.
__device__ __forceinline__ double stage_state(const EnsPar &P, int t, double h, double y, const double *kv)
{
#pragma clang fp contract(off)
    switch (t) {
    case 1: return y + P.T.b21 * h * kv[0];
    case 2: return y + h * (P.T.b31 * kv[0] + P.T.b32 * kv[1]);
    case 3: return y + h * (P.T.b41 * kv[0] + P.T.b42 * kv[1] + P.T.b43 * kv[2]);
    case 4: return y + h * (P.T.b51 * kv[0] + P.T.b52 * kv[1] + P.T.b53 * kv[2] + P.T.b54 * kv[3]);
    default: return y + h * (P.T.b61 * kv[0] + P.T.b62 * kv[1] + P.T.b63 * kv[2] + P.T.b64 * kv[3] + P.T.b65 * kv[4]);
    }
}

struct EPair {
    int g;
    double v;
};
__device__ __forceinline__ EPair e_combine(EPair l, EPair r)
{
    if (r.g) return r;
    return EPair{l.g, fmax(l.v, r.v)};
}

template <int s, bool SC>
__global__ __launch_bounds__(256) void ens_stage_kernel(EnsPar P, const EnsCtl *__restrict__ c, EnsArrays A)
{
    // SC: sigdot is zero for every element and realisation and there is no
    // normal-stress product, so every sigma slope is exactly 0 (deriv's
    // dsigdt = gn = 0 + 0) and the stage states keep sigma = y bit for bit:
    // the sigma slopes are neither stored nor read back (a third of the
    // slope traffic of a stage)
#pragma clang fp contract(off)
    __shared__ int sg[4][kR];
    __shared__ double sv[4][kR];
    if (c->done) return;
    const int r = threadIdx.x % kR, li = threadIdx.x / kR, i = blockIdx.x * 4 + li;
    EPair ep{0, 0.0};
    if (i < P.n && r < P.nreal) {
        const size_t q = (size_t)i * kR + r;
        const int tr = P.tr, t = i / tr, ro = i % tr;
        double hx = 0.0;
        for (int p = A.part0[t]; p < A.part0[t + 1]; p++) hx += A.ypart[((size_t)p * tr + ro) * kR + r];
        double gs = P.hdiv != 0.0 ? hx / P.hdiv : hx;
        double gn = 0.0;
        gn = gn + prm(A, A.sigdot, kPsigdot, i, r);
        gs = gs + prm(A, A.taudot, kPtaudot, i, r);
        const size_t e0 = (3 * (size_t)i) * kR + r;
        const size_t m3 = 3 * (size_t)P.n * kR;
        const double h = c->h[r];
        // y and the earlier slopes, read once; the stage state is formed from them
        // again rather than stored by the previous stage and read back
        double yv[3], kv[3][6], ys[3];
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {
            const size_t e = e0 + cc * kR;
            yv[cc] = A.y[e];
#pragma unroll
            for (int j = 0; j < s; j++) kv[cc][j] = SC && cc == 2 ? 0.0 : A.k[j * m3 + e];
            ys[cc] = s == 0 ? yv[cc] : stage_state(P, s, h, yv[cc], kv[cc]);
        }
        double kk[3];
        const double ai = prm(A, A.a, kPa, i, r);
        deriv(P, gs, gn, ys[0], ys[1], ys[2], A.vel[q], ai, prm(A, A.b, kPb, i, r), prm(A, A.dc, kPdc, i, r),
              prm(A, A.f0, kPf0, i, r), kk);
        double *ks = A.k + s * m3;
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {
            if (!(SC && cc == 2)) ks[e0 + cc * kR] = kk[cc];
            kv[cc][s] = kk[cc];
        }
        if constexpr (s < 5) {
            double yt[3];
#pragma unroll
            for (int cc = 0; cc < 3; cc++) yt[cc] = stage_state(P, s + 1, h, yv[cc], kv[cc]);
            const double v = slip_rate(yt[0], yt[1], yt[2], ai, P.vref);
            A.vel[q] = v;
            const double xv = P.backslip ? v - prm(A, A.vpl, kPvpl, i, r) : v;
            if (A.x32) A.x32[q] = (float)xv;
            else A.x[q] = xv;
        } else {
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                const double *kc = kv[cc];
                const double yo = yv[cc] + h * (P.T.c1 * kc[0] + P.T.c3 * kc[2] + P.T.c4 * kc[3] + P.T.c6 * kc[5]);
                const double ye = h * (P.T.dc1 * kc[0] + P.T.dc3 * kc[2] + P.T.dc4 * kc[3] + P.T.dc5 * kc[4] + P.T.dc6 * kc[5]);
                const size_t e = e0 + cc * kR;
                A.yout[e] = yo;
                const double er = fabs(ye / yo) / P.eps;
                EPair cur{0, er > 0.0 ? er : 0.0};
                if (yo - 1 == yo) cur = EPair{1, 10.0};
                ep = e_combine(ep, cur);
            }
        }
    }
    if constexpr (s != 5) return;
    sg[li][r] = ep.g;
    sv[li][r] = ep.v;
    __syncthreads();
    if (li == 0) {
        EPair b{sg[0][r], sv[0][r]};
        for (int q = 1; q < 4; q++) b = e_combine(b, EPair{sg[q][r], sv[q][r]});
        A.errpart[(2 * (size_t)blockIdx.x) * kR + r] = (double)b.g;
        A.errpart[(2 * (size_t)blockIdx.x + 1) * kR + r] = b.v;
    }
}

// First level of the per-realisation reductions: workgroup g folds a
// contiguous range of the stage blocks' partials into partial g, in order
// (the guarded error combine is associative, so the fold order only has to
// keep the element order).  mode 0: (guard, errmax) pairs, mode 1: max.
constexpr int kFold = 64;

__global__ __launch_bounds__(1024) void ens_fold_kernel(const EnsCtl *__restrict__ c, const double *__restrict__ in,
                                                        int nblk, double *__restrict__ out, int mode)
{
    __shared__ int sg[16][kR];
    __shared__ double sv[16][kR];
    if (c->done) return;
    const int r = threadIdx.x % kR, w = threadIdx.x / kR;
    const int per = (nblk + kFold - 1) / kFold;
    const int g0 = min(nblk, (int)blockIdx.x * per), g1 = min(nblk, g0 + per);
    const int pw = (g1 - g0 + 15) / 16;
    const int b0 = min(g1, g0 + w * pw), b1 = min(g1, b0 + pw);
    EPair e{0, 0.0};
    if (mode == 0) {
        for (int b = b0; b < b1; b++)
            e = e_combine(e, EPair{(int)in[(2 * (size_t)b) * kR + r], in[(2 * (size_t)b + 1) * kR + r]});
    } else {
        for (int b = b0; b < b1; b++) e.v = fmax(e.v, in[(size_t)b * kR + r]);
    }
    sg[w][r] = e.g;
    sv[w][r] = e.v;
    __syncthreads();
    if (w != 0) return;
    EPair t{sg[0][r], sv[0][r]};
    for (int q = 1; q < 16; q++) t = mode == 0 ? e_combine(t, EPair{sg[q][r], sv[q][r]}) : EPair{0, fmax(t.v, sv[q][r])};
    if (mode == 0) {
        out[(2 * (size_t)blockIdx.x) * kR + r] = (double)t.g;
        out[(2 * (size_t)blockIdx.x + 1) * kR + r] = t.v;
    } else {
        out[(size_t)blockIdx.x * kR + r] = t.v;
    }
}

// 1024 threads: lane = realisation, wave w combines a contiguous range of the
// stage blocks' partials, then wave 0 combines the 16 ranges in order
constexpr int kDecThreads = 1024, kDecWaves = kDecThreads / kR;

__global__ __launch_bounds__(kDecThreads) void ens_decide_kernel(EnsPar P, EnsCtl *c, const double *__restrict__ errpart,
                                                                 int nblk)
{
#pragma clang fp contract(off)
    __shared__ int sg[kDecWaves][kR];
    __shared__ double sv[kDecWaves][kR];
    if (c->done) return;
    const int r = threadIdx.x % kR, w = threadIdx.x / kR;
    const int per = (nblk + kDecWaves - 1) / kDecWaves;
    const int b0 = min(nblk, w * per), b1 = min(nblk, b0 + per);
    EPair e{0, 0.0};
    for (int b = b0; b < b1; b++)
        e = e_combine(e, EPair{(int)errpart[(2 * (size_t)b) * kR + r], errpart[(2 * (size_t)b + 1) * kR + r]});
    sg[w][r] = e.g;
    sv[w][r] = e.v;
    __syncthreads();
    if (w != 0) return;
    if (r < P.nreal && !c->frozen[r]) {
        const double SAFETY = (double)0.9f, PGROW = (double)-0.2f, PSHRNK = (double)-0.25f, tiny = (double)1e-20f;
        EPair t{sg[0][r], sv[0][r]};
        for (int q = 1; q < kDecWaves; q++) t = e_combine(t, EPair{sg[q][r], sv[q][r]});
        const double errmax = t.v, h = c->h[r];
        c->errmax[r] = errmax;
        if (P.grid) {
            // replay: the step is taken whatever errmax; the next size comes from the grid
            c->hdid[r] = h;
            c->t[r] = c->t[r] + h;
            c->acc[r] = 1;
            c->accepted[r]++;
            if (c->accepted[r] < P.gcap) c->h[r] = P.grid[(size_t)r * P.gcap + c->accepted[r]];
        } else if (errmax < 1.0 && errmax > tiny) {
            double hn = fmin(2 * h, SAFETY * h * pow(errmax, PGROW));
            hn = fmin(hn, P.dtmax);
            c->hdid[r] = h;
            c->t[r] = c->t[r] + h;
            c->h[r] = hn;
            c->acc[r] = 1;
            c->accepted[r]++;
        } else {
            const double hnew = errmax > tiny ? fmax(0.5 * h, SAFETY * h * pow(errmax, PSHRNK)) : 0.5 * h;
            c->acc[r] = 0;
            c->rejected[r]++;
            c->nrjct[r]++;
            c->h[r] = hnew;
            if (c->t[r] + hnew - c->t[r] < 1.e-15) {
                c->status[r] = HBI_E_STEP;
                c->frozen[r] = 1;
            }
        }
    } else {
        c->acc[r] = 0;
    }
}

// Each wave commits kCommitRun consecutive elements for its 64 realisations
// (lane = realisation).
constexpr int kCommitRun = 4, kCommitPerBlock = 4 * kCommitRun;

__global__ __launch_bounds__(256) void ens_commit_kernel(EnsPar P, const EnsCtl *__restrict__ c, EnsArrays A)
{
#pragma clang fp contract(off)
    __shared__ double vm[4][kR];
    if (c->done) return;
    const int r = threadIdx.x % kR, w = threadIdx.x / kR;
    const int i0 = blockIdx.x * kCommitPerBlock + w * kCommitRun;
    double vmax = 0.0;
    if (r < P.nreal && c->acc[r]) {
        const double dt = c->hdid[r];
        // distinct arrays: the loads of an element need not wait for the stores of the previous one
        double *__restrict__ y = A.y, *__restrict__ slip = A.slip, *__restrict__ velc = A.velc;
        const double *__restrict__ yout = A.yout;
        for (int k = 0; k < kCommitRun; k++) {
            const int i = i0 + k;
            if (i >= P.n) break;
            const size_t q = (size_t)i * kR + r, e0 = (3 * (size_t)i) * kR + r;
            double yv[3];
            for (int cc = 0; cc < 3; cc++) {
                yv[cc] = yout[e0 + cc * kR];
                y[e0 + cc * kR] = yv[cc];
            }
            double sl = slip[q];
            sl = sl + velc[q] * dt * 0.5;
            const double v = slip_rate(yv[0], yv[1], yv[2], prm(A, A.a, kPa, i, r), P.vref);
            sl = sl + v * dt * 0.5;
            slip[q] = sl;
            velc[q] = v;
            vmax = fmax(vmax, fabs(v));
        }
    }
    vm[w][r] = vmax;
    __syncthreads();
    if (w == 0) A.monpart[(size_t)blockIdx.x * kR + r] = fmax(fmax(vm[0][r], vm[1][r]), fmax(vm[2][r], vm[3][r]));
}

// log10 |V| histograms of the accepted realisations (the learned controller's
// features): workgroup (g, m) counts realisations 4 g .. 4 g + 3 over element
// chunk m of kHistChunks into LDS -- each thread a contiguous run of elements
// of one realisation, counted as runs of equal bins (most elements of a fault
// share a few bins) -- then adds its nonzero bins to the global histograms (at
// most kHistChunks adds per address)
// 64 chunks: 1024 workgroups (4 waves per SIMD), each thread at most 8
// elements of one load batch for the 32768-element fault
constexpr int kHistR = 4, kHistChunks = 64;

__global__ __launch_bounds__(256) void ens_hist_kernel(EnsPar P, const EnsCtl *__restrict__ c, EnsArrays A)
{
    __shared__ int hl[kHistR][kBins];
    if (c->done) return;
    const int g = blockIdx.x / kHistChunks, m = blockIdx.x % kHistChunks;
    const int rl = threadIdx.x % kHistR, r = kHistR * g + rl;
    for (int b = threadIdx.x; b < kHistR * kBins; b += blockDim.x) hl[b / kBins][b % kBins] = 0;
    __syncthreads();
    const int per = (P.n + kHistChunks - 1) / kHistChunks;
    const int lanes = (int)blockDim.x / kHistR, run = (per + lanes - 1) / lanes;
    const int i0 = m * per + ((int)threadIdx.x / kHistR) * run;
    const int i1 = min(min(P.n, (m + 1) * per), i0 + run);
    if (r < P.nreal && c->acc[r]) {
        // kHistLd values loaded together (indices clamped into the run), then
        // binned in element order
        constexpr int kHistLd = 8;
        int cur_b = -1, cur_n = 0;
        for (int i = i0; i < i1; i += kHistLd) {
            double v[kHistLd];
#pragma unroll
            for (int k = 0; k < kHistLd; k++) v[k] = A.velc[(size_t)min(i + k, i1 - 1) * kR + r];
#pragma unroll
            for (int k = 0; k < kHistLd; k++) {
                if (i + k >= i1) break;
                const int b = vbin(fabs(v[k]));
                if (b == cur_b) {
                    cur_n++;
                } else {
                    if (cur_n) atomicAdd(&hl[rl][cur_b], cur_n);
                    cur_b = b;
                    cur_n = 1;
                }
            }
        }
        if (cur_n) atomicAdd(&hl[rl][cur_b], cur_n);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kHistR * kBins; b += blockDim.x) {
        const int rr = kHistR * g + b / kBins, n = hl[b / kBins][b % kBins];
        if (n) atomicAdd(&A.bins[(size_t)rr * kBins + b % kBins], n);
    }
}

// history row per accepted step of realisation r: (t, dt, errmax, rejections before it, log10 Vmax),
// then (V, slip) of every probe element
__global__ __launch_bounds__(kDecThreads) void ens_monitor_kernel(EnsPar P, EnsCtl *c, const double *__restrict__ monpart,
                                                                  int nblk, double *__restrict__ hist,
                                                                  const double *__restrict__ velc,
                                                                  const double *__restrict__ slip)
{
    __shared__ double sm[kDecWaves][kR];
    if (c->done) return;
    const int r = threadIdx.x % kR, w = threadIdx.x / kR;
    const int per = (nblk + kDecWaves - 1) / kDecWaves;
    const int b0 = min(nblk, w * per), b1 = min(nblk, b0 + per);
    double m = 0.0;
    for (int b = b0; b < b1; b++) m = fmax(m, monpart[(size_t)b * kR + r]);
    sm[w][r] = m;
    __syncthreads();
    if (w != 0) return;
    if (r < P.nreal) {
        if (c->acc[r]) {
            double v = sm[0][r];
            for (int q = 1; q < kDecWaves; q++) v = fmax(v, sm[q][r]);
            c->vmax[r] = v;
            const long long row = c->accepted[r] - 1 - c->hist_base[r];
            if (hist && row >= 0 && row < c->hist_cap) {
                double *o = hist + ((size_t)r * c->hist_cap + row) * (kHistCols + 2 * P.nprobe);
                o[0] = c->t[r];
                o[1] = c->hdid[r];
                o[2] = c->errmax[r];
                o[3] = (double)c->nrjct[r];
                o[4] = log10(v);
                for (int q = 0; q < P.nprobe; q++) {
                    const size_t e = (size_t)P.probe[q] * kR + r;
                    o[kHistCols + 2 * q] = velc[e];
                    o[kHistCols + 2 * q + 1] = slip[e];
                }
            }
            c->nrjct[r] = 0;
            if (v > P.velmax || v < P.velmin) { c->status[r] = 1; c->frozen[r] = 1; }
            if (c->accepted[r] >= c->target[r]) c->frozen[r] = 1;
        }
    }
    const unsigned long long fz = __ballot(r < P.nreal && c->frozen[r]);
    if (r == 0 && __popcll(fz) >= P.nreal) c->done = 1;
}

// One workgroup per realisation that accepted a step: its step features,
// optional trace row and, with the MLP on, the learned next h (rkqs stays
// the judge; a rejection shrinks h by the rkqs rule).  Clears its histogram.
__global__ __launch_bounds__(kHid) void ens_learn_kernel(EnsPar P, EnsCtl *c, EnsArrays A)
{
    __shared__ int part[kHid + 1], hb[kBins];
    __shared__ float x[kIn], h1[kHid], red[kHid];
    __shared__ double q[3];
    const int r = blockIdx.x, j = threadIdx.x;
    if (c->done || !P.learn || !c->acc[r]) return;
    MlpRow R;
    if (P.use_mlp && !P.grid) R = mlp_load_row(A.mlp);
    int *bins = A.bins + (size_t)r * kBins;
    step_features(bins, P.n, c->hdid[r], c->errmax[r], c->nrjct[r] > 0, c->mem[r], part, hb, x, q);
    if (j == 0 && A.trace && c->trace_n[r] < A.trace_cap) {
        double *row = A.trace + 14 * ((size_t)r * A.trace_cap + c->trace_n[r]);
        for (int k = 0; k < kIn; k++) row[k] = (double)x[k];
        row[12] = c->hdid[r];
        row[13] = c->errmax[r];
        c->trace_n[r]++;
    }
    if (P.use_mlp && !P.grid) {
        const float out = mlp_infer_block(A.mlp, R, x, h1, red);
        if (j == 0) c->h[r] = learned_h(x[11], out, c->h[r], P.safety, P.ratio_lo, P.ratio_hi, P.dtmax);
    }
    if (j == 0) {
        StepMem &m = c->mem[r];
        for (int k = 0; k < 3; k++) m.prev_q[k] = q[k];
        m.prev_h = c->hdid[r];
        m.prev_e = c->errmax[r];
        m.have_prev = 1;
    }
    for (int b = 0; b < kBins / kHid; b++) bins[j * (kBins / kHid) + b] = 0;
}

__global__ void ens_arm_kernel(EnsCtl *c, int nreal, long long nsteps, long long hist_cap, const double *grid,
                               long long gcap)
{
    const int r = threadIdx.x;
    if (r == 0) c->hist_cap = hist_cap;
    if (r < nreal) {
        if (grid && c->accepted[r] < gcap) c->h[r] = grid[(size_t)r * gcap + c->accepted[r]];
        c->hist_base[r] = c->accepted[r];
        c->target[r] = c->accepted[r] + nsteps;
        c->frozen[r] = c->status[r] != 0 ? 1 : 0;
        c->acc[r] = 0;
    }
    if (r == 0) c->done = nsteps > 0 ? 0 : 1;
}

}  // namespace hbi

struct hbi_ens {
    hipStream_t stream = nullptr;
    hbi_hmat *hs = nullptr;
    hbi_rk_opts opts{};
    hbi::EnsPar P{};
    int n = 0, nreal = 0, nblk = 0;
    int shared = 0;   // arrays kept once per element (bit k: EnsArrays order)
    int sig_const = 0;   // every sigdot is 0: sigma slopes are zero and not stored
    hbi::DevBuf<double> a, b, dc, f0, vpl, taudot, sigdot;
    hbi::DevBuf<double> y, yout, k, vel, x, slip, velc, errpart, monpart, z, zpart, ypart;
    hbi::DevBuf<float> x32;      // FP32 copy of x for the FP32-core product
    bool x32_direct = false;     // the stage kernels write x32 themselves (FP32-core product)
    hbi::DevBuf<double> fold;   // first-level reduction partials (kFold x 2 x 64)
    hbi::DevBuf<double> hist;   // nreal-major rows of kHistCols + 2 nprobe (graph argument)
    long long hist_rows = 0;    // rows the buffer holds
    hbi::DevBuf<hbi::EnsCtl> ctl;
    hbi::DevBuf<int> bins;
    hbi::DevBuf<double> grid;   // replayed step sizes, nreal x gcap (hbi_ens_set_step_grid)
    hbi::DevBuf<double> trace;
    int trace_cap = 0;
    hbi::DevBuf<float> mlp_p, mlp_mu, mlp_sd;   // copy of the step-size MLP
    hbi::DevBuf<uint16_t> mlp_w;
    hipGraphExec_t graph = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
};

namespace hbi {

static EnsArrays e_arrays(hbi_ens *E)
{
    EnsArrays A{};
    A.shared = E->shared;
    A.a = E->a.p; A.b = E->b.p; A.dc = E->dc.p; A.f0 = E->f0.p; A.vpl = E->vpl.p; A.taudot = E->taudot.p;
    A.sigdot = E->sigdot.p; A.y = E->y.p; A.yout = E->yout.p; A.k = E->k.p; A.vel = E->vel.p;
    A.x = E->x.p;
    A.x32 = E->x32_direct ? E->x32.p : nullptr; A.slip = E->slip.p; A.velc = E->velc.p; A.errpart = E->errpart.p; A.monpart = E->monpart.p;
    A.ypart = E->ypart.p;
    A.part0 = mr_part0(E->hs, E->opts.prec);
    A.bins = E->bins.p;
    A.trace = E->trace.p;
    A.trace_cap = E->trace_cap;
    A.mlp = MlpDev{E->mlp_w.p, E->mlp_p.p, E->mlp_mu.p, E->mlp_sd.p};
    return A;
}

static void e_attempt(hbi_ens *E)
{
    const int g = E->nblk;
    EnsArrays A = e_arrays(E);
    const int *gate = &E->ctl.p->done;
    ens_stage0_kernel<<<g, 256, 0, E->stream>>>(E->P, E->ctl.p, A);
    using StageFn = void (*)(EnsPar, const EnsCtl *, EnsArrays);
    static const StageFn stage[2][6] = {
        {ens_stage_kernel<0, false>, ens_stage_kernel<1, false>, ens_stage_kernel<2, false>,
         ens_stage_kernel<3, false>, ens_stage_kernel<4, false>, ens_stage_kernel<5, false>},
        {ens_stage_kernel<0, true>, ens_stage_kernel<1, true>, ens_stage_kernel<2, true>,
         ens_stage_kernel<3, true>, ens_stage_kernel<4, true>, ens_stage_kernel<5, true>}};
    for (int s = 0; s < 6; s++) {
        mr_apply_parts(E->hs, E->opts.prec, E->x32_direct ? nullptr : E->x.p, E->x32.p, E->zpart.p, E->z.p,
                       E->ypart.p, E->stream, gate);
        stage[E->sig_const][s]<<<g, 256, 0, E->stream>>>(E->P, E->ctl.p, A);
    }
    ens_fold_kernel<<<kFold, 1024, 0, E->stream>>>(E->ctl.p, E->errpart.p, g, E->fold.p, 0);
    ens_decide_kernel<<<1, kDecThreads, 0, E->stream>>>(E->P, E->ctl.p, E->fold.p, kFold);
    const int gc = (E->n + kCommitPerBlock - 1) / kCommitPerBlock;
    ens_commit_kernel<<<gc, 256, 0, E->stream>>>(E->P, E->ctl.p, A);
    if (E->P.learn) {
        ens_hist_kernel<<<(kR / kHistR) * kHistChunks, 256, 0, E->stream>>>(E->P, E->ctl.p, A);
        ens_learn_kernel<<<E->nreal, kHid, 0, E->stream>>>(E->P, E->ctl.p, A);
    }
    ens_fold_kernel<<<kFold, 1024, 0, E->stream>>>(E->ctl.p, E->monpart.p, gc, E->fold.p, 1);
    ens_monitor_kernel<<<1, kDecThreads, 0, E->stream>>>(E->P, E->ctl.p, E->fold.p, kFold, E->hist.p, E->velc.p,
                                                        E->slip.p);
}

// host (realisation-major, nreal x n x comps) -> device (element-major,
// n x comps x 64); the unused realisation slots r >= nreal get pad[c]
static void upload_rm(DevBuf<double> &d, const double *src, int nreal, int n, int comps, const double *pad,
                      hipStream_t s)
{
    std::vector<double> t((size_t)n * comps * kR);
    for (size_t e = 0; e < (size_t)n * comps; e++)
        for (int r = nreal; r < kR; r++) t[e * kR + r] = pad[e % comps];
    for (int r = 0; r < nreal; r++)
        for (size_t e = 0; e < (size_t)n * comps; e++) t[e * kR + r] = src[(size_t)r * n * comps + e];
    d.upload(t, s);
}
static void download_rm(const DevBuf<double> &d, double *dst, int nreal, int n, int comps, hipStream_t s)
{
    std::vector<double> t((size_t)n * comps * kR);
    d.download(t.data(), t.size(), s);
    HBI_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < nreal; r++)
        for (size_t e = 0; e < (size_t)n * comps; e++) dst[(size_t)r * n * comps + e] = t[e * kR + r];
}

}  // namespace hbi

extern "C" {

int hbi_ens_create(hbi_hmat *hs, const hbi_rk_opts *opts, int nreal, const double *a, const double *b,
                   const double *dc, const double *f0, const double *vpl, const double *taudot,
                   const double *sigdot, hbi_ens **out)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(hs && opts && out && a && b && dc && f0 && vpl && taudot && sigdot, "NULL argument");
        HBI_REQUIRE(nreal >= 1 && nreal <= hbi::kR, "1 <= nreal <= 64");
        HBI_REQUIRE(opts->prec >= 0 && opts->prec <= 2 && opts->eps_r > 0, "bad options");
        *out = nullptr;
        hbi_ens *E = new hbi_ens();
        try {
            E->hs = hs;
            E->opts = *opts;
            const int n = hs->n;
            E->n = n;
            E->nreal = nreal;
            E->nblk = (n + 3) / 4;
            HBI_HIP(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
            HBI_HIP(hipEventCreate(&E->e0));
            HBI_HIP(hipEventCreate(&E->e1));
            hbi::EnsPar &P = E->P;
            P.n = n; P.nreal = nreal; P.law = opts->law; P.backslip = opts->backslip; P.limitsigma = opts->limitsigma;
            P.tr = hs->tr; P.vref = opts->vref; P.minsig = opts->minsig; P.maxsig = opts->maxsig;
            P.dtmax = opts->dtmax; P.hdiv = opts->hdiv; P.eps = opts->eps_r; P.velmin = opts->velmin;
            P.velmax = opts->velmax;
            P.T = hbi::rk_tableau();
            // an array equal for every realisation is kept once per element;
            // unused realisation slots are never evaluated (a = 1 pads the rest)
            const double one = 1.0, zero = 0.0;
            const double *src[hbi::kNPar] = {a, b, dc, f0, vpl, taudot, sigdot};
            hbi::DevBuf<double> *dst[hbi::kNPar] = {&E->a, &E->b, &E->dc, &E->f0, &E->vpl, &E->taudot, &E->sigdot};
            E->shared = 0;
            E->sig_const = 1;
            for (size_t q = 0; q < (size_t)nreal * n && E->sig_const; q++) E->sig_const = sigdot[q] == 0.0;
            for (int k = 0; k < hbi::kNPar; k++) {
                bool same = true;
                for (int r = 1; r < nreal && same; r++)
                    same = std::memcmp(src[k], src[k] + (size_t)r * n, (size_t)n * sizeof(double)) == 0;
                if (same) {
                    dst[k]->upload(src[k], n, E->stream);
                    E->shared |= 1 << k;
                } else {
                    hbi::upload_rm(*dst[k], src[k], nreal, n, 1, (k == hbi::kPa || k == hbi::kPdc) ? &one : &zero,
                                   E->stream);
                }
            }
            const size_t m = (size_t)n * hbi::kR;
            E->y.alloc(3 * m); E->yout.alloc(3 * m); E->k.alloc(18 * m);
            E->vel.alloc(m); E->x.alloc(m); E->slip.alloc(m); E->velc.alloc(m);
            E->errpart.alloc(2 * (size_t)E->nblk * hbi::kR);
            E->monpart.alloc((size_t)E->nblk * hbi::kR);
            E->fold.alloc(2 * (size_t)hbi::kFold * hbi::kR);
            E->bins.alloc((size_t)hbi::kR * hbi::kBins);
            HBI_HIP(hipMemsetAsync(E->bins.p, 0, E->bins.n * sizeof(int), E->stream));
            P.safety = 0.9;
            P.ratio_lo = 0.125;
            P.ratio_hi = 8.0;
            E->z.alloc((size_t)std::max(hs->nz, 1) * hbi::kR);
            E->zpart.alloc(hbi::mr_zpart_size(hs, opts->prec));
            E->x32.alloc(hbi::mr_x32_size(hs, opts->prec));
            E->x32_direct = hbi::mr_x32_size(hs, opts->prec) > 1;
            if (E->x32_direct) HBI_HIP(hipMemsetAsync(E->x32.p, 0, E->x32.n * sizeof(float), E->stream));
            E->ypart.alloc(hbi::mr_ypart_size(hs, opts->prec));
            for (auto *d : {&E->y, &E->yout, &E->k, &E->vel, &E->x, &E->slip, &E->velc})
                HBI_HIP(hipMemsetAsync(d->p, 0, d->n * 8, E->stream));
            E->ctl.alloc(1);
            hbi::EnsCtl c{};
            for (int r = 0; r < hbi::kR; r++) c.h[r] = 1.0;
            hbi::to_device_sync(E->ctl.p, &c, sizeof(c), E->stream);
            HBI_HIP(hipStreamSynchronize(E->stream));
        } catch (...) {
            hbi_ens_destroy(E);
            throw;
        }
        *out = E;
    });
}

int hbi_ens_destroy(hbi_ens *E)
{
    if (!E) return HBI_OK;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    if (E->e0) (void)hipEventDestroy(E->e0);
    if (E->e1) (void)hipEventDestroy(E->e1);
    hipStream_t s = E->stream;
    delete E;
    if (s) (void)hipStreamDestroy(s);
    return HBI_OK;
}

int hbi_ens_set_state(hbi_ens *E, const double *y, double t, double dt_next)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && y, "NULL argument");
        const int n = E->n;
        // unused realisation slots: a finite state (psi 0, tau 0, sigma 1)
        const double pad[3] = {0.0, 0.0, 1.0};
        hbi::upload_rm(E->y, y, E->nreal, n, 3, pad, E->stream);
        HBI_HIP(hipMemsetAsync(E->slip.p, 0, E->slip.n * 8, E->stream));
        hbi::EnsCtl c{};
        for (int r = 0; r < hbi::kR; r++) { c.h[r] = dt_next; c.t[r] = t; }
        hbi::to_device_sync(E->ctl.p, &c, sizeof(c), E->stream);
        // committed slip rate of the initial state
        hbi::EnsArrays A = hbi::e_arrays(E);
        hbi::ens_stage0_kernel<<<E->nblk, 256, 0, E->stream>>>(E->P, E->ctl.p, A);
        HBI_HIP(hipMemcpyAsync(E->velc.p, E->vel.p, E->vel.n * 8, hipMemcpyDeviceToDevice, E->stream));
        HBI_HIP(hipStreamSynchronize(E->stream));
    });
}

int hbi_ens_set_clock(hbi_ens *E, const double *t, const double *dt_next)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && t && dt_next, "NULL argument");
        for (int r = 0; r < E->nreal; r++)
            HBI_REQUIRE(std::isfinite(t[r]) && std::isfinite(dt_next[r]) && dt_next[r] > 0.0,
                        "t must be finite and dt_next finite and > 0");
        hbi::EnsCtl c;
        HBI_HIP(hipMemcpy(&c, E->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
        for (int r = 0; r < E->nreal; r++) { c.t[r] = t[r]; c.h[r] = dt_next[r]; }
        hbi::to_device_sync(E->ctl.p, &c, sizeof(c), E->stream);
    });
}

int hbi_ens_get_clock(hbi_ens *E, double *t, double *dt_next)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && t && dt_next, "NULL argument");
        hbi::EnsCtl c;
        HBI_HIP(hipMemcpy(&c, E->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
        for (int r = 0; r < E->nreal; r++) { t[r] = c.t[r]; dt_next[r] = c.h[r]; }
    });
}

int hbi_ens_get_state(hbi_ens *E, double *y, double *slip, double *vel, double *t, long long *accepted,
                      long long *rejected, int *status)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E, "NULL argument");
        const int n = E->n;
        if (y) hbi::download_rm(E->y, y, E->nreal, n, 3, E->stream);
        if (slip) hbi::download_rm(E->slip, slip, E->nreal, n, 1, E->stream);
        if (vel) hbi::download_rm(E->velc, vel, E->nreal, n, 1, E->stream);
        hbi::EnsCtl c;
        HBI_HIP(hipMemcpy(&c, E->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
        for (int r = 0; r < E->nreal; r++) {
            if (t) t[r] = c.t[r];
            if (accepted) accepted[r] = c.accepted[r];
            if (rejected) rejected[r] = c.rejected[r];
            if (status) status[r] = c.status[r];
        }
    });
}

int hbi_ens_step(hbi_ens *E, int nsteps, double *hist, long long *attempts)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && nsteps >= 0, "bad arguments");
        const size_t hrows = (size_t)E->nreal * nsteps;
        const size_t hcols = hbi::kHistCols + 2 * (size_t)E->P.nprobe;
        if (hist && (long long)hrows > E->hist_rows) {
            // the history buffer is a graph argument: a larger one needs a new capture
            E->hist.alloc(hrows * (hbi::kHistCols + 2 * hbi::kEnsProbes));
            E->hist_rows = (long long)hrows;
            if (E->graph) {
                HBI_HIP(hipGraphExecDestroy(E->graph));
                E->graph = nullptr;
            }
        }
        if (hist) HBI_HIP(hipMemsetAsync(E->hist.p, 0, hrows * hcols * sizeof(double), E->stream));
        if (!E->graph) {
            hipGraph_t gr;
            HBI_HIP(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
            hbi::e_attempt(E);
            HBI_HIP(hipStreamEndCapture(E->stream, &gr));
            HBI_HIP(hipGraphInstantiate(&E->graph, gr, nullptr, nullptr, 0));
            (void)hipGraphDestroy(gr);
        }
        if (E->P.grid) {
            // a replayed grid must cover the steps asked for
            hbi::EnsCtl c0;
            HBI_HIP(hipMemcpyAsync(&c0, E->ctl.p, sizeof(c0), hipMemcpyDeviceToHost, E->stream));
            HBI_HIP(hipStreamSynchronize(E->stream));
            for (int r = 0; r < E->nreal; r++)
                HBI_REQUIRE(c0.status[r] != 0 || c0.accepted[r] + nsteps <= E->P.gcap,
                            "the replayed step grid ends before these steps");
        }
        hbi::ens_arm_kernel<<<1, hbi::kR, 0, E->stream>>>(E->ctl.p, E->nreal, nsteps, hist ? nsteps : 0, E->P.grid,
                                                          E->P.gcap);
        HBI_HIP(hipEventRecord(E->e0, E->stream));
        hbi::EnsCtl c;
        long long launched = 0, budget = 50LL * nsteps + 100;
        int batch = nsteps;
        for (;;) {
            for (int q = 0; q < batch; q++) HBI_HIP(hipGraphLaunch(E->graph, E->stream));
            launched += batch;
            HBI_HIP(hipMemcpyAsync(&c, E->ctl.p, sizeof(c), hipMemcpyDeviceToHost, E->stream));
            HBI_HIP(hipStreamSynchronize(E->stream));
            if (c.done || launched >= budget || nsteps == 0) break;
            long long need = 0;
            for (int r = 0; r < E->nreal; r++)
                if (!c.frozen[r]) need = std::max(need, c.target[r] - c.accepted[r]);
            batch = (int)std::max<long long>(1, std::min<long long>(need + 2, budget - launched));
        }
        HBI_HIP(hipEventRecord(E->e1, E->stream));
        HBI_HIP(hipEventSynchronize(E->e1));
        float ms = 0;
        HBI_HIP(hipEventElapsedTime(&ms, E->e0, E->e1));
        E->last_ms = ms;
        if (attempts) *attempts = launched;
        if (hist) {
            E->hist.download(hist, hrows * hcols, E->stream);
            HBI_HIP(hipStreamSynchronize(E->stream));
        }
    });
}

int hbi_ens_last_ms(const hbi_ens *E, double *ms)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && ms, "NULL argument");
        *ms = E->last_ms;
    });
}

}  // extern "C"

namespace {
void drop_graph(hbi_ens *E)
{
    if (E->graph) {
        (void)hipGraphExecDestroy(E->graph);
        E->graph = nullptr;
    }
}
}  // namespace

extern "C" {

int hbi_ens_set_controller(hbi_ens *E, int mode, const hbi_mlp *m, double safety, double ratio_lo, double ratio_hi)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && (mode == 0 || mode == 1), "mode must be 0 (rkqs) or 1 (learned)");
        HBI_REQUIRE(mode == 0 || m, "learned mode needs an MLP");
        HBI_REQUIRE(mode == 0 || (safety > 0 && ratio_lo > 0 && ratio_hi >= ratio_lo), "bad proposal bounds");
        if (mode == 1) {
            E->mlp_p.alloc(hbi::kNParam);
            E->mlp_w.alloc(hbi::kNParam);
            E->mlp_mu.alloc(hbi::kIn);
            E->mlp_sd.alloc(hbi::kIn);
            HBI_HIP(hipStreamSynchronize(m->stream));
            HBI_HIP(hipMemcpyAsync(E->mlp_p.p, m->p.p, hbi::kNParam * sizeof(float), hipMemcpyDeviceToDevice, E->stream));
            HBI_HIP(hipMemcpyAsync(E->mlp_w.p, m->w.p, hbi::kNParam * sizeof(uint16_t), hipMemcpyDeviceToDevice, E->stream));
            HBI_HIP(hipMemcpyAsync(E->mlp_mu.p, m->mu.p, hbi::kIn * sizeof(float), hipMemcpyDeviceToDevice, E->stream));
            HBI_HIP(hipMemcpyAsync(E->mlp_sd.p, m->sd.p, hbi::kIn * sizeof(float), hipMemcpyDeviceToDevice, E->stream));
            HBI_HIP(hipStreamSynchronize(E->stream));
            E->P.safety = safety;
            E->P.ratio_lo = ratio_lo;
            E->P.ratio_hi = ratio_hi;
        }
        E->P.use_mlp = mode;
        E->P.learn = (mode == 1 || E->trace_cap > 0) ? 1 : 0;
        drop_graph(E);
    });
}

int hbi_ens_set_probes(hbi_ens *E, int nprobe, const int *idx)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && nprobe >= 0 && nprobe <= hbi::kEnsProbes && (nprobe == 0 || idx), "0 <= nprobe <= 16");
        for (int q = 0; q < nprobe; q++) HBI_REQUIRE(idx[q] >= 0 && idx[q] < E->n, "probe index out of range");
        E->P.nprobe = nprobe;
        for (int q = 0; q < nprobe; q++) E->P.probe[q] = idx[q];
        drop_graph(E);
    });
}

int hbi_ens_set_step_grid(hbi_ens *E, long long cap, const double *h)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && cap >= 0 && (cap == 0 || h), "bad arguments");
        if (cap > 0) {
            for (size_t q = 0; q < (size_t)E->nreal * cap; q++)
                HBI_REQUIRE(h[q] > 0.0 && std::isfinite(h[q]), "step sizes must be positive and finite");
            std::vector<double> t((size_t)hbi::kR * cap, 1.0);
            std::memcpy(t.data(), h, sizeof(double) * (size_t)E->nreal * cap);
            E->grid.upload(t, E->stream);
            E->P.grid = E->grid.p;
        } else {
            E->grid.release();
            E->P.grid = nullptr;
        }
        E->P.gcap = cap;
        drop_graph(E);
    });
}

int hbi_ens_set_trace(hbi_ens *E, int cap)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && cap >= 0, "bad arguments");
        if (cap > 0) E->trace.alloc((size_t)hbi::kR * cap * 14);
        else E->trace.release();
        E->trace_cap = cap;
        E->P.learn = (E->P.use_mlp || cap > 0) ? 1 : 0;
        hbi::EnsCtl c;
        HBI_HIP(hipMemcpy(&c, E->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
        for (int r = 0; r < hbi::kR; r++) c.trace_n[r] = 0;
        hbi::to_device_sync(E->ctl.p, &c, sizeof(c), E->stream);
        drop_graph(E);
    });
}

int hbi_ens_get_trace(hbi_ens *E, double *feat, double *raw, int *count)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(E && count, "NULL argument");
        HBI_HIP(hipStreamSynchronize(E->stream));
        hbi::EnsCtl c;
        HBI_HIP(hipMemcpy(&c, E->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
        const int cap = E->trace_cap;
        std::vector<double> t;
        if (cap > 0) {
            t.resize((size_t)hbi::kR * cap * 14);
            E->trace.download(t.data(), t.size(), E->stream);
            HBI_HIP(hipStreamSynchronize(E->stream));
        }
        for (int r = 0; r < E->nreal; r++) {
            const int n = std::min(c.trace_n[r], cap);
            count[r] = n;
            for (int i = 0; i < n; i++) {
                const double *row = t.data() + 14 * ((size_t)r * cap + i);
                if (feat)
                    for (int k = 0; k < hbi::kIn; k++) feat[((size_t)r * cap + i) * hbi::kIn + k] = row[k];
                if (raw) {
                    raw[((size_t)r * cap + i) * 2] = row[12];
                    raw[((size_t)r * cap + i) * 2 + 1] = row[13];
                }
            }
        }
    });
}

}  // extern "C"
