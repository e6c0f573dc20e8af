// rk.hip -- rate-and-state friction and Cash-Karp RK45 on the device.
//
// One RK attempt (rkqs -> rkck, main_LH.f90:2270-2370, 2472-2554) is a fixed
// sequence of kernels, captured once as a HIP graph and replayed:
//   stage0                 V, matvec input from y            derivs 2079-2109
//   6 x [ phase1, phase2 (H-matvec)  ->  stage kernel ]
//       stage kernel s: tile partials -> H x (/ds0) + taudot -> deriv ->
//       k_s, then either the next stage's input (y + h sum b_sj k_j, V, V-Vpl)
//       or, after the 6th stage, yout / yerr and the ordered error partials
//   decide                 errmax with the reference's guard, accept / shrink,
//                          hnext, underflow stop                      2309-2357
//   commit                 y <- yout, trapezoidal slip, V         1266-1287
//   monitor                max V, mean |slip|, history rows, stop criteria
// Every kernel first reads a device flag and returns when the batch target
// is reached or the run stopped, so surplus graph launches cost ~nothing and
// no host synchronisation happens inside a batch.  h, t and the counters
// live in device memory.
#include "rkcore.hpp"
#include "learn.hpp"

#include <algorithm>
#include <cmath>

namespace hbi {

struct RkPar {
    int n, law, backslip, limitsigma, has_hn, tr, tr_n, has_open;
    double vref, minsig, maxsig, dtmax, hdiv, eps, velmin, velmax, sg;
    Tab T;
    int learn, use_mlp;                   // step features on; MLP proposals on
    double safety, ratio_lo, ratio_hi;    // learned proposal: factor and bounds relative to rkqs
    // outputs (snapshots, event list): on, interval, buffer sizes, step offset, thresholds
    int out_on, interval, snap_cap, ev_cap;
    long long k0;
    double velth, dtout_inter, dtout_co;
};

struct RkCtl {
    double h, t, hdid, errmax, vmax, meanslip;
    long long attempts, accepted, rejected, target, hist_base, hist_cap;
    int nrjct, done, status, acc;
    // learned step size: the previous accepted step; trace rows written
    StepMem mem;
    int trace_n;
    // outputs (main_LH.f90:1339-1407): event detector state, the decisions
    // of the last accepted step (pending until the next attempt), buffer fill
    int slipping, hypoloc, out_pending, out_now, onset_now, end_now, snap_slot, ev_slot;
    int nsnap, nev, out_full;
    long long eventcount;
    double moment0, onset_time, tout, dtout;
};

struct RkArrays {
    const double *a, *b, *dc, *f0, *vpl, *taudot, *sigdot, *area;
    double *y, *ytmp, *yout, *yerr, *k, *vel, *x, *slip, *velc;
    const double *hx, *hx_n;    // H x of the shear (and normal) stress matrices
    const double *hxo, *hxo_n;  // opening: H_s2 xn, H_n2 xn
    double *xn, *veln, *slipn;  // opening rate (matvec input), post-step opening velocity and opening
    double *errpart, *monpart;
    double *hist, *vhist;
    const int *rec;
    int nrec;
    double *probe;        // per accepted step, kRkProbe values for each probed element
    const int *probe_idx;
    int nprobe;
    int *bins;            // histogram of log10 |V| over the elements after an accepted step
    double *trace;        // rows of 12 features + (h, errmax)
    int trace_cap;
    MlpDev mlp;
    // outputs: snapshots (slot x [vel, slip, sigma, tau] x n), their step and
    // time; slip at the event onset; finished events and their coseismic slip
    double *snap, *snap_t, *islip, *eqslip;
    long long *snap_k;
    hbi_rk_event *events;
};


constexpr int kRkThreads = 256;
// per-block monitor partials: max |V|, sum |slip| area, sum (tau / sigma) area, max sigma, min sigma,
// the index of the first max |V|
constexpr int kMon = 6;
// history row per accepted step (the quantities of main_LH.f90:1491):
// t, dt, errmax, rejected attempts before the step, log10 max |V|, mean |slip|,
// mean tau / sigma (area weighted), max sigma, min sigma
constexpr int kRkHist = 9;
// probe row per accepted step and probed element (the local output of
// main_LH.f90:1495-1499): V, slip, tau, sigma, mu = tau / sigma, psi, veln, slipn
// (veln and slipn 0 without opening)
constexpr int kRkProbe = 8;
constexpr int kRkMaxProbes = 16;


__global__ void rk_arm_kernel(RkCtl *c, long long nsteps, double h_override, int set_h)
{
    if (set_h) c->h = h_override;
    if (c->status == 0) c->done = 0;
    c->target = c->accepted + nsteps;
    c->hist_base = c->accepted;
    c->hist_cap = nsteps;
    c->acc = 0;
}

__global__ void rk_stage0_kernel(RkPar P, const RkCtl *__restrict__ c, RkArrays A, int from_tmp)
{
#pragma clang fp contract(off)
    if (c->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double *y = from_tmp ? A.ytmp : A.y;
    const double v = slip_rate(y[3 * i], y[3 * i + 1], y[3 * i + 2], A.a[i], P.vref);
    A.vel[i] = v;
    A.x[i] = P.backslip ? v - A.vpl[i] : v;
    // opening rate, main_LH.f90:2096-2100
    if (P.has_open) A.xn[i] = fmax(0.0, P.minsig - y[3 * i + 2]) * fabs(v);
}

// ordered (guard, value) pair of the reference's errmax loop: a guard element
// (ytemp - 1 == ytemp) overwrites errmax with 10, otherwise errmax = max(errmax, e)
struct ErrPair {
    int g;
    double v;
};
__device__ __forceinline__ ErrPair err_combine(ErrPair l, ErrPair r)
{
    if (r.g) return r;
    return ErrPair{l.g, fmax(l.v, r.v)};
}

// stage s in 0..5; mode 0: full attempt, 1: derivs only (writes k_0)
__global__ __launch_bounds__(kRkThreads) void rk_stage_kernel(RkPar P, const RkCtl *__restrict__ c, RkArrays A, int s,
                                                               int mode)
{
#pragma clang fp contract(off)
    __shared__ int sg[kRkThreads / kWave];
    __shared__ double sv[kRkThreads / kWave];
    if (c->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = P.n;
    const Tab &T = P.T;
    ErrPair ep{0, 0.0};
    if (i < n) {
        const double hx = A.hx[i];
        double gs = P.hdiv != 0.0 ? hx / P.hdiv : hx;
        double gn = 0.0;
        if (P.has_hn) {
            const double hy = A.hx_n[i];
            gn = P.hdiv != 0.0 ? hy / P.hdiv : hy;
        }
        if (P.has_open) {   // main_LH.f90:2129-2143
            gs = gs + (P.hdiv != 0.0 ? A.hxo[i] / P.hdiv : A.hxo[i]);
            gn = gn + (P.hdiv != 0.0 ? A.hxo_n[i] / P.hdiv : A.hxo_n[i]);
        }
        gn = gn + A.sigdot[i];
        gs = gs + A.taudot[i];
        const double *ys = (s == 0) ? A.y : A.ytmp;
        double kk[3];
        deriv(P, gs, gn, ys[3 * i], ys[3 * i + 1], ys[3 * i + 2], A.vel[i], A.a[i], A.b[i], A.dc[i], A.f0[i], kk);
        const size_t m3 = 3 * (size_t)n;
        double *ks = A.k + s * m3;
        ks[3 * i] = kk[0];
        ks[3 * i + 1] = kk[1];
        ks[3 * i + 2] = kk[2];
        if (mode == 0) {
            const double h = c->h;
            const double *k1 = A.k, *k2 = A.k + m3, *k3 = A.k + 2 * m3, *k4 = A.k + 3 * m3, *k5 = A.k + 4 * m3,
                         *k6 = A.k + 5 * m3;
            if (s < 5) {
                double yt[3];
                for (int q = 0; q < 3; q++) {
                    const size_t e = 3 * (size_t)i + q;
                    const double y = A.y[e];
                    switch (s) {
                    case 0: yt[q] = y + T.b21 * h * k1[e]; break;
                    case 1: yt[q] = y + h * (T.b31 * k1[e] + T.b32 * k2[e]); break;
                    case 2: yt[q] = y + h * (T.b41 * k1[e] + T.b42 * k2[e] + T.b43 * k3[e]); break;
                    case 3: yt[q] = y + h * (T.b51 * k1[e] + T.b52 * k2[e] + T.b53 * k3[e] + T.b54 * k4[e]); break;
                    default:
                        yt[q] = y + h * (T.b61 * k1[e] + T.b62 * k2[e] + T.b63 * k3[e] + T.b64 * k4[e] + T.b65 * k5[e]);
                        break;
                    }
                    A.ytmp[e] = yt[q];
                }
                const double v = slip_rate(yt[0], yt[1], yt[2], A.a[i], P.vref);
                A.vel[i] = v;
                A.x[i] = P.backslip ? v - A.vpl[i] : v;
                if (P.has_open) A.xn[i] = fmax(0.0, P.minsig - yt[2]) * fabs(v);
            } else {
                for (int q = 0; q < 3; q++) {
                    const size_t e = 3 * (size_t)i + q;
                    const double yo = A.y[e] + h * (T.c1 * k1[e] + T.c3 * k3[e] + T.c4 * k4[e] + T.c6 * k6[e]);
                    const double ye = h * (T.dc1 * k1[e] + T.dc3 * k3[e] + T.dc4 * k4[e] + T.dc5 * k5[e] + T.dc6 * k6[e]);
                    A.yout[e] = yo;
                    A.yerr[e] = ye;
                    const double er = fabs(ye / yo) / P.eps;
                    ErrPair cur{0, er > 0.0 ? er : 0.0};
                    if (er != er) cur.v = 0.0;
                    if (yo - 1 == yo) cur = ErrPair{1, 10.0};
                    ep = err_combine(ep, cur);
                }
            }
        }
    }
    if (mode != 0 || s != 5) return;
    // ordered block combine: lane l absorbs lane l+o (higher indices)
    for (int o = 1; o < kWave; o <<= 1) {
        const int g2 = __shfl_down(ep.g, o, kWave);
        const double v2 = __shfl_down(ep.v, o, kWave);
        const int lane = threadIdx.x % kWave;
        if ((lane & (2 * o - 1)) == 0 && lane + o < kWave) ep = err_combine(ep, ErrPair{g2, v2});
    }
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) { sg[w] = ep.g; sv[w] = ep.v; }
    __syncthreads();
    if (threadIdx.x == 0) {
        ErrPair b{sg[0], sv[0]};
        for (int q = 1; q < kRkThreads / kWave; q++) b = err_combine(b, ErrPair{sg[q], sv[q]});
        A.errpart[2 * blockIdx.x] = (double)b.g;
        A.errpart[2 * blockIdx.x + 1] = b.v;
    }
}

__global__ void rk_decide_kernel(RkPar P, RkCtl *c, const double *__restrict__ errpart, int nblk)
{
#pragma clang fp contract(off)
    if (c->done) return;
    const double SAFETY = (double)0.9f, PGROW = (double)-0.2f, PSHRNK = (double)-0.25f, tiny = (double)1e-20f;
    // the block pairs combined in block order: lane l takes a contiguous run of
    // blocks, then the lanes in order (err_combine is associative)
    const int lane = threadIdx.x, per = (nblk + 63) / 64;
    ErrPair e{0, 0.0};
    for (int b = lane * per; b < min(nblk, (lane + 1) * per); b++)
        e = err_combine(e, ErrPair{(int)errpart[2 * b], errpart[2 * b + 1]});
    for (int o = 1; o < 64; o <<= 1) {
        const int g2 = __shfl_down(e.g, o, 64);
        const double v2 = __shfl_down(e.v, o, 64);
        if ((lane & (2 * o - 1)) == 0 && lane + o < 64) e = err_combine(e, ErrPair{g2, v2});
    }
    if (lane != 0) return;
    c->out_pending = 0;
    const double errmax = e.v;
    const double h = c->h;
    c->attempts++;
    c->errmax = errmax;
    if (errmax < 1.0 && errmax > tiny) {
        double hn = fmin(2 * h, SAFETY * h * pow(errmax, PGROW));
        hn = fmin(hn, P.dtmax);
        c->hdid = h;
        c->t = c->t + h;
        c->h = hn;
        c->acc = 1;
        c->accepted++;
    } else {
        double hnew;
        if (errmax > tiny) hnew = fmax(0.5 * h, SAFETY * h * pow(errmax, PSHRNK));
        else hnew = 0.5 * h;
        c->acc = 0;
        c->nrjct++;
        c->rejected++;
        c->h = hnew;
        const double xnew = c->t + hnew;
        if (xnew - c->t < 1.e-15) {
            c->status = HBI_E_STEP;
            c->done = 1;
        }
    }
}

__global__ __launch_bounds__(kRkThreads) void rk_commit_kernel(RkPar P, const RkCtl *__restrict__ c, RkArrays A)
{
#pragma clang fp contract(off)
    __shared__ double sp[kRkThreads / kWave][kMon];
    __shared__ int lh[kBins];   // the block's log10 |V| histogram (learned step size)
    if (c->done || !c->acc) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double dt = c->hdid;
    if (P.learn) {
        for (int b = threadIdx.x; b < kBins; b += blockDim.x) lh[b] = 0;
        __syncthreads();
    }
    // max |V|, sum |slip| area, sum mu area, max sigma, min sigma, index of the
    // first max |V| (identities for i >= n; -1 < |V| so element 0 wins ties with it)
    double m[kMon] = {-1.0, 0.0, 0.0, -HUGE_VAL, HUGE_VAL, -1.0};
    if (i < P.n) {
        for (int q = 0; q < 3; q++) A.y[3 * (size_t)i + q] = A.yout[3 * (size_t)i + q];
        double sl = A.slip[i];
        sl = sl + A.velc[i] * dt * 0.5;
        const double v = slip_rate(A.y[3 * i], A.y[3 * i + 1], A.y[3 * i + 2], A.a[i], P.vref);
        const double sig = A.y[3 * i + 2];
        if (P.has_open) {   // main_LH.f90:1279-1283
            const double vn = v * (fmax(0.0, P.minsig - sig) + fmin(0.0, P.maxsig - sig));
            A.veln[i] = vn;
            A.slipn[i] = A.slipn[i] + vn * dt;
        }
        sl = sl + v * dt * 0.5;
        A.slip[i] = sl;
        A.velc[i] = v;
        m[0] = fabs(v);
        m[1] = fabs(sl) * A.area[i];
        m[2] = A.y[3 * i + 1] / sig * A.area[i];
        m[3] = sig;
        m[4] = sig;
        m[5] = (double)i;
        if (P.learn) atomicAdd(&lh[vbin(m[0])], 1);
    }
    if (P.learn) {   // most elements share a few bins: one global add per nonzero bin and block
        __syncthreads();
        for (int b = threadIdx.x; b < kBins; b += blockDim.x)
            if (lh[b]) atomicAdd(&A.bins[b], lh[b]);
    }
    // per-block combine in a fixed order (sums ordered by lane, then by wave)
    for (int o = 1; o < kWave; o <<= 1) {
        double o2[kMon];
        for (int k = 0; k < kMon; k++) o2[k] = __shfl_down(m[k], o, kWave);
        const int lane = threadIdx.x % kWave;
        if ((lane & (2 * o - 1)) == 0 && lane + o < kWave) {
            // the higher lanes hold higher indices: take theirs only when larger
            if (o2[0] > m[0]) { m[0] = o2[0]; m[5] = o2[5]; }
            m[1] = m[1] + o2[1];
            m[2] = m[2] + o2[2];
            m[3] = fmax(m[3], o2[3]);
            m[4] = fmin(m[4], o2[4]);
        }
    }
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0)
        for (int k = 0; k < kMon; k++) sp[w][k] = m[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[kMon];
        for (int k = 0; k < kMon; k++) t[k] = sp[0][k];
        for (int q = 1; q < kRkThreads / kWave; q++) {
            if (sp[q][0] > t[0]) { t[0] = sp[q][0]; t[5] = sp[q][5]; }
            t[1] = t[1] + sp[q][1];
            t[2] = t[2] + sp[q][2];
            t[3] = fmax(t[3], sp[q][3]);
            t[4] = fmin(t[4], sp[q][4]);
        }
        for (int k = 0; k < kMon; k++) A.monpart[kMon * (size_t)blockIdx.x + k] = t[k];
    }
}

// After an accepted step: the 12 step features (slip-rate quantiles from the
// histogram, their changes, h and errmax of this and the previous step, the
// rejection flag, the error-optimal step h e^(-1/5)); optional trace row;
// with the MLP on, the next h replaces the rkqs proposal (within bounds).
// Runs on one workgroup of kHid threads; clears the histogram.
__global__ __launch_bounds__(kHid) void rk_learn_kernel(RkPar P, RkCtl *c, RkArrays A)
{
    __shared__ int part[kHid + 1], hb[kBins];
    __shared__ float x[kIn], h1[kHid], red[kHid];
    __shared__ double q[3];
    if (c->done || !c->acc || !P.learn) return;
    const int j = threadIdx.x;
    MlpRow R;
    if (P.use_mlp) R = mlp_load_row(A.mlp);
    step_features(A.bins, P.n, c->hdid, c->errmax, c->nrjct > 0, c->mem, part, hb, x, q);
    if (j == 0 && A.trace && c->trace_n < A.trace_cap) {
        double *row = A.trace + 14 * (size_t)c->trace_n;
        for (int k = 0; k < kIn; k++) row[k] = (double)x[k];
        row[12] = c->hdid;
        row[13] = c->errmax;
        c->trace_n++;
    }
    if (P.use_mlp) {
        const float out = mlp_infer_block(A.mlp, R, x, h1, red);
        if (j == 0) c->h = learned_h(x[11], out, c->h, P.safety, P.ratio_lo, P.ratio_hi, P.dtmax);
    }
    if (j == 0) {
        for (int k = 0; k < 3; k++) c->mem.prev_q[k] = q[k];
        c->mem.prev_h = c->hdid;
        c->mem.prev_e = c->errmax;
        c->mem.have_prev = 1;
    }
    for (int b = 0; b < kBins / kHid; b++) A.bins[j * (kBins / kHid) + b] = 0;
}

__global__ void rk_monitor_kernel(RkPar P, RkCtl *c, RkArrays A, int nblk)
{
#pragma clang fp contract(off)
    if (c->done || !c->acc) return;
    __shared__ double vm, msum, mmu, smax, smin;
    // max |V| with its first element and the sigma extremes across the
    // workgroup (exact in any order; ties go to the first block), the two
    // sums in block order by thread 0 from LDS, kMonChunk blocks at a time
    constexpr int kMonChunk = 512, kMW = 256 / kWave;
    __shared__ double s1[kMonChunk], s2[kMonChunk], wv[kMW], wl[kMW], wh[kMW], wo[kMW];
    __shared__ int wb[kMW];
    const int tid = threadIdx.x, lane = tid % kWave, wv_i = tid / kWave;
    double pa = -1.0, ploc = -1.0, phi_ = -HUGE_VAL, plo = HUGE_VAL;
    int pblk = nblk;
    for (int q = tid; q < nblk; q += blockDim.x) {
        const double *mp = A.monpart + kMon * (size_t)q;
        if (mp[0] > pa) { pa = mp[0]; ploc = mp[5]; pblk = q; }
        phi_ = fmax(phi_, mp[3]);
        plo = fmin(plo, mp[4]);
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double a2 = __shfl_xor(pa, o, kWave), l2 = __shfl_xor(ploc, o, kWave);
        const int b2 = __shfl_xor(pblk, o, kWave);
        if (a2 > pa || (a2 == pa && b2 < pblk)) { pa = a2; ploc = l2; pblk = b2; }
        phi_ = fmax(phi_, __shfl_xor(phi_, o, kWave));
        plo = fmin(plo, __shfl_xor(plo, o, kWave));
    }
    if (lane == 0) { wv[wv_i] = pa; wl[wv_i] = ploc; wb[wv_i] = pblk; wh[wv_i] = phi_; wo[wv_i] = plo; }
    double bsum = 0.0, usum = 0.0;
    for (int q0 = 0; q0 < nblk; q0 += kMonChunk) {
        const int m = min(kMonChunk, nblk - q0);
        __syncthreads();
        for (int q = tid; q < m; q += blockDim.x) {
            s1[q] = A.monpart[kMon * (size_t)(q0 + q) + 1];
            s2[q] = A.monpart[kMon * (size_t)(q0 + q) + 2];
        }
        __syncthreads();
        if (tid == 0)
            for (int q = 0; q < m; q++) {
                bsum = bsum + s1[q];
                usum = usum + s2[q];
            }
    }
    if (threadIdx.x == 0) {
        double a = wv[0], loc = wl[0], hi = wh[0], lo = wo[0];
        int ab = wb[0];
        for (int w = 1; w < kMW; w++) {
            if (wv[w] > a || (wv[w] == a && wb[w] < ab)) { a = wv[w]; loc = wl[w]; ab = wb[w]; }
            hi = fmax(hi, wh[w]);
            lo = fmin(lo, wo[w]);
        }
        const double b = bsum, u = usum;
        vm = a;
        msum = b / P.sg;
        mmu = u / P.sg;
        smax = hi;
        smin = lo;
        c->vmax = a;
        c->meanslip = msum;
        if (P.out_on) {
            // main_LH.f90:1339-1407: event start / end, snapshot decision
            const double t = c->t;
            const long long k = P.k0 + c->accepted;
            int out = 0, onset = 0, end = 0;
            if (!c->slipping && a > P.velth) {
                out = 1; onset = 1;
                c->slipping = 1;
                c->dtout = P.dtout_co;
                c->tout = t + c->dtout;
                c->eventcount++;
                c->moment0 = msum;
                c->hypoloc = (int)loc + 1;
                c->onset_time = t;
            }
            if (c->slipping && a < 0.5 * P.velth) {
                out = 1; end = 1;
                c->slipping = 0;
                c->dtout = P.dtout_inter;
                c->tout = t + c->dtout;
                const double moment = msum - c->moment0;
                hbi_rk_event ev{};
                ev.count = c->eventcount; ev.k = k; ev.onset_time = c->onset_time;
                ev.mw = (log10(moment * kRigid * P.sg) + 5.9) / 1.5;
                ev.moment = moment;
                ev.hypoloc = c->hypoloc;
                c->ev_slot = c->nev;
                A.events[c->nev++] = ev;
            }
            if (P.interval > 0 && k % P.interval == 0) out = 1;
            if (t > c->tout) { out = 1; c->tout = t + c->dtout; }
            c->out_now = out; c->onset_now = onset; c->end_now = end;
            c->out_pending = out | onset | end;
            if (out) {
                c->snap_slot = c->nsnap;
                A.snap_k[c->nsnap] = k;
                A.snap_t[c->nsnap] = t;
                c->nsnap++;
            }
            if (c->nsnap >= P.snap_cap || c->nev >= P.ev_cap) c->out_full = 1;
        }
    }
    __syncthreads();
    const long long row = c->accepted - c->hist_base - 1;
    if (row >= 0 && row < c->hist_cap) {
        if (A.hist && threadIdx.x == 0) {
            double *hr = A.hist + kRkHist * row;
            hr[0] = c->t; hr[1] = c->hdid; hr[2] = c->errmax; hr[3] = (double)c->nrjct; hr[4] = log10(vm);
            hr[5] = msum; hr[6] = mmu; hr[7] = smax; hr[8] = smin;
        }
        if (A.vhist)
            for (int r = threadIdx.x; r < A.nrec; r += blockDim.x) A.vhist[row * A.nrec + r] = A.velc[A.rec[r]];
        if (A.probe)
            for (int q = threadIdx.x; q < A.nprobe; q += blockDim.x) {
                const size_t i = A.probe_idx[q];
                const double tau = A.y[3 * i + 1], sig = A.y[3 * i + 2];
                double *pr = A.probe + ((size_t)row * A.nprobe + q) * kRkProbe;
                pr[0] = A.velc[i]; pr[1] = A.slip[i]; pr[2] = tau; pr[3] = sig; pr[4] = tau / sig;
                pr[5] = A.y[3 * i];
                pr[6] = P.has_open ? A.veln[i] : 0.0;
                pr[7] = P.has_open ? A.slipn[i] : 0.0;
            }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        c->nrjct = 0;
        if (vm > P.velmax || vm < P.velmin) { c->status = 1; c->done = 1; }
        if (c->accepted >= c->target || c->out_full) c->done = 1;
    }
}

// The output decisions of the last accepted step: snapshot of vel, slip,
// sigma, tau into its slot; slip kept at an event onset; coseismic slip
// |slip - slip at onset| at an event end.  Rewriting the same values is
// harmless, so the kernel repeats them until the next attempt clears the flag.
__global__ void rk_output_kernel(RkPar P, const RkCtl *__restrict__ c, RkArrays A)
{
    if (!c->out_pending) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const size_t n = P.n;
    const double sl = A.slip[i];
    if (c->out_now) {
        const int nf = P.has_open ? 6 : 4;
        double *s = A.snap + (size_t)c->snap_slot * nf * n;
        s[i] = A.velc[i];
        s[n + i] = sl;
        s[2 * n + i] = A.y[3 * (size_t)i + 2];
        s[3 * n + i] = A.y[3 * (size_t)i + 1];
        if (P.has_open) {
            s[4 * n + i] = A.veln[i];
            s[5 * n + i] = A.slipn[i];
        }
    }
    if (c->onset_now) A.islip[i] = sl;
    if (c->end_now) A.eqslip[(size_t)c->ev_slot * n + i] = fabs(sl - A.islip[i]);
}

}  // namespace hbi

// ------------------------------------------------------------------ object --
struct hbi_rk {
    hipStream_t stream = nullptr;
    hbi_hmat *hs = nullptr, *hn = nullptr;
    hbi_hmat *hs2 = nullptr, *hn2 = nullptr;   // opening matrices (hbi_rk_set_opening)
    hbi_rk_opts opts{};
    hbi::RkPar P{};
    int n = 0, nblk = 0;
    hbi::DevBuf<double> a, b, dc, f0, vpl, taudot, sigdot, area;
    hbi::DevBuf<double> y, ytmp, yout, yerr, k, vel, x, slip, velc, z, zn, ypart, ypart_n, hx, hx_n, errpart, monpart;
    hbi::DevBuf<double> hist, vhist;
    hbi::DevBuf<int> rec;
    hbi::DevBuf<double> probe;                 // probe rows of the last step call (hbi_rk_set_probes)
    hbi::DevBuf<int> probe_idx;
    long long probe_rows = 0;
    hbi::DevBuf<hbi::RkCtl> ctl;
    hbi::DevBuf<int> bins;
    hbi::DevBuf<double> trace;
    int trace_cap = 0;
    hbi::DevBuf<float> mlp_p, mlp_mu, mlp_sd;   // copy of the step-size MLP
    hbi::DevBuf<uint16_t> mlp_w;
    hbi::DevBuf<double> xn, zo, zo_n, ypart_o, ypart_on, hxo, hxo_n, veln, slipn;   // opening
    hbi::DevBuf<double> snap, snap_t, islip, eqslip;   // outputs (hbi_rk_set_output)
    hbi::DevBuf<long long> snap_k;
    hbi::DevBuf<hbi_rk_event> events;
    bool out_full = false;                     // output held on the device fills its buffers
    bool dirty = false;                        // graph arguments changed: recapture
    hipGraphExec_t graph = nullptr;
    int graph_nrec = -1;
    double *graph_hist = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
};

namespace hbi {

static RkArrays arrays(hbi_rk *r)
{
    RkArrays A{};
    A.a = r->a.p; A.b = r->b.p; A.dc = r->dc.p; A.f0 = r->f0.p; A.vpl = r->vpl.p; A.taudot = r->taudot.p;
    A.sigdot = r->sigdot.p; A.area = r->area.p;
    A.y = r->y.p; A.ytmp = r->ytmp.p; A.yout = r->yout.p; A.yerr = r->yerr.p; A.k = r->k.p; A.vel = r->vel.p;
    A.x = r->x.p; A.slip = r->slip.p; A.velc = r->velc.p;
    A.hx = r->hx.p; A.hx_n = r->hx_n.p;
    A.hxo = r->hxo.p; A.hxo_n = r->hxo_n.p; A.xn = r->xn.p; A.veln = r->veln.p; A.slipn = r->slipn.p;
    A.errpart = r->errpart.p; A.monpart = r->monpart.p;
    A.bins = r->bins.p;
    A.trace = r->trace.p;
    A.trace_cap = r->trace_cap;
    A.mlp = MlpDev{r->mlp_w.p, r->mlp_p.p, r->mlp_mu.p, r->mlp_sd.p};
    A.snap = r->snap.p; A.snap_t = r->snap_t.p; A.snap_k = r->snap_k.p; A.islip = r->islip.p;
    A.eqslip = r->eqslip.p; A.events = r->events.p;
    A.hist = r->hist.n ? r->hist.p : nullptr;
    A.vhist = r->vhist.n ? r->vhist.p : nullptr;
    A.rec = r->rec.p;
    A.nrec = (int)r->rec.n;
    A.probe = r->probe.n ? r->probe.p : nullptr;
    A.probe_idx = r->probe_idx.p;
    A.nprobe = (int)r->probe_idx.n;
    return A;
}

// the six derivs evaluations of one attempt (or one, for derivs only)
static void launch_stages(hbi_rk *r, int nstages, int mode)
{
    const int g = (r->n + kRkThreads - 1) / kRkThreads;
    RkArrays A = arrays(r);
    const int *gate = &r->ctl.p->done;
    rk_stage0_kernel<<<g, kRkThreads, 0, r->stream>>>(r->P, r->ctl.p, A, 0);
    for (int s = 0; s < nstages; s++) {
        hmat_phase1(r->hs, r->opts.prec, r->x.p, r->z.p, r->ypart.p, r->stream, gate);
        hmat_phase2(r->hs, r->opts.prec, r->x.p, r->z.p, r->ypart.p, r->hx.p, r->stream, gate);
        if (r->hn) {
            hmat_phase1(r->hn, r->opts.prec, r->x.p, r->zn.p, r->ypart_n.p, r->stream, gate);
            hmat_phase2(r->hn, r->opts.prec, r->x.p, r->zn.p, r->ypart_n.p, r->hx_n.p, r->stream, gate);
        }
        if (r->P.has_open) {
            hmat_phase1(r->hs2, r->opts.prec, r->xn.p, r->zo.p, r->ypart_o.p, r->stream, gate);
            hmat_phase2(r->hs2, r->opts.prec, r->xn.p, r->zo.p, r->ypart_o.p, r->hxo.p, r->stream, gate);
            hmat_phase1(r->hn2, r->opts.prec, r->xn.p, r->zo_n.p, r->ypart_on.p, r->stream, gate);
            hmat_phase2(r->hn2, r->opts.prec, r->xn.p, r->zo_n.p, r->ypart_on.p, r->hxo_n.p, r->stream, gate);
        }
        rk_stage_kernel<<<g, kRkThreads, 0, r->stream>>>(r->P, r->ctl.p, A, s, mode);
    }
}

static void launch_attempt(hbi_rk *r)
{
    const int g = r->nblk;
    launch_stages(r, 6, 0);
    rk_decide_kernel<<<1, 64, 0, r->stream>>>(r->P, r->ctl.p, r->errpart.p, g);
    RkArrays A = arrays(r);
    rk_commit_kernel<<<g, kRkThreads, 0, r->stream>>>(r->P, r->ctl.p, A);
    if (r->P.learn) rk_learn_kernel<<<1, kHid, 0, r->stream>>>(r->P, r->ctl.p, A);
    rk_monitor_kernel<<<1, 256, 0, r->stream>>>(r->P, r->ctl.p, A, g);
    if (r->P.out_on) rk_output_kernel<<<g, kRkThreads, 0, r->stream>>>(r->P, r->ctl.p, A);
}

static void capture(hbi_rk *r)
{
    if (r->graph) {
        (void)hipGraphExecDestroy(r->graph);
        r->graph = nullptr;
    }
    hipGraph_t gr;
    HBI_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    launch_attempt(r);
    HBI_HIP(hipStreamEndCapture(r->stream, &gr));
    HBI_HIP(hipGraphInstantiate(&r->graph, gr, nullptr, nullptr, 0));
    (void)hipGraphDestroy(gr);
    r->graph_nrec = (int)r->rec.n;
    r->graph_hist = r->hist.p;
}

}  // namespace hbi

extern "C" {

int hbi_rk_create(hbi_hmat *hs, hbi_hmat *hn, const hbi_rk_opts *opts, const double *a, const double *b,
                  const double *dc, const double *f0, const double *vpl, const double *taudot, const double *sigdot,
                  const double *area, hbi_rk **out)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(hs && opts && out && a && b && dc && f0 && vpl && taudot && sigdot && area, "NULL argument");
        HBI_REQUIRE(!hn || hn->n == hs->n, "normal-stress matrix size differs");
        HBI_REQUIRE(opts->prec >= 0 && opts->prec <= 2, "bad precision");
        HBI_REQUIRE(opts->eps_r > 0, "eps_r must be > 0");
        *out = nullptr;
        hbi_rk *r = new hbi_rk();
        try {
            r->hs = hs;
            r->hn = hn;
            r->opts = *opts;
            const int n = hs->n;
            r->n = n;
            r->nblk = (n + hbi::kRkThreads - 1) / hbi::kRkThreads;
            HBI_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
            HBI_HIP(hipEventCreate(&r->e0));
            HBI_HIP(hipEventCreate(&r->e1));
            hbi::RkPar &P = r->P;
            P.n = n; P.law = opts->law; P.backslip = opts->backslip; P.limitsigma = opts->limitsigma;
            P.has_hn = hn ? 1 : 0; P.tr = hs->tr; P.tr_n = hn ? hn->tr : 0;
            P.vref = opts->vref; P.minsig = opts->minsig; P.maxsig = opts->maxsig; P.dtmax = opts->dtmax;
            P.hdiv = opts->hdiv; P.eps = opts->eps_r; P.velmin = opts->velmin; P.velmax = opts->velmax;
            P.T = hbi::rk_tableau();
            double sgsum = 0.0;
            for (int i = 0; i < n; i++) sgsum = sgsum + area[i];
            P.sg = sgsum;
            r->a.upload(a, n, r->stream); r->b.upload(b, n, r->stream); r->dc.upload(dc, n, r->stream);
            r->f0.upload(f0, n, r->stream); r->vpl.upload(vpl, n, r->stream); r->taudot.upload(taudot, n, r->stream);
            r->sigdot.upload(sigdot, n, r->stream); r->area.upload(area, n, r->stream);
            const size_t m3 = 3 * (size_t)n;
            r->y.alloc(m3); r->ytmp.alloc(m3); r->yout.alloc(m3); r->yerr.alloc(m3); r->k.alloc(6 * m3);
            r->vel.alloc(n); r->x.alloc(n + 8); r->slip.alloc(n); r->velc.alloc(n);
            HBI_HIP(hipMemsetAsync(r->x.p, 0, (n + 8) * 8, r->stream));
            HBI_HIP(hipMemsetAsync(r->slip.p, 0, n * 8, r->stream));
            HBI_HIP(hipMemsetAsync(r->velc.p, 0, n * 8, r->stream));
            HBI_HIP(hipMemsetAsync(r->k.p, 0, 6 * m3 * 8, r->stream));
            r->z.alloc(hbi::hmat_z_size(hs, opts->prec));
            r->ypart.alloc(hbi::hmat_ypart_size(hs, opts->prec));
            r->hx.alloc(n);
            if (hn) {
                r->zn.alloc(hbi::hmat_z_size(hn, opts->prec));
                r->ypart_n.alloc(hbi::hmat_ypart_size(hn, opts->prec));
                r->hx_n.alloc(n);
            }
            r->errpart.alloc(2 * (size_t)r->nblk);
            r->monpart.alloc(hbi::kMon * (size_t)r->nblk);
            r->ctl.alloc(1);
            r->bins.alloc(hbi::kBins);
            HBI_HIP(hipMemsetAsync(r->bins.p, 0, hbi::kBins * sizeof(int), r->stream));
            r->P.safety = 0.9;
            r->P.ratio_lo = 0.125;
            r->P.ratio_hi = 8.0;
            hbi::RkCtl c{};
            c.h = 1.0;
            hbi::to_device_sync(r->ctl.p, &c, sizeof(c), r->stream);
            HBI_HIP(hipStreamSynchronize(r->stream));
        } catch (...) {
            hbi_rk_destroy(r);
            throw;
        }
        *out = r;
    });
}

int hbi_rk_destroy(hbi_rk *r)
{
    if (!r) return HBI_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    if (r->graph) (void)hipGraphExecDestroy(r->graph);
    if (r->e0) (void)hipEventDestroy(r->e0);
    if (r->e1) (void)hipEventDestroy(r->e1);
    hipStream_t s = r->stream;
    delete r;
    if (s) (void)hipStreamDestroy(s);
    return HBI_OK;
}

int hbi_rk_set_state(hbi_rk *r, const double *y, const double *slip, const double *vel, double t, double dt_next)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && y, "NULL argument");
        const int n = r->n;
        r->y.upload(y, 3 * (size_t)n, r->stream);
        if (slip) r->slip.upload(slip, n, r->stream);
        else HBI_HIP(hipMemsetAsync(r->slip.p, 0, n * 8, r->stream));
        if (vel) {
            r->velc.upload(vel, n, r->stream);
        } else {
            hbi::RkArrays A = hbi::arrays(r);
            // V of the state, written into velc through the stage-0 kernel
            hbi::RkCtl c{};
            hbi::to_device_sync(r->ctl.p, &c, sizeof(c), r->stream);
            hbi::rk_stage0_kernel<<<r->nblk, hbi::kRkThreads, 0, r->stream>>>(r->P, r->ctl.p, A, 0);
            HBI_HIP(hipMemcpyAsync(r->velc.p, r->vel.p, n * 8, hipMemcpyDeviceToDevice, r->stream));
        }
        hbi::RkCtl c{};
        c.h = dt_next;
        c.t = t;
        if (r->P.has_open) {   // veln = slipn = 0 (main_LH.f90:967, 1004)
            HBI_HIP(hipMemsetAsync(r->veln.p, 0, n * 8, r->stream));
            HBI_HIP(hipMemsetAsync(r->slipn.p, 0, n * 8, r->stream));
        }
        // a fresh event detector (main_LH.f90:1188-1195); undrained output is dropped
        c.dtout = r->P.dtout_inter;
        c.tout = r->P.dtout_inter;
        r->out_full = false;
        hbi::to_device_sync(r->ctl.p, &c, sizeof(c), r->stream);
        HBI_HIP(hipStreamSynchronize(r->stream));
    });
}

int hbi_rk_get_state(hbi_rk *r, double *y, double *slip, double *vel, hbi_rk_info *info)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r, "NULL argument");
        const int n = r->n;
        if (y) r->y.download(y, 3 * (size_t)n, r->stream);
        if (slip) r->slip.download(slip, n, r->stream);
        if (vel) r->velc.download(vel, n, r->stream);
        hbi::RkCtl c;
        HBI_HIP(hipMemcpyAsync(&c, r->ctl.p, sizeof(c), hipMemcpyDeviceToHost, r->stream));
        HBI_HIP(hipStreamSynchronize(r->stream));
        if (info) {
            info->t = c.t; info->dt_next = c.h; info->dt_last = c.hdid; info->errmax_last = c.errmax;
            info->attempts = c.attempts; info->accepted = c.accepted; info->rejected = c.rejected;
            info->vmax = c.vmax; info->status = c.status;
        }
    });
}

int hbi_rk_derivs(hbi_rk *r, const double *y, double *dydx)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && y && dydx, "NULL argument");
        const int n = r->n;
        hbi::to_device_sync(r->y.p, y, 3 * (size_t)n * 8, r->stream);
        hbi::rk_arm_kernel<<<1, 1, 0, r->stream>>>(r->ctl.p, 0, 0.0, 0);
        hbi::launch_stages(r, 1, 1);
        HBI_HIP(hipGetLastError());
        HBI_HIP(hipMemcpyAsync(dydx, r->k.p, 3 * (size_t)n * 8, hipMemcpyDeviceToHost, r->stream));
        HBI_HIP(hipStreamSynchronize(r->stream));
    });
}

int hbi_rk_rkck(hbi_rk *r, const double *y, double h, double *yout, double *yerr)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && y && yout && yerr, "NULL argument");
        const int n = r->n;
        hbi::to_device_sync(r->y.p, y, 3 * (size_t)n * 8, r->stream);
        hbi::rk_arm_kernel<<<1, 1, 0, r->stream>>>(r->ctl.p, 0, h, 1);
        hbi::launch_stages(r, 6, 0);
        HBI_HIP(hipGetLastError());
        HBI_HIP(hipMemcpyAsync(yout, r->yout.p, 3 * (size_t)n * 8, hipMemcpyDeviceToHost, r->stream));
        HBI_HIP(hipMemcpyAsync(yerr, r->yerr.p, 3 * (size_t)n * 8, hipMemcpyDeviceToHost, r->stream));
        HBI_HIP(hipStreamSynchronize(r->stream));
    });
}

}  // extern "C"

namespace hbi {

// history buffers (the graph records their addresses), recapture when any
// graph argument changed, arm the step target, start the timing event
static void step_prepare(hbi_rk *r, int nsteps, bool hist, int nrec, const int *rec)
{
    HBI_REQUIRE(nsteps >= 0, "bad arguments");
    if (r->out_full) throw Error{HBI_E_STATE, "output buffers full: drain them with hbi_rk_get_output"};
    for (int q = 0; q < nrec; q++) HBI_REQUIRE(rec[q] >= 0 && rec[q] < r->n, "record index out of range");
    const size_t hneed = hist ? kRkHist * (size_t)nsteps : 0;
    const size_t vneed = nrec > 0 ? (size_t)nsteps * nrec : 0;
    bool recapture = !r->graph;
    if (hneed > r->hist.n || (hneed == 0 && r->hist.n)) {
        if (hneed) r->hist.alloc(hneed); else r->hist.release();
        recapture = true;
    }
    if (vneed > r->vhist.n || (vneed == 0 && r->vhist.n)) {
        if (vneed) r->vhist.alloc(vneed); else r->vhist.release();
        recapture = true;
    }
    if (nrec != (int)r->rec.n) {
        if (nrec) r->rec.alloc(nrec); else r->rec.release();
        recapture = true;
    }
    if (nrec) r->rec.upload(rec, nrec, r->stream);
    const size_t pneed = (size_t)nsteps * r->probe_idx.n * kRkProbe;
    if (pneed > r->probe.n || (pneed == 0 && r->probe.n)) {
        if (pneed) r->probe.alloc(pneed); else r->probe.release();
        recapture = true;
    }
    r->probe_rows = 0;
    if (r->dirty) recapture = true;
    if (recapture) capture(r);
    r->dirty = false;
    rk_arm_kernel<<<1, 1, 0, r->stream>>>(r->ctl.p, nsteps, 0.0, 0);
    HBI_HIP(hipEventRecord(r->e0, r->stream));
}

static void step_finish(hbi_rk *r, const RkCtl &c, hbi_rk_info *info)
{
    HBI_HIP(hipEventRecord(r->e1, r->stream));
    HBI_HIP(hipEventSynchronize(r->e1));
    float ms = 0;
    HBI_HIP(hipEventElapsedTime(&ms, r->e0, r->e1));
    r->last_ms = ms;
    r->out_full = c.out_full != 0;
    r->probe_rows = std::max<long long>(0, std::min(c.accepted - c.hist_base, c.hist_cap));
    if (info) {
        info->t = c.t; info->dt_next = c.h; info->dt_last = c.hdid; info->errmax_last = c.errmax;
        info->attempts = c.attempts; info->accepted = c.accepted; info->rejected = c.rejected;
        info->vmax = c.vmax; info->status = c.status ? c.status : (c.out_full ? HBI_RK_OUTPUT_FULL : 0);
    }
}

}  // namespace hbi

extern "C" {

int hbi_rk_step(hbi_rk *r, int nsteps, double *hist, int nrec, const int *rec, double *vhist, hbi_rk_info *info)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && nsteps >= 0, "bad arguments");
        HBI_REQUIRE(nrec == 0 || (rec && vhist), "rec and vhist required when nrec > 0");
        hbi::step_prepare(r, nsteps, hist != nullptr, nrec, rec);
        hbi::RkCtl c;
        long long launched = 0, budget = 50LL * nsteps + 100;
        int batch = nsteps;
        for (;;) {
            for (int q = 0; q < batch; q++) HBI_HIP(hipGraphLaunch(r->graph, r->stream));
            launched += batch;
            HBI_HIP(hipMemcpyAsync(&c, r->ctl.p, sizeof(c), hipMemcpyDeviceToHost, r->stream));
            HBI_HIP(hipStreamSynchronize(r->stream));
            if (c.done || launched >= budget || nsteps == 0) break;
            batch = (int)std::max<long long>(1, std::min<long long>(c.target - c.accepted + 2, budget - launched));
        }
        const long long done = c.accepted - c.hist_base;
        if (hist && done > 0)
            r->hist.download(hist, hbi::kRkHist * (size_t)std::min<long long>(done, nsteps), r->stream);
        if (nrec && done > 0)
            r->vhist.download(vhist, (size_t)std::min<long long>(done, nsteps) * nrec, r->stream);
        HBI_HIP(hipStreamSynchronize(r->stream));
        hbi::step_finish(r, c, info);
        if (c.status == HBI_E_STEP) throw hbi::Error{HBI_E_STEP, "Runge-Kutta step size underflow"};
    });
}

int hbi_rk_step_many(hbi_rk *const *rs, int m, int nsteps, hbi_rk_info *infos)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(rs && m > 0 && nsteps >= 0, "bad arguments");
        for (int f = 0; f < m; f++) HBI_REQUIRE(rs[f], "NULL run");
        for (int f = 0; f < m; f++)
            for (int g = f + 1; g < m; g++) HBI_REQUIRE(rs[f] != rs[g], "a run appears twice");
        for (int f = 0; f < m; f++) hbi::step_prepare(rs[f], nsteps, false, 0, nullptr);
        for (int f = 0; f < m; f++) HBI_HIP(hipStreamSynchronize(rs[f]->stream));
        // One graph for one attempt of every run: a branch per run (its own
        // attempt graph's kernels, on its stream), forked from and joined to a
        // launcher stream, so each lockstep round is a single graph launch and
        // the device runs the faults side by side.  A run that reached its
        // target is gated off inside its kernels (its control block's done flag).
        hipStream_t s0 = nullptr;
        hipEvent_t fork = nullptr;
        std::vector<hipEvent_t> join(m, nullptr);
        hipGraphExec_t all = nullptr;
        auto cleanup = [&] {
            if (all) (void)hipGraphExecDestroy(all);
            for (hipEvent_t e : join)
                if (e) (void)hipEventDestroy(e);
            if (fork) (void)hipEventDestroy(fork);
            if (s0) (void)hipStreamDestroy(s0);
        };
        try {
            HBI_HIP(hipStreamCreateWithFlags(&s0, hipStreamNonBlocking));
            HBI_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
            for (int f = 0; f < m; f++) HBI_HIP(hipEventCreateWithFlags(&join[f], hipEventDisableTiming));
            hipGraph_t gr;
            HBI_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
            HBI_HIP(hipEventRecord(fork, s0));
            for (int f = 0; f < m; f++) {
                HBI_HIP(hipStreamWaitEvent(rs[f]->stream, fork, 0));
                hbi::launch_attempt(rs[f]);
                HBI_HIP(hipEventRecord(join[f], rs[f]->stream));
                HBI_HIP(hipStreamWaitEvent(s0, join[f], 0));
            }
            HBI_HIP(hipStreamEndCapture(s0, &gr));
            HBI_HIP(hipGraphInstantiate(&all, gr, nullptr, nullptr, 0));
            (void)hipGraphDestroy(gr);
            std::vector<hbi::RkCtl> c(m);
            long long launched = 0;
            const long long budget = 50LL * nsteps + 100;
            int batch = nsteps;
            for (;;) {
                for (int q = 0; q < batch; q++) HBI_HIP(hipGraphLaunch(all, s0));
                launched += batch;
                for (int f = 0; f < m; f++)
                    HBI_HIP(hipMemcpyAsync(&c[f], rs[f]->ctl.p, sizeof(hbi::RkCtl), hipMemcpyDeviceToHost, s0));
                HBI_HIP(hipStreamSynchronize(s0));
                long long need = 0;
                for (int f = 0; f < m; f++)
                    if (!c[f].done) need = std::max(need, c[f].target - c[f].accepted + 2);
                if (need == 0 || launched >= budget || nsteps == 0) break;
                batch = (int)std::max<long long>(1, std::min<long long>(need, budget - launched));
            }
            for (int f = 0; f < m; f++) hbi::step_finish(rs[f], c[f], infos ? &infos[f] : nullptr);
            for (int f = 0; f < m; f++)
                if (c[f].status == HBI_E_STEP) throw hbi::Error{HBI_E_STEP, "Runge-Kutta step size underflow"};
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

int hbi_rk_set_output(hbi_rk *r, const hbi_rk_out_opts *o)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r, "NULL argument");
        hbi::RkPar &P = r->P;
        if (!o) {
            P.out_on = 0;
            r->snap.release(); r->snap_t.release(); r->snap_k.release(); r->islip.release();
            r->eqslip.release(); r->events.release();
        } else {
            HBI_REQUIRE(o->capacity >= 1 && o->ev_capacity >= 1 && o->velth > 0 && o->dtout_inter > 0 &&
                            o->dtout_co > 0 && o->k0 >= 0,
                        "bad output options");
            const size_t n = r->n;
            r->snap.alloc((size_t)o->capacity * (P.has_open ? 6 : 4) * n);
            r->snap_t.alloc(o->capacity);
            r->snap_k.alloc(o->capacity);
            r->islip.alloc(n);
            r->eqslip.alloc((size_t)o->ev_capacity * n);
            r->events.alloc(o->ev_capacity);
            HBI_HIP(hipMemsetAsync(r->islip.p, 0, n * 8, r->stream));
            P.out_on = 1; P.interval = o->interval; P.snap_cap = o->capacity; P.ev_cap = o->ev_capacity;
            P.k0 = o->k0; P.velth = o->velth; P.dtout_inter = o->dtout_inter; P.dtout_co = o->dtout_co;
        }
        // the detector starts afresh (as a new run would) from the current state
        hbi::RkCtl c;
        HBI_HIP(hipMemcpyAsync(&c, r->ctl.p, sizeof(c), hipMemcpyDeviceToHost, r->stream));
        HBI_HIP(hipStreamSynchronize(r->stream));
        c.slipping = 0; c.eventcount = 0; c.hypoloc = 0; c.moment0 = 0.0; c.onset_time = 0.0;
        c.dtout = P.dtout_inter; c.tout = P.dtout_inter;
        c.out_pending = c.out_now = c.onset_now = c.end_now = 0;
        c.nsnap = c.nev = c.out_full = 0;
        hbi::to_device_sync(r->ctl.p, &c, sizeof(c), r->stream);
        r->out_full = false;
        r->dirty = true;
    });
}

int hbi_rk_set_opening(hbi_rk *r, hbi_hmat *hs2, hbi_hmat *hn2)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && hs2 && hn2, "NULL argument");
        HBI_REQUIRE(r->hn, "opening needs the normal-stress matrix (hn of hbi_rk_create)");
        HBI_REQUIRE(hs2->n == r->n && hn2->n == r->n, "opening matrix size differs");
        HBI_REQUIRE(!r->P.out_on, "set the opening before hbi_rk_set_output");
        const int n = r->n;
        const int prec = r->opts.prec;
        r->hs2 = hs2;
        r->hn2 = hn2;
        r->xn.alloc(n + 8);
        HBI_HIP(hipMemsetAsync(r->xn.p, 0, (n + 8) * 8, r->stream));
        r->zo.alloc(hbi::hmat_z_size(hs2, prec));
        r->zo_n.alloc(hbi::hmat_z_size(hn2, prec));
        r->ypart_o.alloc(hbi::hmat_ypart_size(hs2, prec));
        r->ypart_on.alloc(hbi::hmat_ypart_size(hn2, prec));
        r->hxo.alloc(n);
        r->hxo_n.alloc(n);
        r->veln.alloc(n);
        r->slipn.alloc(n);
        HBI_HIP(hipMemsetAsync(r->veln.p, 0, n * 8, r->stream));
        HBI_HIP(hipMemsetAsync(r->slipn.p, 0, n * 8, r->stream));
        HBI_HIP(hipStreamSynchronize(r->stream));
        r->P.has_open = 1;
        r->dirty = true;
    });
}

int hbi_rk_get_opening(hbi_rk *r, double *veln, double *slipn)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && r->P.has_open, "no opening (hbi_rk_set_opening)");
        if (veln) r->veln.download(veln, r->n, r->stream);
        if (slipn) r->slipn.download(slipn, r->n, r->stream);
        HBI_HIP(hipStreamSynchronize(r->stream));
    });
}

int hbi_rk_set_probes(hbi_rk *r, int nprobe, const int *idx)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && nprobe >= 0 && nprobe <= hbi::kRkMaxProbes && (nprobe == 0 || idx), "bad arguments");
        for (int q = 0; q < nprobe; q++) HBI_REQUIRE(idx[q] >= 0 && idx[q] < r->n, "probe index out of range");
        if (nprobe) {
            r->probe_idx.alloc(nprobe);
            r->probe_idx.upload(idx, nprobe, r->stream);
        } else {
            r->probe_idx.release();
        }
        r->probe.release();
        r->probe_rows = 0;
        r->dirty = true;
    });
}

int hbi_rk_get_probes(hbi_rk *r, long long *nrows, double *rows)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && nrows, "NULL argument");
        *nrows = r->probe_rows;
        if (rows && r->probe_rows > 0) {
            r->probe.download(rows, (size_t)r->probe_rows * r->probe_idx.n * hbi::kRkProbe, r->stream);
            HBI_HIP(hipStreamSynchronize(r->stream));
        }
    });
}

int hbi_rk_get_output(hbi_rk *r, int *nsnap, long long *snap_k, double *snap_t, double *fields, int *nev,
                      hbi_rk_event *events, double *eqslip)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && r->P.out_on, "outputs not set (hbi_rk_set_output)");
        hbi::RkCtl c;
        HBI_HIP(hipMemcpyAsync(&c, r->ctl.p, sizeof(c), hipMemcpyDeviceToHost, r->stream));
        HBI_HIP(hipStreamSynchronize(r->stream));
        const size_t n = r->n;
        if (nsnap) *nsnap = c.nsnap;
        if (nev) *nev = c.nev;
        if (c.nsnap > 0) {
            if (snap_k) r->snap_k.download(snap_k, c.nsnap, r->stream);
            if (snap_t) r->snap_t.download(snap_t, c.nsnap, r->stream);
            if (fields) r->snap.download(fields, (size_t)c.nsnap * (r->P.has_open ? 6 : 4) * n, r->stream);
        }
        if (c.nev > 0) {
            if (events) r->events.download(events, c.nev, r->stream);
            if (eqslip) r->eqslip.download(eqslip, (size_t)c.nev * n, r->stream);
        }
        HBI_HIP(hipStreamSynchronize(r->stream));
        // the copies are taken: clear the buffers (the flags of the last step
        // stay, so its output kernel still finds its slot if launched again)
        c.nsnap = c.nev = c.out_full = 0;
        c.out_pending = 0;
        hbi::to_device_sync(r->ctl.p, &c, sizeof(c), r->stream);
        r->out_full = false;
    });
}

int hbi_rk_last_ms(const hbi_rk *r, double *ms)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && ms, "NULL argument");
        *ms = r->last_ms;
    });
}

}  // extern "C"

extern "C" {

int hbi_rk_set_controller(hbi_rk *r, int mode, const hbi_mlp *m, double safety, double ratio_lo, double ratio_hi)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && (mode == 0 || mode == 1), "mode must be 0 (rkqs) or 1 (learned)");
        HBI_REQUIRE(mode == 0 || m, "learned mode needs an MLP");
        HBI_REQUIRE(mode == 0 || (safety > 0 && ratio_lo > 0 && ratio_hi >= ratio_lo), "bad proposal bounds");
        if (mode == 1) {
            r->mlp_p.alloc(hbi::kNParam);
            r->mlp_w.alloc(hbi::kNParam);
            r->mlp_mu.alloc(hbi::kIn);
            r->mlp_sd.alloc(hbi::kIn);
            HBI_HIP(hipStreamSynchronize(m->stream));
            HBI_HIP(hipMemcpyAsync(r->mlp_p.p, m->p.p, hbi::kNParam * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
            HBI_HIP(hipMemcpyAsync(r->mlp_w.p, m->w.p, hbi::kNParam * sizeof(uint16_t), hipMemcpyDeviceToDevice, r->stream));
            HBI_HIP(hipMemcpyAsync(r->mlp_mu.p, m->mu.p, hbi::kIn * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
            HBI_HIP(hipMemcpyAsync(r->mlp_sd.p, m->sd.p, hbi::kIn * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
            HBI_HIP(hipStreamSynchronize(r->stream));
            r->P.safety = safety;
            r->P.ratio_lo = ratio_lo;
            r->P.ratio_hi = ratio_hi;
        }
        r->P.use_mlp = mode;
        r->P.learn = (mode == 1 || r->trace_cap > 0) ? 1 : 0;
        r->dirty = true;
    });
}

int hbi_rk_set_trace(hbi_rk *r, int cap)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && cap >= 0, "bad arguments");
        if (cap > 0) r->trace.alloc((size_t)cap * 14);
        else r->trace.release();
        r->trace_cap = cap;
        r->P.learn = (r->P.use_mlp || cap > 0) ? 1 : 0;
        // restart the row counter
        hbi::RkCtl c;
        HBI_HIP(hipMemcpy(&c, r->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
        c.trace_n = 0;
        hbi::to_device_sync(r->ctl.p, &c, sizeof(c), r->stream);
        r->dirty = true;
    });
}

int hbi_rk_get_trace(hbi_rk *r, double *feat, double *raw, int *count)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(r && count, "NULL argument");
        hbi::RkCtl c;
        HBI_HIP(hipStreamSynchronize(r->stream));
        HBI_HIP(hipMemcpy(&c, r->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
        const int n = std::min(c.trace_n, r->trace_cap);
        *count = n;
        if (n == 0) return;
        std::vector<double> t((size_t)n * 14);
        r->trace.download(t.data(), t.size(), r->stream);
        HBI_HIP(hipStreamSynchronize(r->stream));
        for (int i = 0; i < n; i++) {
            if (feat)
                for (int k = 0; k < hbi::kIn; k++) feat[(size_t)i * hbi::kIn + k] = t[(size_t)i * 14 + k];
            if (raw) {
                raw[2 * (size_t)i] = t[(size_t)i * 14 + 12];
                raw[2 * (size_t)i + 1] = t[(size_t)i * 14 + 13];
            }
        }
    });
}

}  // extern "C"
