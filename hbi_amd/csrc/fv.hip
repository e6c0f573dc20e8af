// fv.hip -- the fluid-driven cycle of main_fv.f90 on the device.
//
// State per element in cluster order, interleaved (psi, tau, sigma_e, ks)
// as main_fv.f90:995-1003; pf, sigma, disp beside it.  One RK attempt is a
// fixed kernel sequence, captured once as a HIP graph and replayed:
//   stage0                  V, matvec input from y                   2064-2077
//   6 x [ H-matvec (shear, and normal unless sigma is constant) -> stage s ]
//       stage s: derivs (4 components) -> k_s, then the next stage's state
//       and V, or after the 6th yout, yerr and the block maxima of
//       |yerr / yout| / eps                                        2259-2342
//   decide                  accept / shrink, hnext, underflow       2149-2256
//   commit                  y <- yout, V, sigma = sigma_e + pf       1016-1028
//   pressure (1 workgroup)  implicit pore-pressure step: Beuler by
//                           conjugate gradients, Picard iterations when
//                           the diffusivity depends on pf; dt bound   1678-1940
//   fields                  sigma_e = sigma - pf, disp, V, mu, kp,
//                           qflow, monitor partials                   1045-1082
//   monitor                 history row, stop criteria                1083-1152
// Grid mode (main_fv3dp.f90, imax x jmax 3dp fault): the same sequence with
// sigma held (not rebuilt from sigma_e + pf), the pressure solved on the 2-D
// grid by one launch of at most 40 resident workgroups (grid-wide barriers
// and reductions between the conjugate-gradient steps), hnext <= dtmax, the minsig rule and the
// monitor values of main_fv3dp.f90:725-815.
// Every kernel first reads the device flag and returns when the batch target
// is reached or the run stopped (rejected attempts skip commit .. monitor).
#include "rkcore.hpp"

#include <algorithm>
#include <cmath>

namespace hbi {

struct FvPar {
    int n, law, backslip, has_hn, pdep, pfconst, bcl, bcr, setting, outpertime;
    double vref, eps, hdiv, ds0, kL, kpmin, s0, eta, beta, Bs, pbcl, pbcr, q0, pinj, tinj, dtout;
    double velmin, velmax, tmax, sg;
    // grid mode (imax, jmax > 0): main_fv3dp's 2-D pore-pressure grid
    int grid, imax, jmax, bct, bcb, nwell;
    double dtmax, minsig, phi0;
    Tab T;
};

struct FvCtl {
    double h, t, hdid, errmax, vmax;
    double tbreak, vbreak_hi, vbreak_lo;   // the call ends after the first accepted step with t > tbreak,
                                           // max |V| > vbreak_hi or max |V| < vbreak_lo
    long long attempts, accepted, rejected, target, hist_base, hist_cap;
    int done, status, acc, niter;
    int sw, nn;   // grid mode: the well schedule's switch flag and 1-based position (implicitsolver2)
};

struct FvArrays {
    const double *a, *b, *dc, *f0, *vpl, *kpmax, *kT, *phi, *pfhyd, *qin, *area;
    double *y, *ytmp, *yout, *k, *vel, *x, *hx, *hx_n;
    double *pf, *sigma, *disp, *velpre, *qflow, *errpart, *monpart, *hist;
    double *probe;          // per accepted step, kFvProbe values for each probed element
    const int *probe_idx;
    int nprobe;
    // grid mode: cluster index of grid cell l; wells (1-based cells, schedule of
    // HBI_FV_MAX_SCHEDULE entries each)
    const int *cell;
    const int *wi, *wj, *wk;
    const double *wq, *wt;
};

constexpr int kFvThreads = 256;
constexpr int kFvRed = 16;   // reduction slots: waves of the largest solve workgroup
constexpr int kFvMon = 6;    // max |V|, sum disp area, sum mu area, max sigma_e, min sigma_e, first argmax |V|
constexpr int kFvHist = 10;
// probe row (the local output of main_fv.f90:1312-1316): V, disp, sigma_e,
// tau, pf, sigma, mu, psi, kp
constexpr int kFvProbe = 9;
constexpr int kFvMaxProbes = 16;
constexpr int kFvMaxCells = HBI_FV_MAX_CELLS;

// main_fv.f90:2039-2146 for one element: o = (dpsi, dtau, dsigma_e, dks)/dt
__device__ __forceinline__ void fv_deriv(const FvPar &P, double gs, double gn, double psi, double tau, double sig,
                                         double ks, double vel, double a, double b, double dc, double f0,
                                         double kpmax, double kT, double *o)
{
#pragma clang fp contract(off)
    const double rigid = kRigid, vs = kVs, vref = P.vref;
    const double dksdt = -vel / P.kL * (ks - kpmax) - (ks - P.kpmin) / kT;
    double dpsidt;
    if (P.law == 0) {
        if (b == 0) dpsidt = 0.0;
        else dpsidt = b / dc * vref * exp((f0 - psi) / b) - b * vel / dc;
    } else {
        const double fss = f0 + (a - b) * log(fabs(vel) / vref);
        dpsidt = -fabs(vel) / dc * (fabs(tau) / sig - fss);
    }
    const double dsigdt = (1 - P.Bs) * gn;
    const double dvdtau = 2 * vref * exp(-psi / a) * cosh(tau / sig / a) / (a * sig);
    const double dvdsig = -2 * vref * exp(-psi / a) * cosh(tau / sig / a) * tau / (a * (sig * sig));
    const double dvdpsi = -vel / a;
    double dtaudt = gs - 0.5 * rigid / vs * (dvdpsi * dpsidt + dvdsig * dsigdt);
    dtaudt = dtaudt / (1.0 + 0.5 * rigid / vs * dvdtau);
    o[0] = dpsidt;
    o[1] = dtaudt;
    o[2] = dsigdt;
    o[3] = dksdt;
}

__global__ void fv_arm_kernel(FvCtl *c, long long nsteps, double tbreak, double vhi, double vlo)
{
    if (c->status == 0) c->done = 0;
    c->tbreak = tbreak;
    c->vbreak_hi = vhi;
    c->vbreak_lo = vlo;
    c->target = c->accepted + nsteps;
    c->hist_base = c->accepted;
    c->hist_cap = nsteps;
    c->acc = 0;
}

// V and the matvec input from y (or from ytmp for a derivs evaluation)
__global__ void fv_stage0_kernel(FvPar P, const int *__restrict__ gate, FvArrays A, int from_tmp)
{
#pragma clang fp contract(off)
    if (*gate) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double *y = from_tmp ? A.ytmp : A.y;
    const double v = slip_rate(y[4 * i], y[4 * i + 1], y[4 * i + 2], A.a[i], P.vref);
    A.vel[i] = v;
    A.x[i] = P.backslip ? v - A.vpl[i] : v;
}

// stage s in 0..5; mode 0: full attempt, 1: derivs only of ytmp (writes k_0)
__global__ __launch_bounds__(kFvThreads) void fv_stage_kernel(FvPar P, const int *__restrict__ gate,
                                                               const FvCtl *__restrict__ c, FvArrays A, int s,
                                                               int mode)
{
#pragma clang fp contract(off)
    __shared__ double sv[kFvThreads / kWave];
    if (*gate) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = P.n;
    const Tab &T = P.T;
    double emax = 0.0;
    if (i < n) {
        const double hx = A.hx[i];
        const double gs = P.hdiv != 0.0 ? hx / P.hdiv : hx;
        double gn = 0.0;
        if (P.has_hn) gn = P.hdiv != 0.0 ? A.hx_n[i] / P.hdiv : A.hx_n[i];
        const double *ys = (s == 0 && mode == 0) ? A.y : A.ytmp;
        double kk[4];
        fv_deriv(P, gs, gn, ys[4 * i], ys[4 * i + 1], ys[4 * i + 2], ys[4 * i + 3], A.vel[i], A.a[i], A.b[i],
                 A.dc[i], A.f0[i], A.kpmax[i], A.kT[i], kk);
        const size_t m4 = 4 * (size_t)n;
        double *ks = A.k + s * m4;
        for (int q = 0; q < 4; q++) ks[4 * i + q] = kk[q];
        if (mode == 0) {
            const double h = c->h;
            const double *k1 = A.k, *k2 = A.k + m4, *k3 = A.k + 2 * m4, *k4 = A.k + 3 * m4, *k5 = A.k + 4 * m4,
                         *k6 = A.k + 5 * m4;
            if (s < 5) {
                double yt[4];
                for (int q = 0; q < 4; q++) {
                    const size_t e = 4 * (size_t)i + q;
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
            } else {
                // max |yerr / yout| / eps over the 4 components (main_fv.f90:2192-2197);
                // a NaN ratio never wins, as in the reference's comparison
                for (int q = 0; q < 4; q++) {
                    const size_t e = 4 * (size_t)i + q;
                    const double yo = A.y[e] + h * (T.c1 * k1[e] + T.c3 * k3[e] + T.c4 * k4[e] + T.c6 * k6[e]);
                    const double ye = h * (T.dc1 * k1[e] + T.dc3 * k3[e] + T.dc4 * k4[e] + T.dc5 * k5[e] + T.dc6 * k6[e]);
                    A.yout[e] = yo;
                    const double r = fabs(ye / yo) / P.eps;
                    if (r > emax) emax = r;
                }
            }
        }
    }
    if (mode != 0 || s != 5) return;
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(emax, o, kWave);
        if (v2 > emax) emax = v2;
    }
    if (threadIdx.x % kWave == 0) sv[threadIdx.x / kWave] = emax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sv[0];
        for (int q = 1; q < kFvThreads / kWave; q++)
            if (sv[q] > m) m = sv[q];
        A.errpart[blockIdx.x] = m;
    }
}

// rkqs decision, main_fv.f90:2182-2252
__global__ void fv_decide_kernel(FvPar P, FvCtl *c, const double *__restrict__ errpart, int nblk)
{
#pragma clang fp contract(off)
    if (c->done) return;
    const double SAFETY = (double)0.9f, PGROW = (double)-0.2f, PSHRNK = (double)-0.25f;
    // the largest block error over the wave's lanes (a maximum: any order gives the same bits)
    double errmax = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 64)
        if (errpart[b] > errmax) errmax = errpart[b];
    for (int o = 32; o > 0; o >>= 1) errmax = fmax(errmax, __shfl_xor(errmax, o, 64));
    if (threadIdx.x != 0) return;
    const double h = c->h;
    c->attempts++;
    c->errmax = errmax;
    if (errmax < 1.0 && errmax > 1e-15) {
        double hn = fmin(2 * h, SAFETY * h * pow(errmax, PGROW));
        if (P.grid) hn = fmin(hn, P.dtmax);
        else if (P.outpertime) hn = fmin(hn, P.dtout);
        c->hdid = h;
        c->t = c->t + h;
        c->h = hn;
        c->acc = 1;
        c->accepted++;
    } else {
        double hnew;
        if (errmax > 1e-15) hnew = fmax(0.5 * h, SAFETY * h * pow(errmax, PSHRNK));
        else hnew = 0.5 * h;
        c->acc = 0;
        c->rejected++;
        c->h = hnew;
        const double xnew = c->t + hnew;
        if (xnew - c->t < 1.e-15) {
            c->status = HBI_E_STEP;
            c->done = 1;
        }
    }
}

// y <- yout; V of the new state (the first half of the displacement
// update uses it, main_fv.f90:1023, 1054); sigma = sigma_e + pf
__global__ void fv_commit_kernel(FvPar P, const FvCtl *__restrict__ c, FvArrays A)
{
#pragma clang fp contract(off)
    if (c->done || !c->acc) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    double yo[4];
    for (int q = 0; q < 4; q++) {
        yo[q] = A.yout[4 * (size_t)i + q];
        A.y[4 * (size_t)i + q] = yo[q];
    }
    A.velpre[i] = slip_rate(yo[0], yo[1], yo[2], A.a[i], P.vref);
    if (!P.grid) A.sigma[i] = yo[2] + A.pf[i];   // grid mode: sigma (total) is held
}

// ---------------------------------------------------------- pore pressure --
struct FvSolveIO {
    const double *pf;       // pf before the step
    const double *y;        // ks (component 3)
    const double *sigma;    // total normal stress (pressure-dependent diffusivity)
    double *pfnew;
    double *rows;           // 3N: the matrix rows of the solve
    double *work;           // 3N: right-hand side, solution, Picard trial
    double h, t;            // used when not gated (the graph path reads hdid and t)
    double *dtnxt;          // in/out (gated: &ctl->h)
    int *niter;
    int gated;
};

// block sum over NT threads; every thread returns the same value.  Two LDS
// slots alternate so one barrier per reduction suffices.
template <int NT>
__device__ __forceinline__ double fv_bsum(double v, double (*red)[kFvRed], int &slot)
{
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) red[slot][w] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < NT / kWave; q++) s += red[slot][q];
    slot ^= 1;
    return s;
}

template <int NT>
__device__ __forceinline__ double fv_bmax(double v, double (*red)[kFvRed], int &slot)
{
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o, kWave);
        if (v2 > v) v = v2;
    }
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) red[slot][w] = v;
    __syncthreads();
    double s = red[slot][0];
    for (int q = 1; q < NT / kWave; q++)
        if (red[slot][q] > s) s = red[slot][q];
    slot ^= 1;
    return s;
}

// Dxx row e of Beuler (main_fv.f90:1786-1816) from the diffusivities in LDS:
// the coefficients of pf(e-1), pf(e), pf(e+1), before the 1 / ds0^2 scaling
__device__ __forceinline__ void fv_dxx_row(const FvPar &P, const double *cd, int e, int n, double &d1, double &d2,
                                           double &d3)
{
#pragma clang fp contract(off)
    d1 = 0.0; d2 = 0.0; d3 = 0.0;
    if (e == 0) {
        d1 = -cd[0] - cd[1];
        d2 = P.bcl == HBI_FV_DIRICHLET ? cd[1] - cd[0] : -d1;
    } else if (e == 1) {
        d1 = P.bcl == HBI_FV_DIRICHLET ? cd[1] / 2 - cd[0] / 2 : cd[0] / 2 + cd[1] / 2;
        d2 = -cd[0] / 2 - cd[1] - cd[2] / 2;
        d3 = cd[1] / 2 + cd[2] / 2;
    } else if (e == n - 2) {
        d1 = cd[n - 3] / 2 + cd[n - 2] / 2;
        d2 = -cd[n - 3] / 2 - cd[n - 2] - cd[n - 1] / 2;
        d3 = P.bcr == HBI_FV_DIRICHLET ? -cd[n - 1] / 2 + cd[n - 2] / 2 : cd[n - 2] / 2 + cd[n - 1] / 2;
    } else if (e == n - 1) {
        // row n keeps its diagonal in column 3 and the sub-diagonal in column 2
        d3 = -cd[n - 1] - cd[n - 2];
        d2 = P.bcr == HBI_FV_DIRICHLET ? cd[n - 2] - cd[n - 1] : -d3;
    } else {
        d1 = cd[e - 1] / 2 + cd[e] / 2;
        d2 = -cd[e - 1] / 2 - cd[e] - cd[e + 1] / 2;
        d3 = cd[e] / 2 + cd[e + 1] / 2;
    }
}

// One Beuler solve (main_fv.f90:1776-1940) by a whole workgroup of NT
// threads.  The set-up pass writes each element's matrix row (sub, diagonal,
// super) to rows[3e..3e+2], the right-hand side to w[e] and the start x =
// pf - pfhyd to w[n + e]; in the conjugate-gradient loop element e = j * NT +
// threadIdx.x (j < NPT) of x, r and A p lives in registers, p (read at
// neighbours) in ps.  With RL (N <= 4096) the rows are then copied into LDS
// over cd (the diffusivities, no longer needed: 3N doubles from cd on),
// else each thread reads its rows from global memory.  pfnew goes to
// w[n + e]; returns the iteration count.
template <int NT, int NPT, bool RL>
__device__ __forceinline__ int fv_beuler(const FvPar &P, double *cd, double *ps, double (*red)[kFvRed], int &slot,
                                         const double *pf, const double *pfhyd, const double *qin, const double *phi,
                                         double h, double t, double *__restrict__ rows, double *__restrict__ w)
{
#pragma clang fp contract(off)
    const int n = P.n;
    const int tid = threadIdx.x;
    const double ds0 = P.ds0;
    const double c1e9 = (double)1e-9f;
    const int half = n / 2 - 1;    // 1-based row N/2
    for (int e = tid; e < n; e += NT) {
        double d1, d2, d3;
        fv_dxx_row(P, cd, e, n, d1, d2, d3);
        d1 = d1 / ds0 / ds0;
        d2 = d2 / ds0 / ds0;
        d3 = d3 / ds0 / ds0;
        double lo = 0.0, di, up = 0.0;
        if (e == 0) {
            di = 1.0 - h * d1;
            up = -h * d2;
        } else if (e == n - 1) {
            di = 1.0 - h * d3;
            lo = -h * d2;
        } else {
            lo = -h * d1;
            di = 1.0 - h * d2;
            up = -h * d3;
        }
        // penalty vector (main_fv.f90:1842-1868)
        double sat = 0.0;
        if (P.bcl == HBI_FV_DIRICHLET) {
            if (e == 0) sat = -1.0 * cd[0] / ds0 / ds0 * (P.pbcl - pfhyd[0]) * h * 2;
            if (e == 1) sat = -0.5 * cd[0] / ds0 / ds0 * (P.pbcl - pfhyd[0]) * h * 2;
        } else if (P.setting == HBI_FV_STEADYSTATE && e == 0) {
            sat = -P.q0 / P.beta / phi[0] * c1e9 * h / ds0 * 2;
        }
        if (P.bcr == HBI_FV_DIRICHLET) {
            const double satn = -1 * cd[n - 1] / ds0 / ds0 * (P.pbcr - pfhyd[n - 1]) * h * 2;
            if (e == n - 1) sat = satn;
            if (e == n - 2) sat = satn / 2;
        } else if (e == n - 1) {
            if (P.setting == HBI_FV_STEADYSTATE) sat = P.q0 / P.beta / phi[n - 1] * c1e9 * h / ds0 * 2;
            if (P.setting == HBI_FV_THRUST) sat = qin[1999] / P.beta / phi[n - 1] * c1e9 * h / ds0 * 2;
        }
        if (P.setting == HBI_FV_THRUST && e >= 2 && e <= n - 3) sat = -qin[e] / P.beta / phi[e] * c1e9 * h / ds0;
        double b = pf[e] - sat - pfhyd[e];
        if (e == half) {
            if (P.setting == HBI_FV_INJECTIONP) {
                b = b + h * P.pinj / 1e1;
                di = di + h / 1e1;
            } else if (P.setting == HBI_FV_INJECTIONQ && t < P.tinj) {
                b = b + h * qin[half] / P.beta / phi[half] * c1e9 / ds0;
            }
        }
        if (e == 0 || e == n - 1) {
            b = b / 2;
            lo = lo / 2;
            di = di / 2;
            up = up / 2;
        }
        rows[3 * (size_t)e] = lo;
        rows[3 * (size_t)e + 1] = di;
        rows[3 * (size_t)e + 2] = up;
        w[e] = b;
        const double x0 = pf[e] - pfhyd[e];
        w[n + e] = x0;
        ps[e] = x0;
    }
    __syncthreads();   // ps holds x; every row has read cd; the rows are written
    if constexpr (RL) {
        for (int q = tid; q < 3 * n; q += NT) cd[q] = rows[q];
        __syncthreads();
    }
    const double *R = RL ? cd : rows;
    // the matrix product in the reference's order of terms (1895-1903)
    auto apply = [&](const double *v, int e) -> double {
        const double *rw = R + 3 * (size_t)e;
        if (e == 0) return v[0] * rw[1] + v[1] * rw[2];
        if (e == n - 1) return v[n - 1] * rw[1] + v[n - 2] * rw[0];
        return v[e - 1] * rw[0] + v[e] * rw[1] + v[e + 1] * rw[2];
    };
    double x[NPT], r[NPT], rr = 0.0;
#pragma unroll
    for (int j = 0; j < NPT; j++) {
        const int e = j * NT + tid;
        x[j] = r[j] = 0.0;
        if (e < n) {
            x[j] = ps[e];
            r[j] = w[e] - apply(ps, e);
            rr += r[j] * r[j];
        }
    }
    double rsold = fv_bsum<NT>(rr, red, slot);   // its barrier ends the reads of ps
#pragma unroll
    for (int j = 0; j < NPT; j++) {
        const int e = j * NT + tid;
        if (e < n) ps[e] = r[j];
    }
    const double tol = (double)1e-4f;
    const double thr = tol * tol * (double)n;
    const int itermax = 10000;
    int niter = 0;
    if (!(rsold < thr)) {
        niter = itermax;
        for (int it = 1; it <= itermax; it++) {
            const double tmp1 = rsold;   // = sum(r * r) of the current r
            __syncthreads();             // ps holds p
            double mp = 0.0, m[NPT];
#pragma unroll
            for (int j = 0; j < NPT; j++) {
                const int e = j * NT + tid;
                m[j] = 0.0;
                if (e < n) {
                    m[j] = apply(ps, e);
                    mp += m[j] * ps[e];
                }
            }
            const double tmp2 = fv_bsum<NT>(mp, red, slot);
            const double alpha = tmp1 / tmp2;
            double rs = 0.0;
#pragma unroll
            for (int j = 0; j < NPT; j++) {
                const int e = j * NT + tid;
                if (e < n) {
                    x[j] = x[j] + alpha * ps[e];
                    r[j] = r[j] - alpha * m[j];
                    rs += r[j] * r[j];
                }
            }
            const double rsnew = fv_bsum<NT>(rs, red, slot);
            if (rsnew < thr) {
                niter = it;
                break;
            }
#pragma unroll
            for (int j = 0; j < NPT; j++) {
                const int e = j * NT + tid;
                if (e < n) ps[e] = r[j] + (rsnew / rsold) * ps[e];
            }
            rsold = rsnew;
        }
    }
#pragma unroll
    for (int j = 0; j < NPT; j++) {
        const int e = j * NT + tid;
        if (e < n) w[n + e] = x[j] + pfhyd[e];
    }
    return niter;
}

// implicitsolver2 (main_fv.f90:1727-1774) or, pressure dependent,
// implicitsolver (1678-1724): pfnew, CG iterations and the dt bound.  The
// work vector w (3N) holds the right-hand side, pfnew and the Picard trial.
template <int NT, int NPT, bool RL>
__global__ __launch_bounds__(NT) void fv_pressure_kernel(FvPar P, FvCtl *c, FvArrays A, FvSolveIO io)
{
#pragma clang fp contract(off)
    // RL: ps = [0, 4096), cd and then the rows from 4096 on; else cd, ps halves
    __shared__ double lds[2 * kFvMaxCells];
    __shared__ double red[2][kFvRed];
    double *cd = RL ? lds + kFvMaxCells / 2 : lds;
    double *ps = RL ? lds : lds + kFvMaxCells;
    if (io.gated && (c->done || !c->acc)) return;
    const int n = P.n;
    const int tid = threadIdx.x;
    const double h = io.gated ? c->hdid : io.h;
    const double t = io.gated ? c->t : io.t;
    double *dtnxt = io.gated ? &c->h : io.dtnxt;
    double *w = io.work, *pft = io.work + 2 * (size_t)n;
    int slot = 0;
    for (int e = tid; e < n; e += NT) pft[e] = io.pf[e];
    int niter = 0;
    const int kitmax = P.pdep ? 20 : 1;
    for (int kit = 1; kit <= kitmax; kit++) {
        __syncthreads();   // the previous solve's readers of cd and ps are done
        for (int e = tid; e < n; e += NT) {
            const double cc = P.eta * P.beta * A.phi[e];
            const double ks = io.y[4 * (size_t)e + 3];
            cd[e] = P.pdep ? ks * exp(-(io.sigma[e] - pft[e]) / P.s0) / cc * 1e-6 : ks / cc * 1e-6;
        }
        __syncthreads();
        niter = fv_beuler<NT, NPT, RL>(P, cd, ps, red, slot, io.pf, A.pfhyd, A.qin, A.phi, h, t, io.rows, w);
        if (!P.pdep) break;
        double err = 0.0;
        for (int e = tid; e < n; e += NT) {
            const double d = fabs(pft[e] - w[n + e]);
            if (d > err) err = d;
        }
        err = fv_bmax<NT>(err, red, slot);
        if (err < (double)1e-3f) break;
        for (int e = tid; e < n; e += NT) pft[e] = w[n + e];
    }
    // pf(e) is read and written by its own thread only, so pfnew may alias pf
    double dmax = -HUGE_VAL;
    for (int e = tid; e < n; e += NT) {
        const double old = io.pf[e];
        const double x = w[n + e];
        io.pfnew[e] = x;
        const double d = x - old;
        if (d > dmax) dmax = d;
    }
    dmax = fv_bmax<NT>(dmax, red, slot);
    if (tid == 0) {
        const double dpth = P.pdep ? (double)0.3f : (double)0.1f;
        double dt = *dtnxt;
        if (dt / h * dmax > dpth) dt = dpth * h / dmax;
        if (P.pdep && P.outpertime) dt = fmin(dt, P.dtout);
        *dtnxt = dt;
        *io.niter = niter;
    }
}

// ------------------------------------------- the 2-D grid of main_fv3dp --
// Work arrays of the 2-D solve, grid order (cell l = i jmax + j): cd the
// diffusivities, co the six coefficients of each row (Amx along j, Amy along
// i, Beuler's storage: first row (diagonal, right), last row (left at 1,
// diagonal at 2)), b, the directions p / p2 of alternate iterations and the
// residual r (read at neighbours), the pressure at the start pg, the Picard
// trial pt, the new pressure pn, pfhyd hg; red: the per-workgroup partials of
// two reductions in flight.
struct Fv2Work {
    double *cd, *co, *b, *p, *pg, *pt, *pn, *hg, *red, *r, *p2;
    unsigned *bar;   // grid barrier: release word, then one 128-byte arrival slot per workgroup
    int *fail;       // set when a barrier wait ran past its spin limit
};
constexpr int kFv2Threads = 256;
constexpr long long kFv2Spin = 1LL << 24;
constexpr int kFv2Failed = 5;   // run status: a barrier of the 2-D solve timed out
constexpr int kFv2MaxBlocks = 40;     // workgroups of the 2-D solve at most: 40 of 256 threads measured fastest on
                                      // the 200 x 200 grid (barrier cost against cells per thread)

// Grid-wide barrier of a launch whose workgroups are all resident (at most
// kFv2MaxBlocks, or with 16 cells per thread at most one per CU: a fraction of
// what the CUs hold at once, resident once the stream's earlier kernels end),
// without a contended counter: workgroup b publishes generation g in its own
// 128-byte slot of W.bar, workgroup 0 waits (one thread per slot) until every
// slot shows g and then publishes g as the release word W.bar[0]; the others
// wait for it.  g = the release word read at kernel start (every workgroup
// reads it before it first arrives, and the release word cannot move before
// all have arrived) plus the barrier's ordinal.  Each thread releases its
// writes before, and acquires after.  Waits are bounded: past kFv2Spin polls
// W.fail is raised and the wait ends (the host discards the result).
constexpr int kFv2Slot = 32;   // unsigned words per arrival slot (128 bytes)
__device__ __forceinline__ void fv2_barrier(const Fv2Work &W, unsigned &gen)
{
    gen++;
    // every wave's stores complete (workgroup scope: no cache write-back per
    // wave); the arriving thread's release at device scope then publishes the
    // workgroup's writes once
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    const int G = gridDim.x;
    int bad = 0;
    if (threadIdx.x == 0)
        __hip_atomic_store(&W.bar[kFv2Slot * (1 + blockIdx.x)], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (blockIdx.x == 0) {
        for (int q = threadIdx.x; q < G; q += blockDim.x) {
            long long spins = 0;
            while (__hip_atomic_load(&W.bar[kFv2Slot * (1 + q)], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != gen)
                if (++spins > kFv2Spin) {
                    bad = 1;
                    break;
                }
        }
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(&W.bar[0], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
        long long spins = 0;
        while (__hip_atomic_load(&W.bar[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != gen)
            if (++spins > kFv2Spin) {
                bad = 1;
                break;
            }
    }
    if (bad) atomicOr(W.fail, 1);
    // the waiting thread's device-scope acquire invalidated the CU's and the
    // XCD's caches; the barrier orders the other waves' loads after it
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Sum (max) of one value per thread over the whole grid: per workgroup in a
// fixed order, the partials to red[slot], a grid barrier, then every
// workgroup folds all partials in the same order (lane k takes partials k,
// k + 64, ... then a fixed butterfly), so all get the same bits.  Two slots
// alternate: a slot is written again only after the next reduction's barrier.
template <bool MAX>
__device__ __forceinline__ double fv2_reduce(double v, const Fv2Work &W, int &slot, double *lds, unsigned &gen)
{
    double *red = W.red;
    const int tid = threadIdx.x, lane = tid % kWave, w = tid / kWave;
    const int G = gridDim.x;
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double u = __shfl_xor(v, o, kWave);
        v = MAX ? fmax(v, u) : v + u;
    }
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double* out = red + (size_t)slot * G;
    if (tid == 0) {
        double b = lds[0];
        for (int q = 1; q < kFv2Threads / kWave; q++) b = MAX ? fmax(b, lds[q]) : b + lds[q];
        out[blockIdx.x] = b;
    }
    fv2_barrier(W, gen);
    if (w == 0) {
        double a = MAX ? -HUGE_VAL : 0.0;
        for (int q = lane; q < G; q += kWave) a = MAX ? fmax(a, out[q]) : a + out[q];
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const double u = __shfl_xor(a, o, kWave);
            a = MAX ? fmax(a, u) : a + u;
        }
        if (lane == 0) lds[kFv2Threads / kWave] = a;
    }
    __syncthreads();
    const double r = lds[kFv2Threads / kWave];
    __syncthreads();
    slot ^= 1;
    return r;
}

// D(row, 0..2) of one grid line at position k of len (Beuler's Dxx along j
// or Dyy along i, before the 1 / ds0^2 scaling); c(q) the line's q-th
// diffusivity.  The same expressions as the 1-D fv_dxx_row.
template <typename CF>
__device__ __forceinline__ void fv2_line_row(int k, int len, int lo_bc, int hi_bc, CF c, double &d1, double &d2,
                                             double &d3)
{
#pragma clang fp contract(off)
    d1 = 0.0; d2 = 0.0; d3 = 0.0;
    if (k == 0) {
        d1 = -c(0) - c(1);
        d2 = lo_bc == HBI_FV_DIRICHLET ? c(1) - c(0) : -d1;
    } else if (k == 1) {
        d1 = lo_bc == HBI_FV_DIRICHLET ? c(1) / 2 - c(0) / 2 : c(0) / 2 + c(1) / 2;
        d2 = -c(0) / 2 - c(1) - c(2) / 2;
        d3 = c(1) / 2 + c(2) / 2;
    } else if (k == len - 2) {
        d1 = c(len - 3) / 2 + c(len - 2) / 2;
        d2 = -c(len - 3) / 2 - c(len - 2) - c(len - 1) / 2;
        d3 = hi_bc == HBI_FV_DIRICHLET ? -c(len - 1) / 2 + c(len - 2) / 2 : c(len - 2) / 2 + c(len - 1) / 2;
    } else if (k == len - 1) {
        d3 = -c(len - 1) - c(len - 2);
        d2 = hi_bc == HBI_FV_DIRICHLET ? c(len - 2) - c(len - 1) : -d3;
    } else {
        d1 = c(k - 1) / 2 + c(k) / 2;
        d2 = -c(k - 1) / 2 - c(k) - c(k + 1) / 2;
        d3 = c(k) / 2 + c(k + 1) / 2;
    }
}

// Am of a line row from D (already / ds0^2): 0.5 - h D on the diagonal
__device__ __forceinline__ void fv2_line_matrix(int k, int len, double h, double d1, double d2, double d3, double *a)
{
#pragma clang fp contract(off)
    if (k == 0) {
        a[0] = 0.5 - h * d1; a[1] = -h * d2; a[2] = 0.0;
    } else if (k == len - 1) {
        a[0] = 0.0; a[1] = -h * d2; a[2] = 0.5 - h * d3;
    } else {
        a[0] = -h * d1; a[1] = 0.5 - h * d2; a[2] = -h * d3;
    }
}

// (A v)(l): the line products along j and i in the reference's order of terms
// (main_fv3dp.f90:1566-1585), m = mx + my
template <typename VF>
__device__ __forceinline__ double fv2_apply(const FvPar &P, const double *co, VF v, int l)
{
#pragma clang fp contract(off)
    const int jm = P.jmax, i = l / jm, j = l % jm;
    const double *a = co + 6 * (size_t)l;
    const double vl = v(l);
    double mx, my;
    if (j == 0) mx = vl * a[0] + v(l + 1) * a[1];
    else if (j == jm - 1) mx = vl * a[2] + v(l - 1) * a[1];
    else mx = v(l - 1) * a[0] + vl * a[1] + v(l + 1) * a[2];
    if (i == 0) my = vl * a[3] + v(l + jm) * a[4];
    else if (i == P.imax - 1) my = vl * a[5] + v(l - jm) * a[4];
    else my = v(l - jm) * a[3] + vl * a[4] + v(l + jm) * a[5];
    return mx + my;
}

// The rate of well w at time t (main_fv3dp.f90:1553-1565)
__device__ __forceinline__ double fv2_well_rate(const FvArrays &A, int w, double t)
{
#pragma clang fp contract(off)
    const double *qv = A.wq + (size_t)w * HBI_FV_MAX_SCHEDULE, *qt = A.wt + (size_t)w * HBI_FV_MAX_SCHEDULE;
    const int kl = A.wk[w];
    double q = 0.0;
    for (int k = 0; k < kl - 1; k++) {
        const double t0 = qt[k], t1 = qt[k + 1], v0 = qv[k], v1 = qv[k + 1];
        if (t >= t0 && t <= t1) q = (v1 - v0) / (t1 - t0) * (t - t0) + v0;
        else if (t > t1 && k == kl - 2) q = v1;
    }
    return q;
}

// implicitsolver2 / implicitsolver of main_fv3dp (1314-1424) with Beuler
// (1426-1649) on the imax x jmax grid: one launch of resident workgroups, kFv2Threads
// threads per workgroup, NPT cells per thread (cell l = g + q G kFv2Threads,
// g the global thread index); x and r of its cells in registers, p and the
// coefficients in global memory (neighbours).  Gated like the 1-D kernel.
template <int NPT>
__global__ __launch_bounds__(kFv2Threads) void fv2_pressure_kernel(FvPar P, FvCtl *c, FvArrays A, FvSolveIO io,
                                                                    Fv2Work W)
{
#pragma clang fp contract(off)
    __shared__ double lds[kFv2Threads / kWave + 1];
    if (io.gated && (c->done || !c->acc)) return;
    unsigned gen = __hip_atomic_load(&W.bar[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const int n = P.n, im = P.imax, jm = P.jmax;
    const int gt = blockIdx.x * kFv2Threads + threadIdx.x, stride = gridDim.x * kFv2Threads;
    const double h = io.gated ? c->hdid : io.h;
    const double t = io.gated ? c->t : io.t;
    const double ds0 = P.ds0;
    const double cc = P.eta * P.beta * P.phi0;
    const double e9 = (double)1e-9f, e12 = (double)1e-12f;
    int slot = 0;
    // grid-order copies of pf and pfhyd; the Picard trial starts at pf
#pragma unroll
    for (int q = 0; q < NPT; q++) {
        const int l = gt + q * stride;
        if (l < n) {
            const int e = A.cell[l];
            W.pg[l] = io.pf[e];
            W.pt[l] = io.pf[e];
            W.hg[l] = A.pfhyd[e];
        }
    }
    int niter = 0;
    const int kitmax = P.pdep ? 20 : 1;
    for (int kit = 1; kit <= kitmax; kit++) {
#pragma unroll
        for (int q = 0; q < NPT; q++) {
            const int l = gt + q * stride;
            if (l < n) {
                const int e = A.cell[l];
                const double ks = io.y[4 * (size_t)e + 3];
                W.cd[l] = P.pdep ? ks * exp(-(io.sigma[e] - W.pt[l]) / P.s0) / cc * 1e-6 : ks / cc * 1e-6;
            }
        }
        fv2_barrier(W, gen);
        // the rows, the right-hand side and the start x = pf - pfhyd (p holds it for r = b - A x)
        double x[NPT], r[NPT];
#pragma unroll
        for (int q = 0; q < NPT; q++) {
            const int l = gt + q * stride;
            x[q] = r[q] = 0.0;
            if (l >= n) continue;
            const int i = l / jm, j = l % jm;
            const double *cd = W.cd;
            double d1, d2, d3, a[6];
            fv2_line_row(j, jm, P.bct, P.bcb, [&](int k) { return cd[i * jm + k]; }, d1, d2, d3);
            fv2_line_matrix(j, jm, h, d1 / ds0 / ds0, d2 / ds0 / ds0, d3 / ds0 / ds0, a);
            fv2_line_row(i, im, P.bcl, P.bcr, [&](int k) { return cd[k * jm + j]; }, d1, d2, d3);
            fv2_line_matrix(i, im, h, d1 / ds0 / ds0, d2 / ds0 / ds0, d3 / ds0 / ds0, a + 3);
            double *co = W.co + 6 * (size_t)l;
            for (int k = 0; k < 6; k++) co[k] = a[k];
            // penalty vector (1531-1545)
            const double hy = W.hg[l];
            double sat = 0.0;
            if (P.setting == HBI_FV_STEADYSTATE) {
                if (i == 0) sat = -P.q0 / P.beta / P.phi0 * e9 * h / ds0 * 2;
                if (i == im - 1) sat = P.q0 / P.beta / P.phi0 * e9 * h / ds0 * 2;
            }
            if (P.setting == HBI_FV_THRUST) {
                const double c0 = cd[i * jm], h0 = W.hg[i * jm];
                const double cl = cd[i * jm + jm - 1], hl = W.hg[i * jm + jm - 1];
                if (j == 0) sat = -1.0 * c0 / ds0 / ds0 * (P.pbcl - h0) * h * 2;
                if (j == 1) sat = -0.5 * c0 / ds0 / ds0 * (P.pbcl - h0) * h * 2;
                if (P.bcb == HBI_FV_NEUMANN) {
                    if (j == jm - 1) sat = -P.q0 / P.beta / P.phi0 * e9 * h / ds0 * 2;
                } else {
                    if (j == jm - 1) sat = -1.0 * cl / ds0 / ds0 * (P.pbcr - hl) * h * 2;
                    if (j == jm - 2) sat = -0.5 * cl / ds0 / ds0 * (P.pbcr - hl) * h * 2;
                }
            }
            const double pfl = W.pg[l];
            double bb = pfl - sat - hy;
            if (P.setting == HBI_FV_INJECTIONQ) {   // 'injection' (1546-1565)
                if (P.nwell > 0) {
                    for (int w = 0; w < P.nwell; w++)
                        if (A.wi[w] - 1 == i && A.wj[w] - 1 == j)
                            bb = bb + h * fv2_well_rate(A, w, t) / P.beta / P.phi0 * e12 / ds0 / ds0;
                } else if (i == im / 2 - 1 && j == jm / 2 - 1) {
                    bb = bb + h * P.q0 / P.beta / P.phi0 * e12 / ds0 / ds0;
                }
            }
            W.b[l] = bb;
            x[q] = pfl - hy;
            W.p[l] = x[q];
        }
        fv2_barrier(W, gen);
        double rr = 0.0;
#pragma unroll
        for (int q = 0; q < NPT; q++) {
            const int l = gt + q * stride;
            if (l < n) {
                r[q] = W.b[l] - fv2_apply(P, W.co, [&](int k) { return W.p[k]; }, l);
                W.r[l] = r[q];
                rr += r[q] * r[q];
            }
        }
        double rsold = fv2_reduce<false>(rr, W, slot, lds, gen);   // its barrier ends the reads of p = x
        const double tol = (double)1e-6f;
        const double thr = tol * tol * (double)im * (double)jm;
        const int itermax = 1000;
        niter = 0;
        if (!(rsold < thr)) {
            // The reference's p = r + (rsnew / rsold) p at the end of an
            // iteration is taken into the next one: each thread forms the new p
            // of its cells and of their neighbours from r and the previous p
            // (same operands, same bits), so an iteration needs two barriers.
            niter = itermax;
            double beta = 0.0;
            for (int it = 1; it <= itermax; it++) {
                const double tmp1 = rsold;
                const double *pold = (it & 1) ? W.p2 : W.p;
                double *pnew = (it & 1) ? W.p : W.p2;
                const bool first = it == 1;
                auto pval = [&](int k) { return first ? W.r[k] : W.r[k] + beta * pold[k]; };
                double mp = 0.0, m[NPT], pv[NPT];
#pragma unroll
                for (int q = 0; q < NPT; q++) {
                    const int l = gt + q * stride;
                    m[q] = pv[q] = 0.0;
                    if (l < n) {
                        pv[q] = first ? r[q] : r[q] + beta * pold[l];
                        pnew[l] = pv[q];
                        m[q] = fv2_apply(P, W.co, pval, l);
                        mp += m[q] * pv[q];
                    }
                }
                const double tmp2 = fv2_reduce<false>(mp, W, slot, lds, gen);
                const double alpha = tmp1 / tmp2;
                double rs = 0.0;
#pragma unroll
                for (int q = 0; q < NPT; q++) {
                    const int l = gt + q * stride;
                    if (l < n) {
                        x[q] = x[q] + alpha * pv[q];
                        r[q] = r[q] - alpha * m[q];
                        W.r[l] = r[q];
                        rs += r[q] * r[q];
                    }
                }
                const double rsnew = fv2_reduce<false>(rs, W, slot, lds, gen);
                if (rsnew < thr) {
                    niter = it;
                    break;
                }
                beta = rsnew / rsold;
                rsold = rsnew;
            }
        }
        double err = 0.0;
#pragma unroll
        for (int q = 0; q < NPT; q++) {
            const int l = gt + q * stride;
            if (l < n) {
                const double pn = x[q] + W.hg[l];
                W.pn[l] = pn;
                err = fmax(err, fabs(W.pt[l] - pn));
            }
        }
        if (!P.pdep) break;
        err = fv2_reduce<true>(err, W, slot, lds, gen);
        if (err < (double)1e-5f) break;
#pragma unroll
        for (int q = 0; q < NPT; q++) {
            const int l = gt + q * stride;
            if (l < n) W.pt[l] = W.pn[l];
        }
    }
    // pf <- pfnew, the dt bound on max |dpf| (dpth 0.1) and the well schedule
    double amax = 0.0;
#pragma unroll
    for (int q = 0; q < NPT; q++) {
        const int l = gt + q * stride;
        if (l < n) {
            const double d = W.pn[l] - W.pg[l];
            amax = fmax(amax, fabs(d));
            io.pfnew[A.cell[l]] = W.pn[l];
        }
    }
    amax = fv2_reduce<true>(amax, W, slot, lds, gen);
    if (gt == 0) {
        double *dtnxt = io.gated ? &c->h : io.dtnxt;
        double dt = *dtnxt;
        const double dpth = (double)0.1f;
        if (dt / h * amax > dpth) dt = dpth * h / amax;
        if (!P.pdep) {
            if (c->sw) {
                dt = 20.0;
                c->sw = 0;
            }
            if (P.nwell > 0 && c->nn <= A.wk[0]) {
                const double tq = A.wt[c->nn - 1];
                if (t + dt > tq) {
                    dt = tq - t - 1.0;
                    c->sw = 1;
                    c->nn += 2;
                }
            }
        }
        *dtnxt = dt;
        *io.niter = niter;
        if (io.gated && __hip_atomic_load(W.fail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) {
            c->status = kFv2Failed;
            c->done = 1;
        }
    }
}

// sigma_e = sigma - pf, displacement (trapezoid over the post-RK and the final
// V), mu, kp, qflow and the monitor partials (main_fv.f90:1045-1082, 1083-1134)
__global__ __launch_bounds__(kFvThreads) void fv_fields_kernel(FvPar P, const FvCtl *__restrict__ c, FvArrays A)
{
#pragma clang fp contract(off)
    __shared__ double sp[kFvThreads / kWave][kFvMon];
    if (c->done || !c->acc) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = P.n;
    const double dt = c->hdid;
    double m[kFvMon] = {P.grid ? -HUGE_VAL : -1.0, 0.0, 0.0, -HUGE_VAL, HUGE_VAL, -1.0};
    if (i < n) {
        const size_t e = 4 * (size_t)i;
        const double psi = A.y[e], tau = A.y[e + 1], ks = A.y[e + 3];
        // grid mode: where the RK's sigma_e fell below minsig, sigma = pf + minsig
        // (main_fv3dp.f90:765-770: the test reads sigma_e before the pressure step)
        if (P.grid && A.y[e + 2] < P.minsig) A.sigma[i] = A.pf[i] + P.minsig;
        const double sige = A.sigma[i] - A.pf[i];
        A.y[e + 2] = sige;
        double d = A.disp[i];
        d = d + A.velpre[i] * dt * 0.5;
        const double v = slip_rate(psi, tau, sige, A.a[i], P.vref);
        d = d + v * dt * 0.5;
        A.disp[i] = d;
        A.vel[i] = v;
        const double mu = tau / sige;
        if (!P.grid) {
            // qflow with kp of this element and pf of the neighbours (final pf);
            // main_fv3dp keeps qflow = q0
            const double kp = P.pdep ? ks * exp(-sige / P.s0) : ks;
            const double *pf = A.pf, *ph = A.pfhyd;
            double q;
            if (i == 0) q = -kp / P.eta * 1e3 * (pf[1] - ph[1] - pf[0] + ph[0]) / P.ds0;
            else if (i == n - 1) q = -kp / P.eta * 1e3 * (pf[n - 1] - ph[n - 1] - pf[n - 2] + ph[n - 2]) / P.ds0;
            else q = -kp / P.eta * 1e3 * (pf[i + 1] - ph[i + 1] - pf[i - 1] + ph[i - 1]) / 2 / P.ds0;
            A.qflow[i] = q;
        }
        m[0] = P.grid ? v : fabs(v);   // main_fv3dp's maxval(vel); main_fv's max |V|
        m[1] = d * A.area[i];
        m[2] = mu * A.area[i];
        m[3] = sige;
        m[4] = sige;
        m[5] = (double)i;
    }
    for (int o = 1; o < kWave; o <<= 1) {
        double o2[kFvMon];
        for (int k = 0; k < kFvMon; k++) o2[k] = __shfl_down(m[k], o, kWave);
        const int lane = threadIdx.x % kWave;
        if ((lane & (2 * o - 1)) == 0 && lane + o < kWave) {
            if (o2[0] > m[0]) { m[0] = o2[0]; m[5] = o2[5]; }
            m[1] = m[1] + o2[1];
            m[2] = m[2] + o2[2];
            m[3] = fmax(m[3], o2[3]);
            m[4] = fmin(m[4], o2[4]);
        }
    }
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0)
        for (int k = 0; k < kFvMon; k++) sp[w][k] = m[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double b[kFvMon];
        for (int k = 0; k < kFvMon; k++) b[k] = sp[0][k];
        for (int q = 1; q < kFvThreads / kWave; q++) {
            if (sp[q][0] > b[0]) { b[0] = sp[q][0]; b[5] = sp[q][5]; }
            b[1] = b[1] + sp[q][1];
            b[2] = b[2] + sp[q][2];
            b[3] = fmax(b[3], sp[q][3]);
            b[4] = fmin(b[4], sp[q][4]);
        }
        for (int k = 0; k < kFvMon; k++) A.monpart[kFvMon * (size_t)blockIdx.x + k] = b[k];
    }
}

__global__ void fv_monitor_kernel(FvPar P, FvCtl *c, FvArrays A, int nblk)
{
#pragma clang fp contract(off)
    if (c->done || !c->acc) return;
    // max |V| with its first element, max and min sigma_e over the wave's
    // lanes (exact in any order; ties of max |V| go to the first block), then
    // the two sums in block order by lane 0 from LDS, kMonChunk blocks at a time
    const int lane = threadIdx.x;
    double vmx = -HUGE_VAL, vloc = 0.0, smx = -HUGE_VAL, smn = HUGE_VAL;
    int vblk = nblk;
    for (int q = lane; q < nblk; q += 64) {
        const double *mp = A.monpart + kFvMon * (size_t)q;
        if (mp[0] > vmx) { vmx = mp[0]; vloc = mp[5]; vblk = q; }
        smx = fmax(smx, mp[3]);
        smn = fmin(smn, mp[4]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(vmx, o, 64), l2 = __shfl_xor(vloc, o, 64);
        const int b2 = __shfl_xor(vblk, o, 64);
        if (v2 > vmx || (v2 == vmx && b2 < vblk)) { vmx = v2; vloc = l2; vblk = b2; }
        smx = fmax(smx, __shfl_xor(smx, o, 64));
        smn = fmin(smn, __shfl_xor(smn, o, 64));
    }
    constexpr int kMonChunk = 512;
    __shared__ double s1[kMonChunk], s2[kMonChunk];
    double sum1 = 0.0, sum2 = 0.0;
    for (int q0 = 0; q0 < nblk; q0 += kMonChunk) {
        const int m = min(kMonChunk, nblk - q0);
        __syncthreads();
        for (int q = lane; q < m; q += 64) {
            s1[q] = A.monpart[kFvMon * (size_t)(q0 + q) + 1];
            s2[q] = A.monpart[kFvMon * (size_t)(q0 + q) + 2];
        }
        __syncthreads();
        if (lane == 0)
            for (int q = 0; q < m; q++) {
                sum1 = q0 + q == 0 ? s1[q] : sum1 + s1[q];
                sum2 = q0 + q == 0 ? s2[q] : sum2 + s2[q];
            }
    }
    if (lane != 0) return;
    double b[kFvMon];
    b[0] = vmx; b[1] = sum1; b[2] = sum2; b[3] = smx; b[4] = smn; b[5] = vloc;
    const double vm = b[0];
    c->vmax = vm;
    const long long row = c->accepted - c->hist_base - 1;
    if (A.hist && row >= 0 && row < c->hist_cap) {
        double *hr = A.hist + kFvHist * row;
        hr[0] = c->t; hr[1] = log10(vm); hr[2] = b[1] / P.sg; hr[3] = b[2] / P.sg; hr[4] = b[3]; hr[5] = b[4];
        hr[6] = b[5] + 1.0; hr[7] = c->errmax; hr[8] = c->hdid; hr[9] = (double)c->niter;
    }
    if (A.probe && row >= 0 && row < c->hist_cap)
        for (int q = 0; q < A.nprobe; q++) {
            const size_t i = A.probe_idx[q];
            const double *yi = A.y + 4 * i;
            const double sige = yi[2], tau = yi[1], ks = yi[3];
            double *pr = A.probe + ((size_t)row * A.nprobe + q) * kFvProbe;
            pr[0] = A.vel[i]; pr[1] = A.disp[i]; pr[2] = sige; pr[3] = tau; pr[4] = A.pf[i]; pr[5] = A.sigma[i];
            pr[6] = tau / sige; pr[7] = yi[0]; pr[8] = P.pdep || P.grid ? ks * exp(-sige / P.s0) : ks;
        }
    // stop controls, main_fv.f90:1136-1152
    if (vm > P.velmax || vm < P.velmin) c->status = 1;
    else if (c->t > P.tmax) c->status = 3;
    else if (b[4] < (double)0.2f) c->status = 4;
    if (c->status || c->accepted >= c->target || c->t > c->tbreak || vm > c->vbreak_hi || vm < c->vbreak_lo)
        c->done = 1;
}

}  // namespace hbi

// ------------------------------------------------------------------ object --
struct hbi_fv {
    hipStream_t stream = nullptr;
    hbi_hmat *hs = nullptr, *hn = nullptr;
    hbi_fv_opts opts{};
    hbi::FvPar P{};
    int n = 0, nblk = 0;
    hbi::DevBuf<double> a, b, dc, f0, vpl, kpmax, kT, phi, pfhyd, qin, area;
    hbi::DevBuf<double> y, ytmp, yout, k, vel, x, z, zn, ypart, ypart_n, hx, hx_n;
    hbi::DevBuf<double> pf, sigma, disp, velpre, qflow, errpart, monpart, hist, rows, work;
    hbi::DevBuf<double> io_pf, io_y, io_sig, io_out, io_dt;
    hbi::DevBuf<int> io_niter, zero;
    hbi::DevBuf<double> probe;                 // probe rows of the last step call (hbi_fv_set_probes)
    hbi::DevBuf<int> probe_idx;
    long long probe_rows = 0;
    hbi::DevBuf<hbi::FvCtl> ctl;
    // grid mode: the 2-D solve's work arrays, grid -> cluster map, wells, launch shape
    hbi::DevBuf<double> g2;
    hbi::DevBuf<unsigned> g2bar;   // barrier count, generation, failure flag
    hbi::DevBuf<int> cell, wi, wj, wk;
    hbi::DevBuf<double> wq, wt;
    int g2_blocks = 0, g2_npt = 0;
    hipGraphExec_t graph = nullptr;
    double *graph_hist = nullptr, *graph_probe = nullptr;
    bool dirty = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    double tbreak = HUGE_VAL, vbreak_hi = HUGE_VAL, vbreak_lo = 0.0;   // hbi_fv_set_break
};

namespace hbi {

static FvArrays fv_arrays(hbi_fv *f)
{
    FvArrays A{};
    A.a = f->a.p; A.b = f->b.p; A.dc = f->dc.p; A.f0 = f->f0.p; A.vpl = f->vpl.p; A.kpmax = f->kpmax.p;
    A.kT = f->kT.p; A.phi = f->phi.p; A.pfhyd = f->pfhyd.p; A.qin = f->qin.p; A.area = f->area.p;
    A.y = f->y.p; A.ytmp = f->ytmp.p; A.yout = f->yout.p; A.k = f->k.p; A.vel = f->vel.p; A.x = f->x.p;
    A.hx = f->hx.p; A.hx_n = f->hx_n.p;
    A.pf = f->pf.p; A.sigma = f->sigma.p; A.disp = f->disp.p; A.velpre = f->velpre.p; A.qflow = f->qflow.p;
    A.errpart = f->errpart.p; A.monpart = f->monpart.p;
    A.hist = f->hist.n ? f->hist.p : nullptr;
    A.probe = f->probe.n ? f->probe.p : nullptr;
    A.probe_idx = f->probe_idx.p;
    A.nprobe = (int)f->probe_idx.n;
    A.cell = f->cell.p;
    A.wi = f->wi.p; A.wj = f->wj.p; A.wk = f->wk.p; A.wq = f->wq.p; A.wt = f->wt.p;
    return A;
}

static Fv2Work fv2_work(hbi_fv *f)
{
    const size_t n = f->n;
    double *g = f->g2.p;
    unsigned *bar = f->g2bar.p;
    return Fv2Work{g, g + n, g + 7 * n, g + 8 * n, g + 9 * n, g + 10 * n, g + 11 * n, g + 12 * n,
                   g + 15 * n, g + 13 * n, g + 14 * n,
                   bar, bar ? (int *)(bar + kFv2Slot * (1 + (size_t)f->g2_blocks)) : nullptr};
}

// cells per thread and workgroups of the 2-D solve: the fewest cells per
// thread whose grid fits on the device at once
static void fv2_shape(hbi_fv *f)
{
    int dev = 0, cus = 0;
    HBI_HIP(hipGetDevice(&dev));
    HBI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const void *fn[5] = {(const void *)fv2_pressure_kernel<1>, (const void *)fv2_pressure_kernel<2>,
                         (const void *)fv2_pressure_kernel<4>, (const void *)fv2_pressure_kernel<8>,
                         (const void *)fv2_pressure_kernel<16>};
    for (int k = 0; k < 5; k++) {
        const int npt = 1 << k;
        int per = 0;
        HBI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn[k], kFv2Threads, 0));
        const long long need = ((long long)f->n + (long long)kFv2Threads * npt - 1) / ((long long)kFv2Threads * npt);
        if (need <= (long long)std::min(per, 1) * cus && (need <= kFv2MaxBlocks || npt == 16)) {
            f->g2_npt = npt;
            f->g2_blocks = (int)need;
            return;
        }
    }
    throw Error{HBI_E_ARG, "pore-pressure grid too large for one launch of resident workgroups"};
}

static void fv2_launch(hbi_fv *f, FvArrays A, FvSolveIO io)
{
    FvCtl *c = f->ctl.p;
    Fv2Work W = fv2_work(f);
    void *args[] = {&f->P, &c, &A, &io, &W};
    const void *fn = f->g2_npt == 1 ? (const void *)fv2_pressure_kernel<1>
                   : f->g2_npt == 2 ? (const void *)fv2_pressure_kernel<2>
                   : f->g2_npt == 4 ? (const void *)fv2_pressure_kernel<4>
                   : f->g2_npt == 8 ? (const void *)fv2_pressure_kernel<8>
                                    : (const void *)fv2_pressure_kernel<16>;
    // a plain launch: the barriers need every workgroup resident, which at
    // most one workgroup per CU gives; a cooperative launch would also
    // guarantee it, but takes the device's global wave-sync resource, which
    // one process at a time can hold (a second process's launch then fails)
    HBI_HIP(hipLaunchKernel(fn, dim3(f->g2_blocks), dim3(kFv2Threads), args, 0, f->stream));
}

static void fv_launch_pressure(hbi_fv *f, FvArrays A, const FvSolveIO &io)
{
    if (f->P.grid) {
        fv2_launch(f, A, io);
        return;
    }
    FvCtl *c = f->ctl.p;
    // elements per thread: registers hold x, r and A p of NPT elements, so
    // 2049-4096 elements take 512-thread workgroups (256 VGPRs per lane); up
    // to 4096 elements the matrix rows sit in LDS
    if (f->n <= 1024) fv_pressure_kernel<1024, 1, true><<<1, 1024, 0, f->stream>>>(f->P, c, A, io);
    else if (f->n <= 2048) fv_pressure_kernel<1024, 2, true><<<1, 1024, 0, f->stream>>>(f->P, c, A, io);
    else if (f->n <= 4096) fv_pressure_kernel<512, 8, true><<<1, 512, 0, f->stream>>>(f->P, c, A, io);
    else fv_pressure_kernel<1024, 8, false><<<1, 1024, 0, f->stream>>>(f->P, c, A, io);
}

// nstages derivs evaluations: mode 0 a whole attempt from y, mode 1 k_0 of ytmp
static void fv_launch_stages(hbi_fv *f, int nstages, int mode, const int *gate)
{
    const int g = f->nblk;
    FvArrays A = fv_arrays(f);
    const int prec = f->opts.prec;
    fv_stage0_kernel<<<g, kFvThreads, 0, f->stream>>>(f->P, gate, A, mode == 1);
    for (int s = 0; s < nstages; s++) {
        hmat_phase1(f->hs, prec, f->x.p, f->z.p, f->ypart.p, f->stream, gate);
        hmat_phase2(f->hs, prec, f->x.p, f->z.p, f->ypart.p, f->hx.p, f->stream, gate);
        if (f->hn) {
            hmat_phase1(f->hn, prec, f->x.p, f->zn.p, f->ypart_n.p, f->stream, gate);
            hmat_phase2(f->hn, prec, f->x.p, f->zn.p, f->ypart_n.p, f->hx_n.p, f->stream, gate);
        }
        fv_stage_kernel<<<g, kFvThreads, 0, f->stream>>>(f->P, gate, f->ctl.p, A, s, mode);
    }
}

// an attempt: part 1 (stages, decide, commit), the pressure step, part 2
// (fields, monitor)
static void fv_launch_part1(hbi_fv *f)
{
    const int g = f->nblk;
    FvArrays A = fv_arrays(f);
    FvCtl *c = f->ctl.p;
    fv_launch_stages(f, 6, 0, &c->done);
    fv_decide_kernel<<<1, 64, 0, f->stream>>>(f->P, c, f->errpart.p, g);
    fv_commit_kernel<<<g, kFvThreads, 0, f->stream>>>(f->P, c, A);
}

static void fv_launch_gated_pressure(hbi_fv *f)
{
    if (f->P.pfconst) return;
    FvCtl *c = f->ctl.p;
    FvSolveIO io{};
    io.pf = f->pf.p; io.y = f->y.p; io.sigma = f->sigma.p; io.pfnew = f->pf.p; io.rows = f->rows.p;
    io.work = f->work.p;
    io.niter = &c->niter; io.gated = 1;
    fv_launch_pressure(f, fv_arrays(f), io);
}

static void fv_launch_part2(hbi_fv *f)
{
    const int g = f->nblk;
    FvArrays A = fv_arrays(f);
    FvCtl *c = f->ctl.p;
    fv_fields_kernel<<<g, kFvThreads, 0, f->stream>>>(f->P, c, A);
    fv_monitor_kernel<<<1, 64, 0, f->stream>>>(f->P, c, A, g);
}

static void fv_capture_one(hbi_fv *f, hipGraphExec_t *out)
{
    if (*out) {
        (void)hipGraphExecDestroy(*out);
        *out = nullptr;
    }
    hipGraph_t gr;
    HBI_HIP(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
    fv_launch_part1(f);
    fv_launch_gated_pressure(f);
    fv_launch_part2(f);
    HBI_HIP(hipStreamEndCapture(f->stream, &gr));
    HBI_HIP(hipGraphInstantiate(out, gr, nullptr, nullptr, 0));
    (void)hipGraphDestroy(gr);
}

// the whole attempt (stages, decide, commit, the gated pressure step,
// fields, monitor) as one graph, in both modes
static void fv_capture(hbi_fv *f)
{
    fv_capture_one(f, &f->graph);
    f->graph_hist = f->hist.p;
    f->graph_probe = f->probe.p;
}

static void fv_launch_graphs(hbi_fv *f)
{
    HBI_HIP(hipGraphLaunch(f->graph, f->stream));
}

}  // namespace hbi

extern "C" {

int hbi_fv_create(hbi_hmat *hs, hbi_hmat *hn, const hbi_fv_opts *o, const double *a, const double *b,
                  const double *dc, const double *f0, const double *vpl, const double *kpmaxv, const double *kTv,
                  const double *phi, const double *pfhyd, const double *qin, const double *area, hbi_fv **out)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(hs && o && out && a && b && dc && f0 && vpl && kpmaxv && kTv && phi && pfhyd && qin && area,
                    "NULL argument");
        HBI_REQUIRE(!hn || hn->n == hs->n, "normal-stress matrix size differs");
        HBI_REQUIRE(o->prec >= 0 && o->prec <= 2, "bad precision");
        HBI_REQUIRE(o->eps_r > 0 && o->ds0 > 0, "eps_r and ds0 must be > 0");
        HBI_REQUIRE((o->bcl == HBI_FV_NEUMANN || o->bcl == HBI_FV_DIRICHLET) &&
                        (o->bcr == HBI_FV_NEUMANN || o->bcr == HBI_FV_DIRICHLET),
                    "bad boundary condition");
        HBI_REQUIRE(o->setting >= HBI_FV_INJECTIONQ && o->setting <= HBI_FV_OTHER, "bad setting");
        const int n = hs->n;
        const bool grid = o->imax > 0 || o->jmax > 0;
        if (grid) {
            HBI_REQUIRE(o->imax >= 4 && o->jmax >= 4 && (long long)o->imax * o->jmax == n,
                        "grid mode: imax, jmax >= 4 and imax jmax = N");
            HBI_REQUIRE((o->bct == HBI_FV_NEUMANN || o->bct == HBI_FV_DIRICHLET) &&
                            (o->bcb == HBI_FV_NEUMANN || o->bcb == HBI_FV_DIRICHLET),
                        "bad boundary condition");
            HBI_REQUIRE(o->dtmax > 0 && o->phi0 > 0, "grid mode: dtmax and phi0 must be > 0");
            HBI_REQUIRE(o->setting != HBI_FV_INJECTIONP, "grid mode: no injection-pressure setting in main_fv3dp");
            HBI_REQUIRE((int)hs->part.lod.size() == n, "matrix without its cluster permutation");
        } else {
            HBI_REQUIRE(n >= 4 && n <= HBI_FV_MAX_CELLS, "the pore-pressure grid needs 4 to 8192 elements");
            HBI_REQUIRE(o->setting != HBI_FV_THRUST || n >= 2000, "setting thrust reads qin(2000): needs N >= 2000");
        }
        *out = nullptr;
        hbi_fv *f = new hbi_fv();
        try {
            f->hs = hs;
            f->hn = hn;
            f->opts = *o;
            f->n = n;
            f->nblk = (n + hbi::kFvThreads - 1) / hbi::kFvThreads;
            HBI_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
            HBI_HIP(hipEventCreate(&f->e0));
            HBI_HIP(hipEventCreate(&f->e1));
            hbi::FvPar &P = f->P;
            P.n = n; P.law = o->law; P.backslip = o->backslip; P.has_hn = hn ? 1 : 0;
            P.pdep = o->pressure_dependent; P.pfconst = o->pfconst; P.bcl = o->bcl; P.bcr = o->bcr;
            P.setting = o->setting; P.outpertime = o->outpertime;
            P.vref = o->vref; P.eps = o->eps_r; P.hdiv = o->hdiv; P.ds0 = o->ds0; P.kL = o->kL; P.kpmin = o->kpmin;
            P.s0 = o->s0; P.eta = o->eta; P.beta = o->beta; P.Bs = o->Bs; P.pbcl = o->pbcl; P.pbcr = o->pbcr;
            P.q0 = o->q0; P.pinj = o->pinj; P.tinj = o->tinj_s; P.dtout = o->dtout_s;
            P.velmin = o->velmin; P.velmax = o->velmax; P.tmax = o->tmax_s;
            P.grid = grid ? 1 : 0; P.imax = o->imax; P.jmax = o->jmax; P.bct = o->bct; P.bcb = o->bcb;
            P.nwell = 0; P.dtmax = o->dtmax; P.minsig = o->minsig; P.phi0 = o->phi0;
            P.T = hbi::rk_tableau();
            double sg = 0.0;
            for (int i = 0; i < n; i++) sg = sg + area[i];
            P.sg = sg;
            hipStream_t s = f->stream;
            f->a.upload(a, n, s); f->b.upload(b, n, s); f->dc.upload(dc, n, s); f->f0.upload(f0, n, s);
            f->vpl.upload(vpl, n, s); f->kpmax.upload(kpmaxv, n, s); f->kT.upload(kTv, n, s);
            f->phi.upload(phi, n, s); f->pfhyd.upload(pfhyd, n, s); f->qin.upload(qin, n, s);
            f->area.upload(area, n, s);
            const size_t m4 = 4 * (size_t)n;
            f->y.alloc(m4); f->ytmp.alloc(m4); f->yout.alloc(m4); f->k.alloc(6 * m4);
            f->vel.alloc(n); f->x.alloc(n + 8); f->hx.alloc(n);
            HBI_HIP(hipMemsetAsync(f->x.p, 0, (n + 8) * 8, s));
            HBI_HIP(hipMemsetAsync(f->k.p, 0, 6 * m4 * 8, s));
            f->z.alloc(hbi::hmat_z_size(hs, o->prec));
            f->ypart.alloc(hbi::hmat_ypart_size(hs, o->prec));
            if (hn) {
                f->zn.alloc(hbi::hmat_z_size(hn, o->prec));
                f->ypart_n.alloc(hbi::hmat_ypart_size(hn, o->prec));
                f->hx_n.alloc(n);
            }
            f->pf.alloc(n); f->sigma.alloc(n); f->disp.alloc(n); f->velpre.alloc(n); f->qflow.alloc(n);
            HBI_HIP(hipMemsetAsync(f->pf.p, 0, n * 8, s));
            HBI_HIP(hipMemsetAsync(f->disp.p, 0, n * 8, s));
            HBI_HIP(hipMemsetAsync(f->qflow.p, 0, n * 8, s));
            f->rows.alloc(3 * (size_t)n);
            f->work.alloc(3 * (size_t)n);
            f->errpart.alloc(f->nblk);
            f->monpart.alloc(hbi::kFvMon * (size_t)f->nblk);
            f->io_pf.alloc(n); f->io_y.alloc(m4); f->io_sig.alloc(n); f->io_out.alloc(n); f->io_dt.alloc(1);
            f->io_niter.alloc(1);
            f->zero.alloc(1);
            HBI_HIP(hipMemsetAsync(f->zero.p, 0, sizeof(int), s));
            if (grid) {
                std::vector<int> cell(n);
                for (int e = 0; e < n; e++) cell[hs->part.lod[e]] = e;
                f->cell.upload(cell, s);
                hbi::fv2_shape(f);
                f->g2.alloc(15 * (size_t)n + 2 * (size_t)f->g2_blocks);
                const size_t nbar = hbi::kFv2Slot * (2 + (size_t)f->g2_blocks);
                f->g2bar.alloc(nbar);
                HBI_HIP(hipMemsetAsync(f->g2bar.p, 0, nbar * sizeof(unsigned), s));
            }
            f->ctl.alloc(1);
            hbi::FvCtl c{};
            c.h = 1.0;
            c.nn = 2;
            hbi::to_device_sync(f->ctl.p, &c, sizeof(c), s);
        } catch (...) {
            hbi_fv_destroy(f);
            throw;
        }
        *out = f;
    });
}

int hbi_fv_destroy(hbi_fv *f)
{
    if (!f) return HBI_OK;
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    if (f->graph) (void)hipGraphExecDestroy(f->graph);
    if (f->e0) (void)hipEventDestroy(f->e0);
    if (f->e1) (void)hipEventDestroy(f->e1);
    hipStream_t s = f->stream;
    delete f;
    if (s) (void)hipStreamDestroy(s);
    return HBI_OK;
}

int hbi_fv_set_state(hbi_fv *f, const double *y, const double *pf, const double *disp, double t, double dt_next)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && y && pf, "NULL argument");
        const int n = f->n;
        f->y.upload(y, 4 * (size_t)n, f->stream);
        f->pf.upload(pf, n, f->stream);
        if (disp) f->disp.upload(disp, n, f->stream);
        else HBI_HIP(hipMemsetAsync(f->disp.p, 0, n * 8, f->stream));
        hbi::FvCtl c{};
        c.nn = 2;
        hbi::to_device_sync(f->ctl.p, &c, sizeof(c), f->stream);
        // V of the state (main_fv.f90:995-1003)
        hbi::fv_stage0_kernel<<<f->nblk, hbi::kFvThreads, 0, f->stream>>>(f->P, f->zero.p, hbi::fv_arrays(f), 0);
        HBI_HIP(hipGetLastError());
        c.h = dt_next;
        c.t = t;
        hbi::to_device_sync(f->ctl.p, &c, sizeof(c), f->stream);
        // grid mode: sigma (total, held through the steps) = sigma_e + pf of the state
        if (f->P.grid) {
            std::vector<double> sg(n);
            for (int i = 0; i < n; i++) sg[i] = y[4 * (size_t)i + 2] + pf[i];
            f->sigma.upload(sg.data(), n, f->stream);
        }
        HBI_HIP(hipStreamSynchronize(f->stream));
    });
}

int hbi_fv_set_wells(hbi_fv *f, int nwell, const int *iw, const int *jw, const int *kleng, const double *qvals,
                     const double *qtimes)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && nwell >= 0 && (nwell == 0 || (iw && jw && kleng && qvals && qtimes)), "bad arguments");
        HBI_REQUIRE(f->P.grid || nwell == 0, "wells need the 2-D grid mode");
        for (int w = 0; w < nwell; w++) {
            HBI_REQUIRE(iw[w] >= 1 && iw[w] <= f->P.imax && jw[w] >= 1 && jw[w] <= f->P.jmax, "well outside the grid");
            HBI_REQUIRE(kleng[w] >= 1 && kleng[w] <= HBI_FV_MAX_SCHEDULE, "schedule length 1..50");
        }
        HBI_HIP(hipStreamSynchronize(f->stream));
        if (nwell) {
            f->wi.upload(iw, nwell, f->stream);
            f->wj.upload(jw, nwell, f->stream);
            f->wk.upload(kleng, nwell, f->stream);
            f->wq.upload(qvals, (size_t)nwell * HBI_FV_MAX_SCHEDULE, f->stream);
            f->wt.upload(qtimes, (size_t)nwell * HBI_FV_MAX_SCHEDULE, f->stream);
        }
        f->P.nwell = nwell;
        hbi::FvCtl c;
        HBI_HIP(hipMemcpy(&c, f->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
        c.sw = 0;
        c.nn = 2;
        hbi::to_device_sync(f->ctl.p, &c, sizeof(c), f->stream);
        HBI_HIP(hipStreamSynchronize(f->stream));
        f->dirty = true;
    });
}

int hbi_fv_get_state(hbi_fv *f, double *y, double *pf, double *disp, double *vel, double *qflow, hbi_fv_info *info)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f, "NULL argument");
        const int n = f->n;
        hbi::FvCtl c;
        HBI_HIP(hipMemcpyAsync(&c, f->ctl.p, sizeof(c), hipMemcpyDeviceToHost, f->stream));
        if (y) f->y.download(y, 4 * (size_t)n, f->stream);
        if (pf) f->pf.download(pf, n, f->stream);
        if (disp) f->disp.download(disp, n, f->stream);
        if (vel) f->vel.download(vel, n, f->stream);
        if (qflow) f->qflow.download(qflow, n, f->stream);
        HBI_HIP(hipStreamSynchronize(f->stream));
        if (info) {
            info->t = c.t; info->dt_next = c.h; info->dt_last = c.hdid; info->errmax_last = c.errmax;
            info->attempts = c.attempts; info->accepted = c.accepted; info->rejected = c.rejected;
            info->vmax = c.vmax; info->niter = c.niter; info->status = c.status;
        }
    });
}

int hbi_fv_derivs(hbi_fv *f, const double *y, double *dydx)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && y && dydx, "NULL argument");
        const size_t m4 = 4 * (size_t)f->n;
        hbi::to_device_sync(f->ytmp.p, y, m4 * 8, f->stream);
        hbi::fv_launch_stages(f, 1, 1, f->zero.p);
        HBI_HIP(hipGetLastError());
        HBI_HIP(hipMemcpyAsync(dydx, f->k.p, m4 * 8, hipMemcpyDeviceToHost, f->stream));
        HBI_HIP(hipStreamSynchronize(f->stream));
    });
}

int hbi_fv_diffuse(hbi_fv *f, const double *pf, const double *y, const double *sigma, double h, double t,
                   double *pfnew, int *niter, double *dtnxt)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && pf && y && sigma && pfnew && niter && dtnxt, "NULL argument");
        HBI_REQUIRE(h > 0, "h must be > 0");
        const int n = f->n;
        f->io_pf.upload(pf, n, f->stream);
        f->io_y.upload(y, 4 * (size_t)n, f->stream);
        f->io_sig.upload(sigma, n, f->stream);
        f->io_dt.upload(dtnxt, 1, f->stream);
        hbi::FvSolveIO io{};
        io.pf = f->io_pf.p; io.y = f->io_y.p; io.sigma = f->io_sig.p; io.pfnew = f->io_out.p; io.rows = f->rows.p;
        io.work = f->work.p;
        io.h = h; io.t = t; io.dtnxt = f->io_dt.p; io.niter = f->io_niter.p; io.gated = 0;
        hbi::fv_launch_pressure(f, hbi::fv_arrays(f), io);
        HBI_HIP(hipGetLastError());
        if (f->P.grid) {
            int failed = 0;
            HBI_HIP(hipMemcpyAsync(&failed, hbi::fv2_work(f).fail, sizeof(int), hipMemcpyDeviceToHost, f->stream));
            HBI_HIP(hipStreamSynchronize(f->stream));
            if (failed) throw hbi::Error{HBI_E_HIP, "pore-pressure solve: a grid barrier timed out"};
        }
        f->io_out.download(pfnew, n, f->stream);
        f->io_niter.download(niter, 1, f->stream);
        f->io_dt.download(dtnxt, 1, f->stream);
        HBI_HIP(hipStreamSynchronize(f->stream));
    });
}

int hbi_fv_step(hbi_fv *f, int nsteps, double *hist, hbi_fv_info *info)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && nsteps >= 0, "bad arguments");
        const size_t hneed = hist ? hbi::kFvHist * (size_t)nsteps : 0;
        bool recapture = !f->graph;
        if (hneed > f->hist.n || (hneed == 0 && f->hist.n)) {
            if (hneed) f->hist.alloc(hneed); else f->hist.release();
            recapture = true;
        }
        const size_t pneed = (size_t)nsteps * f->probe_idx.n * hbi::kFvProbe;
        if (pneed > f->probe.n || (pneed == 0 && f->probe.n)) {
            if (pneed) f->probe.alloc(pneed); else f->probe.release();
            recapture = true;
        }
        f->probe_rows = 0;
        if (f->dirty) recapture = true;
        f->dirty = false;
        if (recapture) hbi::fv_capture(f);
        hbi::fv_arm_kernel<<<1, 1, 0, f->stream>>>(f->ctl.p, nsteps, f->tbreak, f->vbreak_hi, f->vbreak_lo);
        HBI_HIP(hipEventRecord(f->e0, f->stream));
        hbi::FvCtl c;
        long long launched = 0;
        const long long budget = 50LL * nsteps + 100;
        int batch = nsteps;
        for (;;) {
            for (int q = 0; q < batch; q++) hbi::fv_launch_graphs(f);
            launched += batch;
            HBI_HIP(hipMemcpyAsync(&c, f->ctl.p, sizeof(c), hipMemcpyDeviceToHost, f->stream));
            HBI_HIP(hipStreamSynchronize(f->stream));
            if (c.done || launched >= budget || nsteps == 0) break;
            batch = (int)std::max<long long>(1, std::min<long long>(c.target - c.accepted + 2, budget - launched));
        }
        HBI_HIP(hipEventRecord(f->e1, f->stream));
        HBI_HIP(hipEventSynchronize(f->e1));
        float ms = 0;
        HBI_HIP(hipEventElapsedTime(&ms, f->e0, f->e1));
        f->last_ms = ms;
        const long long done = c.accepted - c.hist_base;
        f->probe_rows = std::max<long long>(0, std::min<long long>(done, nsteps));
        if (hist && done > 0) {
            f->hist.download(hist, hbi::kFvHist * (size_t)std::min<long long>(done, nsteps), f->stream);
            HBI_HIP(hipStreamSynchronize(f->stream));
        }
        if (info) {
            info->t = c.t; info->dt_next = c.h; info->dt_last = c.hdid; info->errmax_last = c.errmax;
            info->attempts = c.attempts; info->accepted = c.accepted; info->rejected = c.rejected;
            info->vmax = c.vmax; info->niter = c.niter; info->status = c.status;
        }
        if (c.status == HBI_E_STEP) throw hbi::Error{HBI_E_STEP, "Runge-Kutta step size underflow"};
        if (c.status == hbi::kFv2Failed) throw hbi::Error{HBI_E_HIP, "pore-pressure solve: a grid barrier timed out"};
    });
}

int hbi_fv_set_break(hbi_fv *f, double t, double vmax_above, double vmax_below)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && !std::isnan(t) && !std::isnan(vmax_above) && !std::isnan(vmax_below), "bad arguments");
        f->tbreak = t;
        f->vbreak_hi = vmax_above;
        f->vbreak_lo = vmax_below;
    });
}

int hbi_fv_set_probes(hbi_fv *f, int nprobe, const int *idx)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && nprobe >= 0 && nprobe <= hbi::kFvMaxProbes && (nprobe == 0 || idx), "bad arguments");
        for (int q = 0; q < nprobe; q++) HBI_REQUIRE(idx[q] >= 0 && idx[q] < f->n, "probe index out of range");
        if (nprobe) {
            f->probe_idx.alloc(nprobe);
            f->probe_idx.upload(idx, nprobe, f->stream);
        } else {
            f->probe_idx.release();
        }
        f->probe.release();
        f->probe_rows = 0;
        f->dirty = true;
    });
}

int hbi_fv_get_probes(hbi_fv *f, long long *nrows, double *rows)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && nrows, "NULL argument");
        *nrows = f->probe_rows;
        if (rows && f->probe_rows > 0) {
            f->probe.download(rows, (size_t)f->probe_rows * f->probe_idx.n * hbi::kFvProbe, f->stream);
            HBI_HIP(hipStreamSynchronize(f->stream));
        }
    });
}

int hbi_fv_last_ms(const hbi_fv *f, double *ms)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(f && ms, "NULL argument");
        *ms = f->last_ms;
    });
}

}  // extern "C"
