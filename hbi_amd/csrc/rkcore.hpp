// rkcore.hpp -- pieces of the rate-and-state ODE shared by the single-run
// integrator (rk.hip) and the ensemble integrator (ens.hip): the Cash-Karp
// tableau, the slip-rate law and the state derivative.
#pragma once
#include "hmat.hpp"

namespace hbi {

struct Tab {   // Cash-Karp tableau of main_LH.f90:2490-2496 (see rk_tableau)
    double a2, a3, a4, a5, a6, b21, b31, b32, b41, b42, b43, b51, b52, b53, b54,
        b61, b62, b63, b64, b65, c1, c3, c4, c6, dc1, dc3, dc4, dc5, dc6;
};

// The reference declares these real(8) parameters but writes them as
// default-real (single precision) literals, so b_ij, c_i, dc_i carry
// single-precision rounding; A2..A6 and B21 are double literals.
inline Tab rk_tableau()
{
    Tab t;
    t.a2 = .2; t.a3 = .3; t.a4 = .6; t.a5 = 1.0; t.a6 = .875; t.b21 = .2;
    t.b31 = (double)(3.f / 40.f); t.b32 = (double)(9.f / 40.f); t.b41 = (double).3f; t.b42 = (double)-.9f;
    t.b43 = (double)1.2f; t.b51 = (double)(-11.f / 54.f); t.b52 = (double)2.5f; t.b53 = (double)(-70.f / 27.f);
    t.b54 = (double)(35.f / 27.f); t.b61 = (double)(1631.f / 55296.f); t.b62 = (double)(175.f / 512.f);
    t.b63 = (double)(575.f / 13824.f); t.b64 = (double)(44275.f / 110592.f); t.b65 = (double)(253.f / 4096.f);
    t.c1 = (double)(37.f / 378.f); t.c3 = (double)(250.f / 621.f); t.c4 = (double)(125.f / 594.f);
    t.c6 = (double)(512.f / 1771.f);
    t.dc1 = t.c1 - (double)(2825.f / 27648.f); t.dc3 = t.c3 - (double)(18575.f / 48384.f);
    t.dc4 = t.c4 - (double)(13525.f / 55296.f); t.dc5 = (double)(-277.f / 14336.f); t.dc6 = t.c6 - (double).25f;
    return t;
}

// V = 2 V0 exp(-psi/a) sinh(tau/(sigma a))   main_LH.f90:2084
__device__ __forceinline__ double slip_rate(double psi, double tau, double sig, double a, double vref)
{
#pragma clang fp contract(off)
    return 2 * vref * exp(-psi / a) * sinh(tau / sig / a);
}

// deriv, main_LH.f90:2165-2244 (aging and slip laws)
template <class Par>
__device__ __forceinline__ void deriv(const Par &P, double gs, double gn, double psi, double tau, double sig, double vel,
                                      double a, double b, double dc, double f0, double *o)
{
#pragma clang fp contract(off)
    const double rigid = kRigid, vs = kVs, vref = P.vref;
    double dsigdt = gn;
    if (P.limitsigma) {
        if (sig < P.minsig) dsigdt = 0;
        if (sig > P.maxsig) dsigdt = 0;
    }
    double dpsidt;
    if (P.law == 0) {
        if (b == 0) dpsidt = 0.0;
        else dpsidt = b / dc * vref * exp((f0 - psi) / b) - b * fabs(vel) / dc;
    } else {
        const double fss = f0 + (a - b) * log(fabs(vel) / vref);
        dpsidt = -fabs(vel) / dc * (fabs(tau) / sig - fss);
    }
    const double e = exp(-psi / a), ch = cosh(tau / sig / a);
    const double dvdtau = 2 * vref * e * ch / (a * sig);
    const double dvdsig = -2 * vref * e * ch * tau / (a * (sig * sig));
    const double dvdpsi = -vel / a;
    double dtaudt = gs - 0.5 * rigid / vs * (dvdpsi * dpsidt + dvdsig * dsigdt);
    dtaudt = dtaudt / (1.0 + 0.5 * rigid / vs * dvdtau);
    o[0] = dpsidt;
    o[1] = dtaudt;
    o[2] = dsigdt;
}

}  // namespace hbi
