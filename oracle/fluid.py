"""CPU restatement of the fluid-driven cycle of main_fv.f90 (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the cpu_baseline leg of
bench.py, never by the product package.

Follows, line for line in its arithmetic (left-to-right operation order,
single-precision literals where the reference writes them, sequential sums):
  derivs          main_fv.f90:2039-2146   4-component RHS (psi, tau, sigma_e, ks)
  rkck            main_fv.f90:2259-2342   Cash-Karp stages (tableau of main_LH)
  rkqs            main_fv.f90:2149-2256   error max |yerr / yout| / eps over 4N
  implicitsolver2 main_fv.f90:1727-1774   pressure-independent diffusivity
  implicitsolver  main_fv.f90:1678-1724   Picard iterations on the diffusivity
  Beuler          main_fv.f90:1776-1940   backward Euler + conjugate gradients
  step            main_fv.f90:1009-1152   post-RK update, pressure step, fields,
                                          monitor values and stop criteria
The H-matrix products come from the C oracle (oracle_lib.HMatrix).

How each piece is pinned (the reference has no test or fixture for this
driver; tests/test_fluid_oracle.py):
  * derivs: the psi, tau, sigma_e components equal the C oracle's main_LH
    derivs (already pinned against the compiled reference) on the same state;
    dks/dt vanishes at the steady permeability of main_fv.f90:1420;
  * rkck: its psi, tau, sigma_e components equal the C oracle's rkck;
  * Beuler: the assembled system conserves the half-weighted pressure sum
    under Neumann boundaries without sources (the column sums of the
    finite-volume operator vanish), its Dirichlet steady state is the linear
    profile through the boundary pressures, the injection-rate source adds
    exactly its volume while t < tinj, the injection-pressure penalty pulls
    the centre to pinj; the conjugate-gradient result agrees with a direct
    banded solve of the same system to the residual bound;
  * implicitsolver: the Picard result is a fixed point to the 1e-3 test.
  * rkqs: the controller of main_fv.f90:2208-2248 by hand-worked values
    (rejection to max(h/2, 0.9f h e^-0.25), a tiny errmax halving h, growth
    min(2h, 0.9f h e^-0.2f), the outpertime cap, underflow);
  * the post-step update on a scripted step: sigma held through the
    pressure step, the slip's trapezoid of V before and after it, Darcy's
    qflow of the new pf (one-sided at the ends).
"""
from __future__ import annotations

import numpy as np

RIGID = 32.04     # m_const.f90:7
VS = 3.464        # m_const.f90:8
YEAR_S = 365 * 24 * 3600

NEUMANN, DIRICHLET = 0, 1
INJECTIONQ, INJECTIONP, STEADYSTATE, THRUST, OTHER = 0, 1, 2, 3, 4


def _f32(x):
    return float(np.float32(x))


def _q(num, den):
    """A default-real quotient literal of the reference (num./den.)."""
    return float(np.float32(num) / np.float32(den))


def tableau():
    """Cash-Karp coefficients as main_fv.f90:2275-2283 declares them: A2..A6
    and B21 double literals, the others single-precision quotients."""
    t = dict(a2=.2, a3=.3, a4=.6, a5=1.0, a6=.875, b21=.2, b31=_q(3, 40), b32=_q(9, 40), b41=_f32(.3),
             b42=_f32(-.9), b43=_f32(1.2), b51=_q(-11, 54), b52=_f32(2.5), b53=_q(-70, 27), b54=_q(35, 27),
             b61=_q(1631, 55296), b62=_q(175, 512), b63=_q(575, 13824), b64=_q(44275, 110592), b65=_q(253, 4096),
             c1=_q(37, 378), c3=_q(250, 621), c4=_q(125, 594), c6=_q(512, 1771))
    t["dc1"] = t["c1"] - _q(2825, 27648)
    t["dc3"] = t["c3"] - _q(18575, 48384)
    t["dc4"] = t["c4"] - _q(13525, 55296)
    t["dc5"] = _q(-277, 14336)
    t["dc6"] = t["c6"] - _f32(.25)
    return t


def seqsum(a) -> float:
    """Sequential left-to-right sum (Fortran sum)."""
    a = np.asarray(a, dtype=np.float64)
    return float(np.cumsum(a)[-1]) if a.size else 0.0


class FvOracle:
    """main_fv.f90 on one fault: hs, hn oracle H-matrices (hn None when sigma
    is constant), opts as hbi_fv_opts, arrays (cluster order) as hbi_fv_create."""

    def __init__(self, hs, opts: dict, arrays: dict, hn=None):
        self.hs, self.hn = hs, hn
        self.o = dict(opts)
        self.arr = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in arrays.items()}
        self.n = hs.n
        self.T = tableau()
        self.sg = seqsum(self.arr["area"])

    # ------------------------------------------------------------ derivs --
    def vel(self, psi, tau, sig):
        a, vref = self.arr["a"], self.o["vref"]
        return 2 * vref * np.exp(-psi / a) * np.sinh(tau / sig / a)

    def derivs(self, y):
        """main_fv.f90:2039-2146."""
        o, A = self.o, self.arr
        a, b, dc, f0 = A["a"], A["b"], A["dc"], A["f0"]
        vref = o["vref"]
        psi, tau, sig, ks = y[0::4], y[1::4], y[2::4], y[3::4]
        v = self.vel(psi, tau, sig)
        x = v - A["vpl"] if o["backslip"] else v
        gs = self.hs.matvec(x)
        if o["hdiv"] != 0.0:
            gs = gs / o["hdiv"]
        if self.hn is not None:
            gn = self.hn.matvec(x)
            if o["hdiv"] != 0.0:
                gn = gn / o["hdiv"]
        else:
            gn = np.zeros(self.n)
        dks = -v / o["kL"] * (ks - A["kpmaxv"]) - (ks - o["kpmin"]) / A["kTv"]
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            if o["law"] == 0:
                dpsi = np.where(b == 0, 0.0, b / dc * vref * np.exp((f0 - psi) / b) - b * v / dc)
            else:
                fss = f0 + (a - b) * np.log(np.abs(v) / vref)
                dpsi = -np.abs(v) / dc * (np.abs(tau) / sig - fss)
        dsig = (1 - o["Bs"]) * gn
        dvdtau = 2 * vref * np.exp(-psi / a) * np.cosh(tau / sig / a) / (a * sig)
        dvdsig = -2 * vref * np.exp(-psi / a) * np.cosh(tau / sig / a) * tau / (a * (sig * sig))
        dvdpsi = -v / a
        dtau = gs - 0.5 * RIGID / VS * (dvdpsi * dpsi + dvdsig * dsig)
        dtau = dtau / (1.0 + 0.5 * RIGID / VS * dvdtau)
        d = np.empty(4 * self.n)
        d[0::4], d[1::4], d[2::4], d[3::4] = dpsi, dtau, dsig, dks
        return d

    def rkck(self, y, h):
        """main_fv.f90:2259-2342: (yout, yerr)."""
        T = self.T
        k1 = self.derivs(y)
        k2 = self.derivs(y + T["b21"] * h * k1)
        k3 = self.derivs(y + h * (T["b31"] * k1 + T["b32"] * k2))
        k4 = self.derivs(y + h * (T["b41"] * k1 + T["b42"] * k2 + T["b43"] * k3))
        k5 = self.derivs(y + h * (T["b51"] * k1 + T["b52"] * k2 + T["b53"] * k3 + T["b54"] * k4))
        k6 = self.derivs(y + h * (T["b61"] * k1 + T["b62"] * k2 + T["b63"] * k3 + T["b64"] * k4 + T["b65"] * k5))
        yout = y + h * (T["c1"] * k1 + T["c3"] * k3 + T["c4"] * k4 + T["c6"] * k6)
        yerr = h * (T["dc1"] * k1 + T["dc3"] * k3 + T["dc4"] * k4 + T["dc5"] * k5 + T["dc6"] * k6)
        return yout, yerr

    def rkqs(self, y, t, htry):
        """main_fv.f90:2149-2256: (yout, t + hdid, hdid, hnext, errmax, rejected)."""
        eps = self.o["eps_r"]
        SAFETY, PGROW, PSHRNK = _f32(0.9), _f32(-0.2), -0.25
        h = htry
        nrej = 0
        while True:
            yout, yerr = self.rkck(y, h)
            with np.errstate(divide="ignore", invalid="ignore"):
                r = np.abs(yerr / yout) / eps
            r = r[~np.isnan(r)]
            errmax = max(0.0, float(r.max())) if r.size else 0.0
            if errmax < 1.0 and errmax > 1e-15:
                break
            if errmax > 1e-15:
                h = max(0.5 * h, SAFETY * h * errmax ** PSHRNK)
            else:
                h = 0.5 * h
            nrej += 1
            if (t + h) - t < 1e-15:
                raise FloatingPointError("dt is too small")
        hnext = min(2 * h, SAFETY * h * errmax ** PGROW)
        if self.o["outpertime"]:
            hnext = min(hnext, self.o["dtout_s"])
        return yout, t + h, h, hnext, errmax, nrej

    # -------------------------------------------------------- pore pressure --
    def beuler_system(self, pf, cdiff, h, t):
        """The tridiagonal system of Beuler (main_fv.f90:1786-1893) after the
        halving of its end rows: (lo, di, up, b, x0) with row e reading
        lo x[e-1] + di x[e] + up x[e+1]."""
        o, A = self.o, self.arr
        n = self.n
        ds0 = o["ds0"]
        c = np.asarray(cdiff, dtype=np.float64)
        phi, pfhyd, qin = A["phi"], A["pfhyd"], A["qin"]
        D = np.zeros((n, 3))
        if o["bcl"] == DIRICHLET:
            D[0, 0] = -c[0] - c[1]
            D[0, 1] = c[1] - c[0]
            D[1] = (c[1] / 2 - c[0] / 2, -c[0] / 2 - c[1] - c[2] / 2, c[1] / 2 + c[2] / 2)
        else:
            D[0, 0] = -c[0] - c[1]
            D[0, 1] = -D[0, 0]
            D[1] = (c[0] / 2 + c[1] / 2, -c[0] / 2 - c[1] - c[2] / 2, c[1] / 2 + c[2] / 2)
        i = np.arange(2, n - 2)
        D[i, 0] = c[i - 1] / 2 + c[i] / 2
        D[i, 1] = -c[i - 1] / 2 - c[i] - c[i + 1] / 2
        D[i, 2] = c[i] / 2 + c[i + 1] / 2
        if o["bcr"] == DIRICHLET:
            D[n - 2] = (c[n - 3] / 2 + c[n - 2] / 2, -c[n - 3] / 2 - c[n - 2] - c[n - 1] / 2,
                        -c[n - 1] / 2 + c[n - 2] / 2)
            D[n - 1, 2] = -c[n - 1] - c[n - 2]
            D[n - 1, 1] = c[n - 2] - c[n - 1]
        else:
            D[n - 2] = (c[n - 3] / 2 + c[n - 2] / 2, -c[n - 3] / 2 - c[n - 2] - c[n - 1] / 2,
                        c[n - 2] / 2 + c[n - 1] / 2)
            D[n - 1, 2] = -c[n - 1] - c[n - 2]
            D[n - 1, 1] = -D[n - 1, 2]
        D = D / ds0 / ds0
        Am = np.zeros((n, 3))
        Am[0, 0] = 1.0 - h * D[0, 0]
        Am[0, 1] = -h * D[0, 1]
        Am[n - 1, 2] = 1.0 - h * D[n - 1, 2]
        Am[n - 1, 1] = -h * D[n - 1, 1]
        Am[1:n - 1, 0] = -h * D[1:n - 1, 0]
        Am[1:n - 1, 1] = 1.0 - h * D[1:n - 1, 1]
        Am[1:n - 1, 2] = -h * D[1:n - 1, 2]
        sat = np.zeros(n)
        e9 = _f32(1e-9)
        if o["bcl"] == DIRICHLET:
            sat[0] = -1.0 * c[0] / ds0 / ds0 * (o["pbcl"] - pfhyd[0]) * h * 2
            sat[1] = -0.5 * c[0] / ds0 / ds0 * (o["pbcl"] - pfhyd[0]) * h * 2
        elif o["setting"] == STEADYSTATE:
            sat[0] = -o["q0"] / o["beta"] / phi[0] * e9 * h / ds0 * 2
        if o["bcr"] == DIRICHLET:
            sat[n - 1] = -1 * c[n - 1] / ds0 / ds0 * (o["pbcr"] - pfhyd[n - 1]) * h * 2
            sat[n - 2] = sat[n - 1] / 2
        else:
            if o["setting"] == STEADYSTATE:
                sat[n - 1] = o["q0"] / o["beta"] / phi[n - 1] * e9 * h / ds0 * 2
            if o["setting"] == THRUST:
                sat[n - 1] = qin[1999] / o["beta"] / phi[n - 1] * e9 * h / ds0 * 2
        if o["setting"] == THRUST:
            k = np.arange(2, n - 2)
            sat[k] = -qin[k] / o["beta"] / phi[k] * e9 * h / ds0
        x0 = pf - pfhyd
        b = pf - sat - pfhyd
        hf = n // 2 - 1
        if o["setting"] == INJECTIONP:
            b[hf] = b[hf] + h * o["pinj"] / 1e1
            Am[hf, 1] = Am[hf, 1] + h / 1e1
        elif o["setting"] == INJECTIONQ and t < o["tinj_s"]:
            b[hf] = b[hf] + h * qin[hf] / o["beta"] / phi[hf] * e9 / ds0
        b[0] = b[0] / 2
        b[n - 1] = b[n - 1] / 2
        Am[0] = Am[0] / 2
        Am[n - 1] = Am[n - 1] / 2
        lo = np.concatenate([[0.0], Am[1:n - 1, 0], [Am[n - 1, 1]]])
        di = np.concatenate([[Am[0, 0]], Am[1:n - 1, 1], [Am[n - 1, 2]]])
        up = np.concatenate([[Am[0, 1]], Am[1:n - 1, 2], [0.0]])
        return lo, di, up, b, x0

    @staticmethod
    def apply(lo, di, up, v):
        """The product in the reference's order of terms (main_fv.f90:1895-1903)."""
        n = v.size
        m = np.empty(n)
        m[0] = v[0] * di[0] + v[1] * up[0]
        m[1:n - 1] = v[0:n - 2] * lo[1:n - 1] + v[1:n - 1] * di[1:n - 1] + v[2:n] * up[1:n - 1]
        m[n - 1] = v[n - 1] * di[n - 1] + v[n - 2] * lo[n - 1]
        return m

    def beuler(self, pf, cdiff, h, t):
        """main_fv.f90:1776-1940: (pfnew, CG iterations)."""
        lo, di, up, b, x = self.beuler_system(pf, cdiff, h, t)
        n = self.n
        tol = _f32(1e-4)
        thr = tol ** 2 * n
        r = b - self.apply(lo, di, up, x)
        p = r.copy()
        rsold = seqsum(r * r)
        niter = 0
        if not rsold < thr:
            niter = 10000
            for it in range(1, 10001):
                tmp1 = seqsum(r * r)
                m = self.apply(lo, di, up, p)
                tmp2 = seqsum(m * p)
                alpha = tmp1 / tmp2
                x = x + alpha * p
                r = r - alpha * m
                rsnew = seqsum(r * r)
                if rsnew < thr:
                    niter = it
                    break
                p = r + (rsnew / rsold) * p
                rsold = rsnew
        return x + self.arr["pfhyd"], niter

    def cdiff(self, ks, sigma, pftry):
        o, phi = self.o, self.arr["phi"]
        cc = o["eta"] * o["beta"] * phi
        if o["pressure_dependent"]:
            return ks * np.exp(-(sigma - pftry) / o["s0"]) / cc * 1e-6
        return ks / cc * 1e-6

    def diffuse(self, pf, y, sigma, h, t, dtnxt):
        """implicitsolver2 (main_fv.f90:1727-1774) or implicitsolver
        (1678-1724): (pfnew, CG iterations of the last solve, dtnxt)."""
        o = self.o
        ks = y[3::4]
        if o["pressure_dependent"]:
            pftry = pf.copy()
            for _ in range(20):
                pfnew, niter = self.beuler(pf, self.cdiff(ks, sigma, pftry), h, t)
                err = float(np.max(np.abs(pftry - pfnew)))
                if err < _f32(1e-3):
                    break
                pftry = pfnew
            dpth = _f32(0.3)
        else:
            pfnew, niter = self.beuler(pf, self.cdiff(ks, sigma, pf), h, t)
            dpth = _f32(0.1)
        dmax = float(np.max(pfnew - pf))
        if dtnxt / h * dmax > dpth:
            dtnxt = dpth * h / dmax
        if o["pressure_dependent"] and o["outpertime"]:
            dtnxt = min(dtnxt, o["dtout_s"])
        return pfnew, niter, dtnxt

    # ----------------------------------------------------------------- run --
    def step(self, st: dict) -> dict:
        """One accepted step of the main loop (main_fv.f90:1009-1152) on the
        state dict (y, pf, disp, t, h); returns the monitor row values."""
        o, A = self.o, self.arr
        y, t = st["y"], st["t"]
        yout, t, hdid, hnext, errmax, nrej = self.rkqs(y, t, st["h"])
        psi, tau, sige, ks = yout[0::4], yout[1::4], yout[2::4].copy(), yout[3::4]
        pf = st["pf"]
        velpre = self.vel(psi, tau, sige)
        sigma = sige + pf
        niter = st.get("niter", 0)
        if not o["pfconst"]:
            pf, niter, hnext = self.diffuse(pf, yout, sigma, hdid, t, hnext)
        sige = sigma - pf
        yout[2::4] = sige
        disp = st["disp"] + velpre * hdid * 0.5
        v = self.vel(psi, tau, sige)
        disp = disp + v * hdid * 0.5
        mu = tau / sige
        kp = ks * np.exp(-sige / o["s0"]) if o["pressure_dependent"] else ks
        ph = A["pfhyd"]
        n = self.n
        q = np.empty(n)
        q[0] = -kp[0] / o["eta"] * 1e3 * (pf[1] - ph[1] - pf[0] + ph[0]) / o["ds0"]
        q[1:n - 1] = -kp[1:n - 1] / o["eta"] * 1e3 * (pf[2:n] - ph[2:n] - pf[0:n - 2] + ph[0:n - 2]) / 2 / o["ds0"]
        q[n - 1] = -kp[n - 1] / o["eta"] * 1e3 * (pf[n - 1] - ph[n - 1] - pf[n - 2] + ph[n - 2]) / o["ds0"]
        st.update(y=yout, pf=pf, disp=disp, vel=v, qflow=q, t=t, h=hnext, niter=niter)
        av = np.abs(v)
        mvel = float(av.max())
        row = dict(t=t, log10vmax=np.log10(mvel), meandisp=seqsum(disp * A["area"]) / self.sg,
                   meanmu=seqsum(mu * A["area"]) / self.sg, maxnorm=float(sige.max()), minnorm=float(sige.min()),
                   argmax=int(np.argmax(av)) + 1, errmax=errmax, dt=hdid, niter=niter, rejected=nrej)
        status = 0
        if mvel > o["velmax"] or mvel < o["velmin"]:
            status = 1
        elif t > o["tmax_s"]:
            status = 3
        elif sige.min() < _f32(0.2):
            status = 4
        row["status"] = status
        return row
