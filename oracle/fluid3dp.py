"""CPU restatement of main_fv3dp.f90: the fluid-driven cycle on a planar
fault of imax x jmax square elements (problem 3dp) with the pore pressure
diffused on the fault's own 2-D grid (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows, in its arithmetic (left-to-right operation order, single-precision
literals where the reference writes them):
  Beuler          main_fv3dp.f90:1426-1649  the 2-D backward-Euler system: the
                                            finite-volume operators along j
                                            (Dxx, top / bottom boundaries) and
                                            along i (Dyy, left / right), the
                                            penalty vector of the setting, the
                                            injection sources (one at the
                                            centre, or wells on a rate
                                            schedule), conjugate gradients to
                                            |r|^2 < tol^2 imax jmax
  implicitsolver  main_fv3dp.f90:1314-1366  Picard iterations on the
                                            pressure-dependent diffusivity
  implicitsolver2 main_fv3dp.f90:1368-1424  constant diffusivity, and the step
                                            bound at the injection schedule's
                                            rate changes
  rkqs            main_fv3dp.f90:1743-1829  as main_fv's, hnext <= dtmax
  step            main_fv3dp.f90:725-800    post-RK update, pressure step, the
                                            minsig rule, fields, monitor
                                            values and stop criteria
The derivs and rkck are main_fv's (oracle.fluid.FvOracle: the 3dp products
are not divided by ds0, sigma is constant).  Grid cell (i, j) (1-based) is
mesh element (i - 1) jmax + j (coordinate3dp, main_fv3dp.f90:1174-1202);
the state arrays are in cluster order and map to the grid through lod.

Pinned (tests/test_fluid3dp_oracle.py): each grid line of the operator is
the 1-D operator of main_fv's Beuler (oracle.fluid, pinned there) on the
same diffusivities; the Neumann operator conserves the boundary-halved
pressure sum and a centre source adds exactly its volume; a uniform pressure
is a fixed point without sources; the 'ss' setting's linear profile is
steady; the conjugate-gradient result agrees with a direct solve of the
assembled system to the residual bound; mirror symmetry of a centred
source on a uniform grid; the rate schedule's interpolation by hand; the
'thrust' penalties add exactly their hand-computed volumes (top rows from
pbcl, bottom flux from q0); the minsig rule and the post-step order on a
scripted step (sigma held, sigma_e = minsig where the RK's falls below it,
the slip's trapezoid of V before and after the pressure step, plain means).
"""
from __future__ import annotations

import numpy as np

from .fluid import FvOracle, NEUMANN, DIRICHLET, _f32

INJECTION, SS, THRUST = 0, 2, 3     # settings of main_fv3dp: 'injection', 'ss', 'thrust'


def fsum(a) -> float:
    """Sequential sum in Fortran (column-major) order of a 2-D array."""
    v = np.asarray(a, dtype=np.float64).ravel(order="F")
    return float(np.cumsum(v)[-1]) if v.size else 0.0


def line_operator(c, bc_lo, bc_hi):
    """D(., 1:3) of one grid line (Beuler's Dxx along j or Dyy along i)
    before the 1 / ds0^2 scaling: row 0 keeps (diagonal, right), the last
    row (left, diagonal) in columns 2 and 3 (1-based), the others (left,
    diagonal, right).  c: diffusivities along the line."""
    n = c.size
    D = np.zeros((n, 3))
    if bc_lo == DIRICHLET:
        D[0, 0] = -c[0] - c[1]
        D[0, 1] = c[1] - c[0]
        D[1] = (c[1] / 2 - c[0] / 2, -c[0] / 2 - c[1] - c[2] / 2, c[1] / 2 + c[2] / 2)
    else:
        D[0, 0] = -c[0] - c[1]
        D[0, 1] = -D[0, 0]
        D[1] = (c[0] / 2 + c[1] / 2, -c[0] / 2 - c[1] - c[2] / 2, c[1] / 2 + c[2] / 2)
    k = np.arange(2, n - 2)
    D[k, 0] = c[k - 1] / 2 + c[k] / 2
    D[k, 1] = -c[k - 1] / 2 - c[k] - c[k + 1] / 2
    D[k, 2] = c[k] / 2 + c[k + 1] / 2
    if bc_hi == DIRICHLET:
        D[n - 2] = (c[n - 3] / 2 + c[n - 2] / 2, -c[n - 3] / 2 - c[n - 2] - c[n - 1] / 2,
                    -c[n - 1] / 2 + c[n - 2] / 2)
        D[n - 1, 2] = -c[n - 1] - c[n - 2]
        D[n - 1, 1] = c[n - 2] - c[n - 1]
    else:
        D[n - 2] = (c[n - 3] / 2 + c[n - 2] / 2, -c[n - 3] / 2 - c[n - 2] - c[n - 1] / 2,
                    c[n - 2] / 2 + c[n - 1] / 2)
        D[n - 1, 2] = -c[n - 1] - c[n - 2]
        D[n - 1, 1] = -D[n - 1, 2]
    return D


def line_matrix(D, h):
    """Am of one line: 0.5 - h D on the diagonal, -h D off it."""
    n = D.shape[0]
    A = np.zeros((n, 3))
    A[0, 0] = 0.5 - h * D[0, 0]
    A[0, 1] = -h * D[0, 1]
    A[1:n - 1, 0] = -h * D[1:n - 1, 0]
    A[1:n - 1, 1] = 0.5 - h * D[1:n - 1, 1]
    A[1:n - 1, 2] = -h * D[1:n - 1, 2]
    A[n - 1, 2] = 0.5 - h * D[n - 1, 2]
    A[n - 1, 1] = -h * D[n - 1, 1]
    return A


def line_apply(A, v):
    """The line product in the reference's order of terms (1566-1583)."""
    n = v.size
    m = np.empty(n)
    m[0] = v[0] * A[0, 0] + v[1] * A[0, 1]
    m[1:n - 1] = v[0:n - 2] * A[1:n - 1, 0] + v[1:n - 1] * A[1:n - 1, 1] + v[2:n] * A[1:n - 1, 2]
    m[n - 1] = v[n - 1] * A[n - 1, 2] + v[n - 2] * A[n - 1, 1]
    return m


def well_rate(qv, qt, kleng, time):
    """The rate of one well at time (main_fv3dp.f90:1553-1565): the
    interpolated segment holding time (the last such), or after the last
    time its last value; 0 before the schedule."""
    q = 0.0
    for k in range(kleng - 1):
        t0, t1, v0, v1 = qt[k], qt[k + 1], qv[k], qv[k + 1]
        if t0 <= time <= t1:
            with np.errstate(divide="ignore", invalid="ignore"):
                q = (v1 - v0) / (t1 - t0) * (time - t0) + v0
        elif time > t1 and k == kleng - 2:
            q = v1
    return q


class Grid2D:
    """Beuler of main_fv3dp on an imax x jmax grid: opts (ds0, bcl, bcr, bct,
    bcb, setting, q0, beta, phi0, eta, pbcl, pbcr), wells: None or a list of
    (i, j, qvals, qtimes) with 1-based grid indices."""

    def __init__(self, imax, jmax, opts: dict, wells=None):
        self.imax, self.jmax = imax, jmax
        self.o = dict(opts)
        self.wells = wells

    def system(self, pf, cdiff, h, time, pfhyd):
        """(Amx, Amy, b, x0) of Beuler (1437-1565) on grid arrays (imax, jmax)."""
        o = self.o
        im, jm = self.imax, self.jmax
        ds0 = o["ds0"]
        c = np.asarray(cdiff, dtype=np.float64)
        Dxx = np.stack([line_operator(c[i], o["bct"], o["bcb"]) for i in range(im)])          # (im, jm, 3)
        Dxx = Dxx / ds0 / ds0
        Dyy = np.stack([line_operator(c[:, j], o["bcl"], o["bcr"]) for j in range(jm)], 1)   # (im, jm, 3)
        Dyy = Dyy / ds0 / ds0
        Amx = np.stack([line_matrix(Dxx[i], h) for i in range(im)])
        Amy = np.stack([line_matrix(Dyy[:, j], h) for j in range(jm)], 1)
        SAT = np.zeros((im, jm))
        x0 = pf - pfhyd
        e9, e12 = _f32(1e-9), _f32(1e-12)
        if o["setting"] == SS:
            SAT[0, :] = -o["q0"] / o["beta"] / o["phi0"] * e9 * h / ds0 * 2
            SAT[im - 1, :] = o["q0"] / o["beta"] / o["phi0"] * e9 * h / ds0 * 2
        if o["setting"] == THRUST:
            SAT[:, 0] = -1.0 * c[:, 0] / ds0 / ds0 * (o["pbcl"] - pfhyd[:, 0]) * h * 2
            SAT[:, 1] = -0.5 * c[:, 0] / ds0 / ds0 * (o["pbcl"] - pfhyd[:, 0]) * h * 2
            if o["bcb"] == NEUMANN:
                SAT[:, jm - 1] = -o["q0"] / o["beta"] / o["phi0"] * e9 * h / ds0 * 2
            else:
                SAT[:, jm - 1] = -1.0 * c[:, jm - 1] / ds0 / ds0 * (o["pbcr"] - pfhyd[:, jm - 1]) * h * 2
                SAT[:, jm - 2] = -0.5 * c[:, jm - 1] / ds0 / ds0 * (o["pbcr"] - pfhyd[:, jm - 1]) * h * 2
        b = pf - SAT - pfhyd
        if o["setting"] == INJECTION:
            if self.wells:
                for (iw, jw, qv, qt) in self.wells:
                    q = well_rate(qv, qt, len(qv), time)
                    b[iw - 1, jw - 1] = b[iw - 1, jw - 1] + h * q / o["beta"] / o["phi0"] * e12 / ds0 / ds0
            else:
                i0, j0 = im // 2 - 1, jm // 2 - 1
                b[i0, j0] = b[i0, j0] + h * o["q0"] / o["beta"] / o["phi0"] * e12 / ds0 / ds0
        return Amx, Amy, b, x0

    def apply(self, Amx, Amy, v):
        """m = mx + my (1566-1585, 1597-1615)."""
        mx = np.stack([line_apply(Amx[i], v[i]) for i in range(self.imax)])
        my = np.stack([line_apply(Amy[:, j], v[:, j]) for j in range(self.jmax)], 1)
        return mx + my

    def beuler(self, pf, cdiff, h, time, pfhyd, itermax=1000):
        """(pfnew, CG iterations) of Beuler (1426-1649)."""
        Amx, Amy, b, x = self.system(pf, cdiff, h, time, pfhyd)
        tol = _f32(1e-6)
        thr = tol ** 2 * self.imax * self.jmax
        r = b - self.apply(Amx, Amy, x)
        p = r.copy()
        rsold = fsum(r * r)
        niter = 0
        if not rsold < thr:
            niter = itermax
            for it in range(1, itermax + 1):
                tmp1 = fsum(r * r)
                m = self.apply(Amx, Amy, p)
                tmp2 = fsum(m * p)
                alpha = tmp1 / tmp2
                x = x + alpha * p
                r = r - alpha * m
                rsnew = fsum(r * r)
                if rsnew < thr:
                    niter = it
                    break
                p = r + (rsnew / rsold) * p
                rsold = rsnew
        return x + pfhyd, niter


class Fv3Oracle(FvOracle):
    """main_fv3dp on a 3dp fault: hs the oracle H-matrix, opts as hbi_fv_opts
    (with imax, jmax, bct, bcb, dtmax, minsig, phi0), arrays in cluster
    order as hbi_fv_create, lod the cluster permutation, wells as Grid2D."""

    def __init__(self, hs, opts: dict, arrays: dict, lod, wells=None):
        super().__init__(hs, opts, arrays)
        self.lod = np.asarray(lod)
        self.im, self.jm = int(opts["imax"]), int(opts["jmax"])
        assert self.im * self.jm == self.n
        self.grid = Grid2D(self.im, self.jm, opts, wells)
        self.wells = wells
        self.switch = False
        self.nn = 2            # 1-based index into the first well's times (implicitsolver2)

    def to_grid(self, v):
        m = np.empty(self.n)
        m[self.lod] = v
        return m.reshape(self.im, self.jm)

    def from_grid(self, G):
        return np.asarray(G).reshape(self.n)[self.lod]

    def rkqs(self, y, t, htry):
        """main_fv3dp.f90:1743-1829: as main_fv's, hnext <= dtmax."""
        saved = self.o["outpertime"]
        self.o["outpertime"] = 0
        try:
            yout, t1, h, hnext, errmax, nrej = super().rkqs(y, t, htry)
        finally:
            self.o["outpertime"] = saved
        return yout, t1, h, min(hnext, self.o["dtmax"]), errmax, nrej

    def diffuse(self, pf, y, sigma, h, t, dtnxt):
        """implicitsolver (pressure dependent, 1314-1366) or implicitsolver2
        (1368-1424) at time t after the step: (pfnew, CG iterations, dtnxt)."""
        o = self.o
        ks = y[3::4]
        pfd, hyd = self.to_grid(pf), self.to_grid(self.arr["pfhyd"])
        ksd, sigd = self.to_grid(ks), self.to_grid(sigma)
        cc = o["eta"] * o["beta"] * o["phi0"]
        if o["pressure_dependent"]:
            pftry = pfd.copy()
            for _ in range(20):
                cdiff = ksd * np.exp(-(sigd - pftry) / o["s0"]) / cc * 1e-6
                pfnew, niter = self.grid.beuler(pfd, cdiff, h, t, hyd)
                err = float(np.max(np.abs(pftry - pfnew)))
                if err < _f32(1e-5):
                    break
                pftry = pfnew
        else:
            pfnew, niter = self.grid.beuler(pfd, ksd / cc * 1e-6, h, t, hyd)
        dpf = self.from_grid(pfnew) - pf
        amax = float(np.max(np.abs(dpf)))
        dpth = _f32(0.1)
        if dtnxt / h * amax > dpth:
            dtnxt = dpth * h / amax
        if not o["pressure_dependent"]:
            if self.switch:
                dtnxt = 20.0
                self.switch = False
            if self.wells:
                qt = self.wells[0][3]
                if self.nn <= len(qt) and t + dtnxt > qt[self.nn - 1]:
                    dtnxt = qt[self.nn - 1] - t - 1.0
                    self.switch = True
                    self.nn += 2
        return self.from_grid(pfnew), niter, dtnxt

    def step(self, st: dict) -> dict:
        """One accepted step of main_fv3dp's loop (725-800, 806-815)."""
        o = self.o
        y, t = st["y"], st["t"]
        yout, t, hdid, hnext, errmax, nrej = self.rkqs(y, t, st["h"])
        psi, tau, sige_rk, ks = yout[0::4], yout[1::4], yout[2::4].copy(), yout[3::4]
        velpre = self.vel(psi, tau, sige_rk)
        sigma = st["sigma"].copy()
        pf, niter = st["pf"], st.get("niter", 0)
        if not o["pfconst"]:
            pf, niter, hnext = self.diffuse(pf, yout, sigma, hdid, t, hnext)
        low = sige_rk < o["minsig"]
        sigma[low] = pf[low] + o["minsig"]
        sige = sigma - pf
        yout[2::4] = sige
        disp = st["disp"] + velpre * hdid * 0.5
        v = self.vel(psi, tau, sige)
        disp = disp + v * hdid * 0.5
        mu = tau / sige
        st.update(y=yout, pf=pf, sigma=sigma, disp=disp, vel=v, t=t, h=hnext, niter=niter)
        mvel = float(v.max())
        n = self.n
        row = dict(t=t, log10vmax=np.log10(mvel), meandisp=float(np.cumsum(disp)[-1]) / n,
                   meanmu=float(np.cumsum(mu)[-1]) / n, maxnorm=float(sige.max()), minnorm=float(sige.min()),
                   argmax=int(np.argmax(v)) + 1, errmax=errmax, dt=hdid, niter=niter, rejected=nrej)
        status = 0
        if mvel > o["velmax"] or mvel < o["velmin"]:
            status = 1
        elif t > o["tmax_s"]:
            status = 3
        elif sige.min() < _f32(0.2):
            status = 4
        row["status"] = status
        return row
