"""CPU: the main_fv3dp oracle (oracle/fluid3dp.py): its 2-D pore-pressure
step pinned against the 1-D step of main_fv's oracle (pinned in
tests/test_fluid_oracle.py) and against what the mathematics of the
finite-volume system fixes (the reference ships no test or fixture for this
driver):

* a grid whose data vary along one direction only, Neumann across it: the
  2-D solution is uniform across and equals main_fv's 1-D solution along it
  (both directions, Neumann and Dirichlet ends) -- pins both line operators,
  the 0.5 + 0.5 split of the identity and the assembly;
* all-Neumann, no sources: the boundary-halved pressure sum is conserved (a
  planted sign error breaks it); a uniform pressure is a fixed point; a
  centre source adds exactly its volume h q / (beta phi0) 1e-12 / ds0^2;
* a well at the centre of an odd grid gives a mirror-symmetric pressure;
* conjugate gradients agree with a direct solve of the assembled system to
  the residual bound;
* the wells' rate schedule by hand (examples/wells.dat's first well);
* implicitsolver's Picard result is a fixed point to its 1e-5 test.
"""
import types

import numpy as np
import pytest

from hbi_amd import problems as P
from oracle import fluid as F
from oracle import fluid3dp as F3

IM, JM = 7, 16
DS0 = 0.02


def _opts(**kw):
    o = dict(ds0=DS0, bcl=F.NEUMANN, bcr=F.NEUMANN, bct=F.NEUMANN, bcb=F.NEUMANN, setting=F3.INJECTION,
             q0=0.0, beta=5e-8, phi0=0.2, eta=0.89e-3, pbcl=0.0, pbcr=0.0)
    o.update(kw)
    return o


def _dense(G, Amx, Amy):
    n = G.imax * G.jmax
    A = np.empty((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        A[:, k] = G.apply(Amx, Amy, e.reshape(G.imax, G.jmax)).ravel()
    return A


def _weights(im, jm):
    wi, wj = np.ones(im), np.ones(jm)
    wi[[0, -1]] = 0.5
    wj[[0, -1]] = 0.5
    return np.outer(wi, wj)


def _fv1d(n, bcl, bcr):
    """main_fv's 1-D oracle on n elements, no sources (setting 'other')."""
    one = np.ones(n)
    arr = dict(a=one, b=one, dc=one, f0=one, vpl=one, kpmaxv=one, kTv=one, phi=0.2 * one, pfhyd=np.zeros(n),
               qin=np.zeros(n), area=one)
    opts = dict(bcl=bcl, bcr=bcr, ds0=DS0, setting=F.OTHER, q0=0.0, beta=5e-8, pbcl=0.0, pbcr=0.0, pinj=0.0,
                tinj_s=0.0)
    return F.FvOracle(types.SimpleNamespace(n=n), opts, arr)


def _banded(lo, di, up, b):
    from scipy.linalg import solve_banded
    n = b.size
    ab = np.zeros((3, n))
    ab[0, 1:] = up[:-1]
    ab[1] = di
    ab[2, :-1] = lo[1:]
    return solve_banded((1, 1), ab, b)


@pytest.mark.parametrize("axis,bc", [(1, F.NEUMANN), (1, F.DIRICHLET), (0, F.NEUMANN), (0, F.DIRICHLET)])
def test_grid_lines_reduce_to_the_1d_step(axis, bc):
    rng = np.random.default_rng(3 + axis + 2 * bc)
    h = 3e3
    if axis == 1:      # variation along j (Dxx: top / bottom), Neumann along i
        o = _opts(bct=bc, bcb=bc)
        c1 = 1e-7 * (1 + rng.random(JM))
        p1 = 5 + rng.standard_normal(JM)
        c, pf = np.tile(c1, (IM, 1)), np.tile(p1, (IM, 1))
    else:              # variation along i (Dyy: left / right), Neumann along j
        o = _opts(bcl=bc, bcr=bc)
        c1 = 1e-7 * (1 + rng.random(IM))
        p1 = 5 + rng.standard_normal(IM)
        c, pf = np.tile(c1[:, None], (1, JM)), np.tile(p1[:, None], (1, JM))
    G = F3.Grid2D(IM, JM, o)
    Amx, Amy, b, x0 = G.system(pf, c, h, 0.0, np.zeros((IM, JM)))
    x = np.linalg.solve(_dense(G, Amx, Amy), b.ravel()).reshape(IM, JM)
    line = x[0] if axis == 1 else x[:, 0]
    across = x - (x[0:1] if axis == 1 else x[:, 0:1])
    assert np.abs(across).max() <= 1e-12 * np.abs(x).max()
    fo = _fv1d(c1.size, bc, bc)
    lo, di, up, b1, _ = fo.beuler_system(p1, c1, h, 0.0)
    x1 = _banded(lo, di, up, b1)
    assert np.abs(line - x1).max() <= 1e-12 * np.abs(x1).max()


@pytest.mark.parametrize("plant", [False, True])
def test_grid_neumann_conserves_the_pressure_sum(plant):
    rng = np.random.default_rng(5)
    G = F3.Grid2D(IM, JM, _opts())
    c = 1e-7 * (1 + rng.random((IM, JM)))
    pf = 5 + rng.standard_normal((IM, JM))
    Amx, Amy, b, x0 = G.system(pf, c, 3e3, 0.0, np.zeros((IM, JM)))
    A = _dense(G, Amx, Amy)
    if plant:   # a sign error in one coupling along i
        k = 3 * JM + 5
        A[k - JM, k] = -A[k - JM, k]
    x = np.linalg.solve(A, b.ravel()).reshape(IM, JM)
    w = _weights(IM, JM)
    ok = abs((w * x).sum() - (w * pf).sum()) <= 1e-12 * (w * np.abs(pf)).sum()
    assert ok != plant


def test_grid_uniform_pressure_is_a_fixed_point_and_a_source_adds_its_volume():
    G = F3.Grid2D(IM, JM, _opts())
    c = np.full((IM, JM), 1e-7)
    pf = np.full((IM, JM), 2.5)
    pnew, it = G.beuler(pf, c, 1e4, 0.0, np.zeros((IM, JM)))
    assert it == 0 and np.array_equal(pnew, pf)
    q0, h = 3e-4, 1e4
    G = F3.Grid2D(IM, JM, _opts(q0=q0))
    pnew, it = G.beuler(pf, c, h, 0.0, np.zeros((IM, JM)))
    vol = h * q0 / 5e-8 / 0.2 * float(np.float32(1e-12)) / DS0 / DS0
    w = _weights(IM, JM)
    assert it > 0 and pnew[IM // 2 - 1, JM // 2 - 1] == pnew.max()
    tol = float(np.float32(1e-6))
    assert abs((w * (pnew - pf)).sum() - vol) <= 2 * tol * np.sqrt(IM * JM)


def test_centred_well_is_mirror_symmetric_and_cg_matches_a_direct_solve():
    im, jm = 9, 11
    wells = [((im + 1) // 2, (jm + 1) // 2, [2e-4, 2e-4], [0.0, 1e9])]
    G = F3.Grid2D(im, jm, _opts(), wells)
    c = np.full((im, jm), 2e-7)
    pf = np.zeros((im, jm))
    h = 5e3
    pnew, it = G.beuler(pf, c, h, 1e3, np.zeros((im, jm)))
    assert it > 0 and pnew.max() > 0
    assert np.abs(pnew - pnew[::-1]).max() <= 1e-9 * pnew.max()
    assert np.abs(pnew - pnew[:, ::-1]).max() <= 1e-9 * pnew.max()
    Amx, Amy, b, _ = G.system(pf, c, h, 1e3, np.zeros((im, jm)))
    x = np.linalg.solve(_dense(G, Amx, Amy), b.ravel()).reshape(im, jm)
    tol = float(np.float32(1e-6))
    assert np.linalg.norm(pnew - x) <= 2 * tol * np.sqrt(im * jm)


def test_well_rate_schedule():
    # examples/wells.dat, first well: rates 3e-4 3e-4 6e-4 6e-4 3e-4 3e-4 0 0 at 0 1e7 1e7 2e7 2e7 3e7 3e7 99e7
    qv = [3e-4, 3e-4, 6e-4, 6e-4, 3e-4, 3e-4, 0.0, 0.0]
    qt = [0.0, 1e7, 1e7, 2e7, 2e7, 3e7, 3e7, 99e7]
    assert F3.well_rate(qv, qt, 8, 5e6) == 3e-4
    assert F3.well_rate(qv, qt, 8, 1.5e7) == 6e-4
    assert F3.well_rate(qv, qt, 8, 2.5e7) == 3e-4
    assert F3.well_rate(qv, qt, 8, 5e7) == 0.0
    assert F3.well_rate(qv, qt, 8, 1e9) == 0.0
    # a ramp: halfway between 1 and 3 at t = 15 of [10, 20]
    assert F3.well_rate([1.0, 3.0], [10.0, 20.0], 2, 15.0) == 2.0
    assert F3.well_rate([1.0, 3.0], [10.0, 20.0], 2, 25.0) == 3.0
    assert F3.well_rate([1.0, 3.0], [10.0, 20.0], 2, 5.0) == 0.0


def test_picard_result_is_a_fixed_point():
    im, jm = 8, 10
    n = im * jm
    lod = np.random.default_rng(1).permutation(n)
    one = np.ones(n)
    arr = dict(a=0.013 * one, b=0.0 * one, dc=1e-4 * one, f0=0.6 * one, vpl=0.0 * one, kpmaxv=3e-14 * one,
               kTv=1e15 * one, phi=0.2 * one, pfhyd=np.zeros(n), qin=np.zeros(n), area=one)
    o = _opts(q0=3e-4, imax=im, jmax=jm, pressure_dependent=1, s0=5.0, dtmax=1e6, minsig=0.0, outpertime=0,
              pfconst=0, kL=1e-3, kpmin=1e-14, vref=1e-6, law=0, backslip=1, hdiv=0.0, Bs=0.0, eps_r=1e-6)
    fo = F3.Fv3Oracle(types.SimpleNamespace(n=n), o, arr, lod)
    y = np.zeros(4 * n)
    y[3::4] = 2e-14
    sigma = np.full(n, 30.0)
    pf = np.zeros(n)
    pnew, it, dt = fo.diffuse(pf, y, sigma, 1e5, 1e5, 1e6)
    assert it > 0 and pnew.max() > 0 and dt <= 1e6
    # one more Picard pass from the result changes it by less than the test
    G = fo.grid
    cc = o["eta"] * o["beta"] * o["phi0"]
    cd = fo.to_grid(y[3::4]) * np.exp(-(fo.to_grid(sigma) - fo.to_grid(pnew)) / o["s0"]) / cc * 1e-6
    again, _ = G.beuler(fo.to_grid(pf), cd, 1e5, 1e5, np.zeros((im, jm)))
    assert np.abs(again - fo.to_grid(pnew)).max() < float(np.float32(1e-5)) * 1.01


def test_ss_setting_keeps_the_linear_background_flow():
    """'ss' (initcond_bgflow, main_fv3dp.f90:1063-1099): the linear profile
    carrying q0 between the flux penalties of rows i = 1 and imax is steady;
    what a step changes is only the single-precision literal 1e-9 of the
    penalties (relative 2.8e-8 of the penalty)."""
    im, jm = 12, 9
    q0, eta, beta, phi0, ks = 3e-4, 0.89e-3, 5e-8, 0.2, 1e-14
    G = F3.Grid2D(im, jm, _opts(setting=F3.SS, q0=q0, eta=eta, beta=beta, phi0=phi0))
    lf = im * DS0
    pbcl = q0 * lf / ks * eta * 1e-3
    xcol = (np.arange(1, im + 1) - im // 2 - 0.5) * DS0
    pf = np.tile((pbcl + (0.0 - pbcl) * xcol / lf)[:, None], (1, jm))
    c = np.full((im, jm), ks / (eta * beta * phi0) * 1e-6)
    for h in (1e2, 1e4, 1e6):
        Amx, Amy, b, _ = G.system(pf, c, h, 0.0, np.zeros((im, jm)))
        x = np.linalg.solve(_dense(G, Amx, Amy), b.ravel()).reshape(im, jm)
        sat = 2 * h * q0 * 1e-9 / (beta * phi0 * DS0)
        assert np.abs(x - pf).max() <= 3e-8 * sat


@pytest.mark.parametrize("plant", [False, True])
def test_thrust_penalties_add_their_volumes(plant):
    """'thrust' (main_fv3dp.f90:1531-1545) by hand: with Neumann boundaries
    the operator keeps the half-weighted sum of pf - pfhyd, so a step adds
    exactly the penalties' volumes: at the top, 1.0 x (j = 1) and 0.5 x
    (j = 2) of 2 h c(i, 1) / ds0^2 (pbcl - pfhyd(i, 1)), i.e. 2 h c (pbcl -
    pfhyd) / ds0^2 per unit weight of column i; at the bottom the flux
    2 h q0 / (beta phi0) 1e-9 / ds0 on row jmax (1e-9 the single-precision
    literal), i.e. h q0 / (beta phi0) 1e-9 / ds0 per unit weight.  A planted
    mistake (the top penalty from c(i, 2)) breaks it."""
    rng = np.random.default_rng(11)
    q0, beta, phi0, pbcl, h = 2e-4, 5e-8, 0.2, 1.5, 2e3
    G = F3.Grid2D(IM, JM, _opts(setting=F3.THRUST, q0=q0, beta=beta, phi0=phi0, pbcl=pbcl))
    c = 1e-7 * (1 + rng.random((IM, JM)))
    hyd = 0.3 * rng.random((IM, JM))
    pf = hyd + 1 + rng.standard_normal((IM, JM))
    cc = c.copy()
    if plant:
        cc[:, 0] = c[:, 1]
    Amx, Amy, b, x0 = G.system(pf, cc, h, 0.0, hyd)
    x = np.linalg.solve(_dense(G, Amx, Amy), b.ravel()).reshape(IM, JM)
    w = _weights(IM, JM)
    wi = w[:, 1]
    top = (wi * 2 * h * c[:, 0] / DS0 ** 2 * (pbcl - hyd[:, 0])).sum()
    bottom = wi.sum() * h * q0 / (beta * phi0) * float(np.float32(1e-9)) / DS0
    got = (w * (x - (pf - hyd))).sum()
    ok = abs(got - (top + bottom)) <= 1e-10 * (abs(top) + abs(bottom))
    assert ok != plant


def test_minsig_rule_and_post_step_order():
    """main_fv3dp.f90:761-787 by hand on a scripted step: the pressure step
    acts on the held sigma; where the RK's sigma_e is below minsig, sigma =
    pf(new) + minsig (so sigma_e = minsig there), elsewhere sigma_e = sigma -
    pf(new); the slip grows by the trapezoid of V from the RK state and V
    from the new sigma_e; the monitor takes the plain means of N cells."""
    im, jm = 6, 5
    g = P.mesh_3dp(im, jm, DS0)
    hs = types.SimpleNamespace(n=g.n)
    _, opts, arr, y, pf = P.fluid3dp_injection(im, jm, DS0, minsig=20.0)
    y = y.copy()
    low = np.zeros(g.n, dtype=bool)
    low[[3, 11, 17]] = True
    y[2::4][low] = 15.0                      # the RK's sigma_e below minsig there
    dpf = np.linspace(0.1, 0.9, g.n)

    class Fixed(F3.Fv3Oracle):
        def rkqs(self, yy, t, htry):
            return yy.copy(), t + htry, htry, 2 * htry, 0.5, 0

        def diffuse(self, pf_, yy, sigma, h, t, dtnxt):
            return pf_ + dpf, 4, dtnxt

    S = Fixed(hs, opts, arr, np.arange(g.n))
    sigma = y[2::4] + pf + 5.0               # the held total normal stress
    h = 1e3
    st = dict(y=y.copy(), pf=pf.copy(), sigma=sigma.copy(), disp=np.zeros(g.n), t=0.0, h=h)
    row = S.step(st)
    p1 = pf + dpf
    sig1 = sigma.copy()
    sig1[low] = p1[low] + 20.0
    sige1 = sig1 - p1
    assert np.array_equal(st["sigma"], sig1) and np.allclose(st["y"][2::4], sige1, rtol=1e-15)
    assert np.all(st["y"][2::4][low] == sig1[low] - p1[low])
    psi, tau = y[0::4], y[1::4]
    v0 = S.vel(psi, tau, y[2::4])
    v1 = S.vel(psi, tau, sige1)
    assert np.allclose(st["disp"], v0 * h * 0.5 + v1 * h * 0.5, rtol=1e-14)
    assert np.isclose(row["minnorm"], sige1.min()) and np.isclose(row["meandisp"], st["disp"].mean(), rtol=1e-14)
    assert np.isclose(row["meanmu"], (tau / sige1).mean(), rtol=1e-14)
