"""CPU: the fluid-driven cycle oracle (oracle/fluid.py, main_fv.f90) pinned
against the already-pinned main_LH oracle and against what the mathematics
of its finite-volume pore-pressure step fixes (the reference ships no test
or fixture for this driver):

* derivs / rkck: the psi, tau, sigma_e components equal the C oracle's
  main_LH derivs / rkck on the same state; dks/dt = 0 at the steady
  permeability of main_fv.f90:1420
* Beuler: Neumann boundaries without sources conserve the half-weighted
  pressure sum; the Dirichlet steady state is the linear profile through the
  boundary pressures; the injection rate adds exactly its volume while t <
  tinj; the injection-pressure penalty pulls the centre to pinj; conjugate
  gradients agree with a direct banded solve to the residual bound; a
  planted sign error in the operator breaks the conservation test
* implicitsolver: the Picard result is a fixed point to the 1e-3 test
* a 30-step run holds the injected volume and raises pf at the source
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded

from hbi_amd import problems as P
from oracle import fluid as F
from oracle import oracle_lib as O

N = 128


@pytest.fixture(scope="module")
def hs():
    g = P.mesh_2dp(N, 0.01)
    return g, O.HMatrix(g, leaf=16, eps=1e-8, nthreads=4)


def _oracle(hs, **kw):
    g, H = hs
    opts, arr, y, pf = P.fluid_injection(g, H.lod, **kw)
    return F.FvOracle(H, opts, arr), y, pf


def _direct(lo, di, up, b):
    n = b.size
    ab = np.zeros((3, n))
    ab[0, 1:] = up[:-1]
    ab[1] = di
    ab[2, :-1] = lo[1:]
    return solve_banded((1, 1), ab, b)


def _weights(n):
    w = np.ones(n)
    w[0] = w[-1] = 0.5
    return w


def test_fluid_derivs_and_rkck_match_the_main_lh_oracle(hs):
    g, H = hs
    fo, y, pf = _oracle(hs)
    rng = np.random.default_rng(1)
    y[1::4] *= 1 + 0.01 * rng.standard_normal(N)     # V off steady state, still > 0
    fr = P.Friction(a=fo.arr["a"], b=fo.arr["b"], dc=fo.arr["dc"], f0=fo.arr["f0"], vpl=fo.arr["vpl"],
                    tau=y[1::4], sigma=y[2::4], vel=np.ones(N), taudot=np.zeros(N), sigdot=np.zeros(N), law=0,
                    vref=fo.o["vref"])
    sim = O.Sim(H, fr, np.arange(N), limitsigma=False)
    y3 = np.empty(3 * N)
    for c in range(3):
        y3[c::3] = y[c::4]
    d4, d3 = fo.derivs(y), sim.derivs(y3)
    for c in range(3):
        assert np.allclose(d4[c::4], d3[c::3], rtol=1e-12, atol=1e-300), c
    h = 100.0
    (yo4, ye4), (yo3, ye3) = fo.rkck(y, h), sim.rkck(y3, h)
    for c in range(3):
        assert np.allclose(yo4[c::4], yo3[c::3], rtol=1e-13, atol=0), c
        # yerr is a difference of nearly equal sums: compare on the scale of y
        assert np.abs(ye4[c::4] - ye3[c::3]).max() <= 1e-14 * np.abs(y[c::4]).max(), c


def test_fluid_permeability_rate_vanishes_at_the_steady_permeability(hs):
    fo, y, pf = _oracle(hs)
    o, A = fo.o, fo.arr
    v = fo.vel(y[0::4], y[1::4], y[2::4])
    kss = (A["kpmaxv"] * v / o["kL"] + o["kpmin"] / A["kTv"]) / (v / o["kL"] + 1 / A["kTv"])
    y[3::4] = kss
    d = fo.derivs(y)[3::4]
    # both terms cancel up to the rounding of kss itself
    ulp = 8 * np.finfo(float).eps * kss * (1 / A["kTv"] + v / o["kL"])
    assert np.all(np.abs(d) <= ulp)
    assert np.all(np.abs(v / o["kL"] * (kss - A["kpmaxv"])) > 1e3 * ulp)
    y[3::4] = kss * 1.5
    assert np.all(fo.derivs(y)[3::4] < 0)


def _rand_system(fo, rng, h=1e4, t=0.0):
    c = 1e-7 * (1 + rng.random(N))
    pf = 5 + rng.standard_normal(N)
    return c, pf, fo.beuler_system(pf, c, h, t)


@pytest.mark.parametrize("plant", [False, True])
def test_beuler_neumann_conserves_the_pressure_sum(hs, plant):
    fo, _, _ = _oracle(hs, setting="other")
    rng = np.random.default_rng(2)
    c, pf, (lo, di, up, b, x0) = _rand_system(fo, rng)
    if plant:   # a sign error in one coupling of the operator
        up = up.copy()
        up[N // 3] = -up[N // 3]
    x = _direct(lo, di, up, b)
    w = _weights(N)
    ok = abs(np.dot(w, x) - np.dot(w, x0)) <= 1e-12 * np.dot(w, np.abs(x0))
    assert ok != plant
    if not plant:
        pnew, it = fo.beuler(pf, c, 1e4, 0.0)
        tol = float(np.float32(1e-4))
        assert it > 0 and abs(np.dot(w, pnew) - np.dot(w, pf)) <= tol * N
        assert np.linalg.norm(pnew - x) <= 2 * tol * np.sqrt(N)


def test_beuler_dirichlet_steady_state_is_linear(hs):
    fo, _, _ = _oracle(hs, setting="other", bcl="Dirichlet", bcr="Dirichlet", pbcl=3.0, pbcr=1.0)
    c = np.full(N, 1e-7)
    pf = np.zeros(N)
    h = 1e13 * 0.01 ** 2 / 1e-7   # h c / ds0^2 = 1e13: every mode has relaxed
    lo, di, up, b, x0 = fo.beuler_system(pf, c, h, 0.0)
    x = _direct(lo, di, up, b)
    lin = 3.0 + (1.0 - 3.0) * np.arange(N) / (N - 1)
    assert np.abs(x - lin).max() <= 1e-7


def test_beuler_injection_rate_adds_its_volume(hs):
    fo, _, _ = _oracle(hs, setting="injectionq", tinj=1.0)
    rng = np.random.default_rng(3)
    c = 1e-7 * (1 + rng.random(N))
    pf = rng.random(N)
    w = _weights(N)
    o, A = fo.o, fo.arr
    h = 5e3
    k = N // 2 - 1
    vol = h * A["qin"][k] / o["beta"] / A["phi"][k] * float(np.float32(1e-9)) / o["ds0"]
    for t, expect in ((1e3, vol), (2 * o["tinj_s"], 0.0)):
        lo, di, up, b, x0 = fo.beuler_system(pf, c, h, t)
        x = _direct(lo, di, up, b)
        assert abs((np.dot(w, x) - np.dot(w, x0)) - expect) <= 1e-12 * np.dot(w, np.abs(x0)) + 1e-15, t
    assert vol > 0


def test_beuler_injection_pressure_penalty_holds_the_centre(hs):
    fo, _, _ = _oracle(hs, setting="injectionp", pinj=7.0)
    c = np.full(N, 1e-30)
    pf = np.zeros(N)
    x = _direct(*fo.beuler_system(pf, c, 1e9, 0.0)[:4])
    k = N // 2 - 1
    assert abs(x[k] - 7.0) <= 1e-6 and np.abs(np.delete(x, k)).max() <= 1e-12


def test_picard_result_is_a_fixed_point(hs):
    fo, y, pf = _oracle(hs, setting="injectionq", pressure_dependent=True)
    rng = np.random.default_rng(4)
    pf = 2 * rng.random(N)
    sigma = y[2::4] + pf
    y[3::4] = 1e-12
    pnew, it, dt = fo.diffuse(pf, y, sigma, 3e4, 0.0, 1e9)
    again, _ = fo.beuler(pf, fo.cdiff(y[3::4], sigma, pnew), 3e4, 0.0)
    assert np.abs(again - pnew).max() < 1e-3
    assert dt < 1e9   # the dpf bound of main_fv.f90:1717


def test_fluid_oracle_run_holds_the_injected_volume(hs):
    fo, y, pf = _oracle(hs, setting="injectionq")
    st = dict(y=y.copy(), pf=pf.copy(), disp=np.zeros(N), t=0.0, h=1.0)
    w = _weights(N)
    o, A = fo.o, fo.arr
    k = N // 2 - 1
    rate = A["qin"][k] / o["beta"] / A["phi"][k] * float(np.float32(1e-9)) / o["ds0"]
    vol, pk = 0.0, [0.0]
    tol = float(np.float32(1e-4))
    for _ in range(30):
        row = fo.step(st)
        assert row["status"] == 0
        vol += row["dt"] * rate
        pk.append(st["pf"][k])
    # a step whose change leaves the residual below tol^2 N is not applied
    # (main_fv.f90:1908-1910): the first, short steps leave pf as it is
    assert np.all(np.diff(pk) >= 0) and pk[-1] > pk[10] > 0
    assert abs(np.dot(w, st["pf"]) - vol) <= 30 * tol * N


class _Scripted(F.FvOracle):
    """The oracle's rkqs driven by scripted error ratios: the k-th rkck call
    returns yerr / yout = errs[k] * eps_r on every component."""

    def __init__(self, hs, errs, **opts):
        g, H = hs
        o, arr, _, _ = P.fluid_injection(g, H.lod)
        o.update(opts)
        super().__init__(H, o, arr)
        self.errs = list(errs)
        self.calls = []

    def rkck(self, y, h):
        self.calls.append(h)
        e = self.errs[len(self.calls) - 1]
        return np.ones_like(y), np.full_like(y, e * self.o["eps_r"])


def test_fluid_rkqs_controller_known_answers(hs):
    """main_fv.f90:2208-2248 by hand: SAFETY = 0.9 and PGROW = -0.2 are
    default-real literals, PSHRNK = -0.25; a rejection (errmax >= 1) shrinks
    h to max(h / 2, 0.9 h errmax^-0.25); errmax <= 1e-15 is a rejection too
    (h / 2), unlike main_LH; an accepted step proposes min(2 h, 0.9 h
    errmax^-0.2), then min(., dtout) with outpertime."""
    y0 = np.ones(4 * N)
    # errmax 2 then 0.5: one rejection to 0.9f * 2^-0.25 = 0.75680675 h, then growth 0.9f * 0.5^-0.2f
    S = _Scripted(hs, [2.0, 0.5])
    _, t1, hdid, hnext, errmax, nrej = S.rkqs(y0, 0.0, 100.0)
    assert nrej == 1 and S.calls == [100.0, hdid]
    assert abs(hdid - 75.680675367981) <= 1e-9 and t1 == hdid
    assert abs(hnext / hdid - 1.033828494245849) <= 1e-14 and errmax == 0.5
    # errmax 1e-16: h / 2 (not a growth step), then 0.001: growth capped at 2 h
    S = _Scripted(hs, [1e-16, 0.001])
    _, _, hdid, hnext, _, nrej = S.rkqs(y0, 0.0, 64.0)
    assert nrej == 1 and hdid == 32.0 and hnext == 64.0
    # errmax 0.5 with outpertime: the proposal is capped at dtout
    S = _Scripted(hs, [0.5], outpertime=1, dtout_s=10.0)
    _, _, hdid, hnext, _, nrej = S.rkqs(y0, 0.0, 100.0)
    assert nrej == 0 and hdid == 100.0 and hnext == 10.0
    # a rejection from errmax 1.5: max(h / 2, 0.9f * 1.5^-0.25) = 0.81324178 h
    S = _Scripted(hs, [1.5, 0.9])
    _, _, hdid, _, _, nrej = S.rkqs(y0, 0.0, 1.0)
    assert nrej == 1 and abs(hdid - 0.8132417817053098) <= 1e-15
    # step-size underflow at large t
    S = _Scripted(hs, [2.0] * 8)
    with pytest.raises(FloatingPointError):
        S.rkqs(y0, 1e12, 1e-5)


def test_fluid_post_step_update_order(hs):
    """main_fv.f90:1045-1082 by hand on a scripted step: the pore-pressure
    update acts on sigma = sigma_e(RK) + pf(old) held fixed, so sigma_e(new)
    = sigma - pf(new); the slip grows by the trapezoid of V before (RK state)
    and after (new sigma_e) the pressure update; qflow is Darcy's flux of
    the new pf in excess of pfhyd (one-sided at the ends, centred inside).
    A planted order mistake (V after the update from the RK sigma_e) moves
    the slip."""
    g, H = hs
    o, arr, y, pf = P.fluid_injection(g, H.lod, tauinit=29.0, muinit=0.0, velinit=1e-9)
    O_ = F.FvOracle(H, o, arr)
    dpf = np.linspace(0.5, 1.5, N)

    class Fixed(F.FvOracle):
        def rkqs(self, yy, t, htry):
            return yy.copy(), t + htry, htry, 2 * htry, 0.5, 0

        def diffuse(self, pf_, yy, sigma, h, t, dtnxt):
            return pf_ + dpf, 3, dtnxt

    S = Fixed(H, o, arr)
    h = 1e4
    st = dict(y=y.copy(), pf=pf.copy(), disp=np.zeros(N), t=0.0, h=h)
    S.step(st)
    psi, tau, sige0, ks = y[0::4], y[1::4], y[2::4], y[3::4]
    sige1 = sige0 + pf - (pf + dpf)
    v0 = O_.vel(psi, tau, sige0)
    v1 = O_.vel(psi, tau, sige1)
    assert np.allclose(st["y"][2::4], sige1, rtol=1e-15)
    assert np.allclose(st["disp"], (v0 * h * 0.5) + v1 * h * 0.5, rtol=1e-14)
    assert not np.allclose(st["disp"], v0 * h, rtol=1e-6)          # the planted order mistake differs
    assert np.allclose(st["vel"], v1, rtol=1e-15)
    p1 = pf + dpf
    dx = o["ds0"]
    k = ks / o["eta"] * 1e3
    assert np.isclose(st["qflow"][0], -k[0] * (p1[1] - p1[0]) / dx, rtol=1e-13)
    assert np.isclose(st["qflow"][5], -k[5] * (p1[6] - p1[4]) / 2 / dx, rtol=1e-13)
    assert np.isclose(st["qflow"][-1], -k[-1] * (p1[-1] - p1[-2]) / dx, rtol=1e-13)
    assert st["h"] == 2 * h and st["t"] == h
