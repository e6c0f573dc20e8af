"""GPU: opening (main_LH.f90:705-716, 2096-2143, 1279-1283) on a small rough
triangular fault: the tensile-dislocation matrices onto shear and normal
traction (TD kernel, slip mode 'open'), derivs and an RK step against the
oracle with the same opening terms, the first steps of a run (identical step
grids), and veln / slipn after every step against the post-step formula
applied to that step's snapshot."""
import dataclasses

import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import problems as P
from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu

MINSIG, MAXSIG = 1.0, 300.0


def _setup():
    tri = P.rough_surface_triangles(6.0, 3.0, 0.25, hurst=0.8, alpha=1e-2, seed=1)
    gs = P.mesh_triangles(tri)
    gn = dataclasses.replace(gs, vcomp=1)
    gos = dataclasses.replace(gs, md=3)
    gon = dataclasses.replace(gs, md=3, vcomp=1)
    kw = dict(leaf=32, eps=1e-6)
    Hs, Hn, Hos, Hon = (H.HMatrix(g, **kw) for g in (gs, gn, gos, gon))
    Os, On, Oos, Oon = (O.HMatrix(g, nthreads=8, **kw) for g in (gs, gn, gos, gon))
    fr = P.friction_c3(gs)
    lod = Hs.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, gs, lod, amp=0.02, depth=2.0, width=1.0)
    # a patch below minsig, tau / sigma kept (same V): the opening rate
    # max(0, minsig - sigma) |V| is nonzero there
    zc = np.asarray(gs.zcol)[lod]
    low = zc > -1.0
    scale = np.where(low, 0.6 / y0[2::3], 1.0)
    y0[1::3] *= scale
    y0[2::3] *= scale
    return gs, fr, lod, y0, (Hs, Hn, Hos, Hon), (Os, On, Oos, Oon), low


def test_gpu_opening_against_oracle():
    gs, fr, lod, y0, (Hs, Hn, Hos, Hon), (Os, On, Oos, Oon), low = _setup()
    assert low.sum() > 10 and (~low).sum() > 10
    rng = np.random.default_rng(3)
    x = rng.standard_normal(gs.n)
    for Hg, Ho in ((Hos, Oos), (Hon, Oon)):
        yg, yo = Hg.matvec(x, H.PREC_F64), Ho.matvec(x)
        assert np.linalg.norm(yg - yo) <= 1e-6 * np.linalg.norm(yo) and np.abs(yo).max() > 0
    rk = H.RK(Hs, fr, lod, gs.ds, eps_r=1e-6, hn=Hn, minsig=MINSIG, maxsig=MAXSIG)
    rk.set_opening(Hos, Hon)
    so = O.Sim(Os, fr, lod, hn=On, minsig=MINSIG, maxsig=MAXSIG)
    so_plain = O.Sim(Os, fr, lod, hn=On, minsig=MINSIG, maxsig=MAXSIG)
    so.set_opening(Oos, Oon)
    dg, do, dp = rk.derivs(y0), so.derivs(y0), so_plain.derivs(y0)
    for c in range(3):
        assert np.abs(dg[c::3] - do[c::3]).max() <= 1e-7 * np.abs(do[c::3]).max() + 1e-300, c
    # the opening terms change the shear stress rate
    assert np.abs(do[1::3] - dp[1::3]).max() > 1e-6 * np.abs(dp[1::3]).max()
    yg, eg = rk.rkck(y0, 1e2)
    yo, eo = so.rkck(y0, 1e2)
    assert np.abs(yg - yo).max() <= 1e-9 * np.abs(yo).max()
    # first steps of a run: the step grids coincide
    rk.set_state(y0, 0.0, 1.0)
    info, hist, _ = rk.step(12, history=True)
    yr, _, _, _ = rk.get_state()
    yo = y0.copy()
    a = fr.a[lod]
    vel = 2 * fr.vref * np.exp(-yo[0::3] / a) * np.sinh(yo[1::3] / yo[2::3] / a)
    done, _, _, ho, _ = so.run(yo, 0.0, 1.0, 1e-6, 12, np.zeros(gs.n), vel, np.asarray(gs.ds)[lod])
    assert done == 12 and info["accepted"] == 12
    assert np.allclose(hist[:, 0], ho[:, 0], rtol=1e-9)
    for c in range(3):
        assert np.abs(yr[c::3] - yo[c::3]).max() <= 1e-9 * np.abs(yo[c::3]).max(), c


def test_gpu_opening_post_step_update():
    gs, fr, lod, y0, (Hs, Hn, Hos, Hon), _, low = _setup()
    rk = H.RK(Hs, fr, lod, gs.ds, eps_r=1e-6, hn=Hn, minsig=MINSIG, maxsig=MAXSIG)
    rk.set_opening(Hos, Hon)
    rk.set_output(velth=1e3, interval=1, capacity=64)
    rk.set_state(y0, 0.0, 1.0)
    info, hist, _ = rk.step(20, history=True)
    o = rk.get_output()
    assert list(o["k"]) == list(range(1, 21))
    sn = np.zeros(gs.n)
    for j in range(20):
        sig, v = o["sigma"][j], o["vel"][j]
        vn = v * (np.maximum(0.0, MINSIG - sig) + np.minimum(0.0, MAXSIG - sig))
        assert np.array_equal(o["veln"][j], vn)
        sn = sn + vn * hist[j, 1]
        assert np.array_equal(o["slipn"][j], sn)
    assert np.abs(o["slipn"][-1][low]).min() > 0 and np.abs(o["slipn"][-1][~low]).max() == 0
    vn, snf = rk.get_opening()
    assert np.array_equal(vn, o["veln"][-1]) and np.array_equal(snf, o["slipn"][-1])


def test_gpu_probes_record_the_local_output():
    """hbi_rk_set_probes / hbi_rk_get_probes (output_local, main_LH.f90:1495-1499):
    after every accepted step, V, slip, tau, sigma, mu, psi, veln, slipn of
    each probed element, equal to that step's snapshot (psi: to the state
    after single-step calls); rows belong to the last step call only."""
    gs, fr, lod, y0, (Hs, Hn, Hos, Hon), _, low = _setup()
    rk = H.RK(Hs, fr, lod, gs.ds, eps_r=1e-6, hn=Hn, minsig=MINSIG, maxsig=MAXSIG)
    rk.set_opening(Hos, Hon)
    rk.set_output(velth=1e3, interval=1, capacity=64)
    idx = np.array([int(np.where(low)[0][0]), int(np.where(low)[0][-1]), int(np.where(~low)[0][3])])
    rk.set_probes(idx)
    rk.set_state(y0, 0.0, 1.0)
    rk.step(12)
    pr = rk.get_probes()
    o = rk.get_output()
    assert pr.shape == (12, 3, 8)
    for c, f in enumerate(("vel", "slip", "tau", "sigma")):
        assert np.array_equal(pr[:, :, c], o[f][:, idx]), f
    assert np.array_equal(pr[:, :, 4], o["tau"][:, idx] / o["sigma"][:, idx])
    assert np.array_equal(pr[:, :, 6], o["veln"][:, idx]) and np.array_equal(pr[:, :, 7], o["slipn"][:, idx])
    assert np.abs(pr[-1, :2, 7]).min() > 0 and pr[-1, 2, 7] == 0
    y, slip, vel, _ = rk.get_state()
    assert np.array_equal(pr[-1, :, 5], y[0::3][idx]) and np.array_equal(pr[-1, :, 0], vel[idx])
    # a later call replaces the rows; psi after each single step
    for _ in range(3):
        rk.step(1)
        pr1 = rk.get_probes()
        y, slip, vel, _ = rk.get_state()
        assert pr1.shape == (1, 3, 8)
        assert np.array_equal(pr1[0, :, 5], y[0::3][idx]) and np.array_equal(pr1[0, :, 1], slip[idx])
    rk.get_output()
    rk.set_probes([])
    rk.step(2)
    assert rk.get_probes().shape == (2, 0, 8)


def test_gpu_probes_without_opening():
    """Without opening the veln and slipn columns are 0; rows equal the state
    after the step; a probe index outside the elements is refused."""
    g = P.mesh_2dvs(512, 40.0 / 512)
    fr = P.friction_c1(g)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod)
    rk = H.RK(Hm, fr, lod, g.ds, eps_r=1e-6)
    with pytest.raises(H.HbiError):
        rk.set_probes([0, 512])
    rk.set_probes([0, 100, 511])
    rk.set_state(y0, 0.0, 1.0)
    info, hist, _ = rk.step(30, history=True)
    pr = rk.get_probes()
    assert pr.shape == (30, 3, 8)
    assert (pr[:, :, 6:] == 0).all()
    y, slip, vel, _ = rk.get_state()
    idx = [0, 100, 511]
    assert np.array_equal(pr[-1, :, 0], vel[idx]) and np.array_equal(pr[-1, :, 1], slip[idx])
    assert np.array_equal(pr[-1, :, 2], y[1::3][idx]) and np.array_equal(pr[-1, :, 3], y[2::3][idx])
    # log10 max |V| of the history bounds each probe's V
    assert (np.log10(np.abs(pr[:, :, 0])).max(axis=1) <= hist[:, 4] + 1e-12).all()
