"""GPU parity: libhbi.so (through the C ABI) against the oracle and the
reference's golden vectors.  Tolerances are written in each test:

* kernel entries (FP64, device libm): |gpu - ref| <= 1e-11 max|ref| per fixture
* partition / permutation: identical integers
* H x (FP64): GPU and oracle are both ACA+ approximations of the same dense
  matrix; each is checked against the dense product to its ACA tolerance and
  against each other
* FP32 / mixed factors: relative to the FP64 product
* derivs, rkck, trajectories: relative error bounds stated inline
"""
import math

import numpy as np
import pytest

from conftest import GOLDEN_FILES, halfspace_tolerance, load_golden
from hbi_amd import lib as H
from hbi_amd import problems as P
from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu


def _relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name", GOLDEN_FILES)
def test_gpu_entries_match_reference_golden(name):
    g, ii, jj, vals = load_golden(name)
    print(name, flush=True)
    got = H.entries(g, ii, jj)
    scale = np.nanmax(np.abs(vals))
    assert np.array_equal(np.isnan(got), np.isnan(vals))
    ok = ~np.isnan(vals)
    err = np.abs(got[ok] - vals[ok])
    if "3dht" in name:
        # conditioned by the steepest side of the source triangle
        assert (err <= halfspace_tolerance(g, jj[ok], vals[ok])).all()
    else:
        assert err.max() <= 1e-11 * scale, (err.max(), scale)


@pytest.mark.parametrize("make,leaf", [
    (lambda: P.mesh_2dvs(2048, 40 / 2048), 32),
    (lambda: P.mesh_3dph(64, 32, 1.0), 32),
    (lambda: P.mesh_triangles(P.rough_surface_triangles(6.0, 3.0, 0.25, seed=4)), 16)])
def test_gpu_partition_identical_to_oracle(make, leaf):
    g = make()
    Hg = H.HMatrix(g, leaf=leaf, eps=1e-4)
    Ho = O.HMatrix(g, leaf=leaf, partition_only=True)
    assert np.array_equal(Hg.lod, Ho.lod)
    assert np.array_equal(Hg.leaves[:, :5], Ho.leaves[:, :5])


def _curve(n, length, amp, y0):
    x = np.linspace(0.0, length, n + 1)
    y = y0 + amp * np.sin(6 * np.pi * x / length) + 0.3 * amp * np.sin(22 * np.pi * x / length)
    return x[:-1], x[1:], y[:-1], y[1:]


CASES = [
    ("2dvs", lambda: P.mesh_2dvs(2048, 40 / 2048), 32, 1e-6, 1e-9),
    ("3dph", lambda: P.mesh_3dph(48, 24, 1.0), 32, 1e-5, 1e-6),
    ("3dnt", lambda: P.mesh_triangles(P.rough_surface_triangles(8.0, 4.0, 0.25, seed=9)), 32, 1e-4, 1e-4),
    ("3dht", lambda: P.mesh_triangles(P.rough_surface_triangles(8.0, 4.0, 0.25, seed=9), halfspace=True), 32, 1e-4,
     1e-4),
    ("2dp", lambda: P.mesh_2dp(1000, 0.1), 32, 1e-6, 1e-6),
    ("3dp", lambda: P.mesh_3dp(48, 24, 0.5), 32, 1e-5, 1e-5),
    ("2dph", lambda: P.mesh_2dph(1600, 0.025, 30.0), 32, 1e-6, 1e-6),
    ("2dn", lambda: P.mesh_segments(*_curve(800, 30.0, 0.5, 0.0)), 32, 1e-6, 1e-6),
    ("25d", lambda: P.mesh_segments(*_curve(800, 30.0, 0.5, 0.0), kind=P.KIND_25D, w=10.0), 32, 1e-6, 1e-6),
    ("2dpfz", lambda: P.mesh_2dpfz(1000, 0.1, 0.5), 32, 1e-6, 1e-6),
    ("2dn_vector", lambda: P.mesh_segments(*_curve(800, 30.0, 0.5, 0.0), kind=P.KIND_2DNV), 32, 1e-6, 1e-6),
    ("2dna", lambda: P.mesh_segments(*_curve(800, 30.0, 0.5, 0.0), kind=P.KIND_2DNA), 32, 1e-6, 1e-6),
]


@pytest.mark.parametrize("label,make,leaf,eps,tol", CASES)
def test_gpu_hmatrix_fp64_against_dense_and_oracle(label, make, leaf, eps, tol):
    g = make()
    print(label, g.n, flush=True)
    Hg = H.HMatrix(g, leaf=leaf, eps=eps)
    Ho = O.HMatrix(g, leaf=leaf, eps=eps, nthreads=8)
    A = Ho.dense()
    rng = np.random.default_rng(5)
    for _ in range(3):
        x = rng.standard_normal(g.n)
        yd = A @ x
        yg = Hg.matvec(x, H.PREC_F64)
        yo = Ho.matvec(x)
        assert _relerr(yg, yd) <= tol, _relerr(yg, yd)
        assert _relerr(yg, yo) <= 2 * tol
    # ranks: the same ACA+ on the same entries; pivots may differ only where
    # the device libm moves a near-tie
    kg, ko = Hg.leaves[:, 5], Ho.leaves[:, 5]
    lr = Hg.leaves[:, 0] == 1
    assert np.mean(kg[lr] == ko[lr]) >= 0.9
    assert np.abs(kg[lr] - ko[lr]).max() <= 3
    st = Hg.stats
    assert st["nlr"] == int(lr.sum()) and st["ncapped"] == 0


def test_gpu_lowrank_blocks_match_oracle_factors():
    g = P.mesh_2dvs(1024, 40 / 1024)
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    Ho = O.HMatrix(g, leaf=32, eps=1e-6)
    Lg = Hg.leaves
    for ip in np.nonzero(Lg[:, 0] == 1)[0][:40]:
        _, _, nr, _, nc, kg = Lg[ip]
        dg = Hg.leaf_data(ip)
        do = Ho.leaf_data(ip)
        ko = Ho.leaves[ip, 5]
        Bg = dg[nc * kg:].reshape(kg, nr).T @ dg[:nc * kg].reshape(kg, nc)
        Bo = do[nc * ko:].reshape(ko, nr).T @ do[:nc * ko].reshape(ko, nc)
        assert np.linalg.norm(Bg - Bo) <= 1e-8 * np.linalg.norm(Bo) + 1e-12


@pytest.mark.parametrize("leaf,chunk_kb", [(32, 0), (64, 0), (64, 32)])
def test_gpu_dense_blocks_in_either_launch(leaf, chunk_kb):
    """The single-vector product with its dense blocks streamed by the first
    launch (partial slots per tile) or by the second (inside the row-tile
    item): the same sum up to its order, against the oracle's product."""
    g = P.mesh_3dph(64, 32, 1.0)
    Ho = O.HMatrix(g, leaf=leaf, eps=1e-5, nthreads=8)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(g.n)
    yo = Ho.matvec(x)
    ys = {}
    for dl in (1, 2):
        Hg = H.HMatrix(g, leaf=leaf, eps=1e-5, bf16_margin=1.0, chunk_kb=chunk_kb, dense_launch=dl)
        for prec in (H.PREC_F64, H.PREC_F32, H.PREC_MIXED):
            print("leaf", leaf, "chunk_kb", chunk_kb, "dense_launch", dl, "prec", prec, flush=True)
            assert Hg.stats["dense_launch"][prec] == dl
            ys[dl, prec] = Hg.matvec(x, prec)
        Hg.close()
    for prec in (H.PREC_F64, H.PREC_F32, H.PREC_MIXED):
        assert _relerr(ys[1, prec], ys[2, prec]) <= 1e-13
    assert _relerr(ys[2, H.PREC_F64], yo) <= 1e-5
    assert _relerr(ys[2, H.PREC_F32], yo) <= 1e-5


def test_gpu_dense_leaves_match_oracle_entries():
    g = P.mesh_3dph(32, 16, 1.0)
    Hg = H.HMatrix(g, leaf=32, eps=1e-5)
    Ho = O.HMatrix(g, leaf=32, eps=1e-5)
    L = Hg.leaves
    for ip in np.nonzero(L[:, 0] == 2)[0][:30]:
        a, b = Hg.leaf_data(ip), Ho.leaf_data(ip)
        assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max()


@pytest.mark.parametrize("label,make,leaf,eps,tol", CASES)
def test_gpu_reduced_precision_products(label, make, leaf, eps, tol):
    g = make()
    Hg = H.HMatrix(g, leaf=leaf, eps=eps, bf16_margin=1.0)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(g.n)
    y64 = Hg.matvec(x, H.PREC_F64)
    y32 = Hg.matvec(x, H.PREC_F32)
    ymx = Hg.matvec(x, H.PREC_MIXED)
    assert _relerr(y32, y64) <= 2e-6
    # BF16 only on blocks whose rounding stays below their ACA tolerance
    assert _relerr(ymx, y64) <= max(2e-6, 10 * tol)
    # deterministic: the same product twice is bitwise identical
    assert np.array_equal(Hg.matvec(x, H.PREC_MIXED), ymx)


def _c1(n=2048, leaf=32, eps=1e-6):
    g = P.mesh_2dvs(n, 40.0 / n)
    fr = P.friction_c1(g)
    return g, fr


def test_gpu_derivs_and_rkck_match_oracle():
    g, fr = _c1()
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    Ho = O.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    y = P.init_state(fr, lod)
    rng = np.random.default_rng(1)
    y[0::3] += 0.01 * rng.standard_normal(g.n)
    rk = H.RK(Hg, fr, lod, g.ds, eps_r=1e-6)
    so = O.Sim(Ho, fr, lod)
    dg, do = rk.derivs(y), so.derivs(y)
    for c in range(3):
        assert np.abs(dg[c::3] - do[c::3]).max() <= 1e-7 * np.abs(do[c::3]).max() + 1e-300
    for h in (1e2, 1e4):   # |errmax| ~ 3 and 2e4 at eps 1e-6: an accepted and a rejected step
        yg, eg = rk.rkck(y, h)
        yo, eo = so.rkck(y, h)
        assert np.abs(yg - yo).max() <= 1e-9 * np.abs(yo).max()
        assert np.abs(eg - eo).max() <= 1e-6 * np.abs(eo).max() + 1e-14


def _onset(hist, thresh=-3.0):
    k = int(np.argmax(hist[:, 4] > thresh))
    return hist[k, 0] if hist[k, 4] > thresh else None


def _window_rel(h0, v0, h1, v1, t_end):
    """max relative slip-rate difference on a common time grid up to t_end"""
    tg = np.linspace(max(h0[0, 0], h1[0, 0]), t_end, 4000)
    return max(np.abs(np.interp(tg, h0[:, 0], np.log(v0[:, r])) - np.interp(tg, h1[:, 0], np.log(v1[:, r]))).max()
               for r in range(v0.shape[1]))


def test_gpu_rk_trajectory_c1_matches_oracle():
    """2-D antiplane C1 setup (seeded nucleation), 600 accepted steps through
    the first event.  Trajectories are compared in time, not by step index:
    an accept/reject decision flipped by a last-bit difference re-grids the
    whole run (a 1e-12 relative perturbation of tau does that to the oracle
    itself), so the checks are the event onset time and the slip-rate history
    before the fast nucleation phase."""
    g, fr = _c1()
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    Ho = O.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod)
    nsteps = 600
    rec = np.arange(0, g.n, 64, dtype=np.int32)
    rk = H.RK(Hg, fr, lod, g.ds, eps_r=1e-6)
    rk.set_state(y0, 0.0, 1.0)
    info, hist, vh = rk.step(nsteps, history=True, rec=rec)
    so = O.Sim(Ho, fr, lod)
    yo = y0.copy()
    a = fr.a[lod]
    vel = 2 * fr.vref * np.exp(-yo[0::3] / a) * np.sinh(yo[1::3] / yo[2::3] / a)
    done, x, dt, ho, vho = so.run(yo, 0.0, 1.0, 1e-6, nsteps, np.zeros(g.n), vel, np.asarray(g.ds)[lod], rec=rec)
    assert done == nsteps and info["accepted"] == nsteps
    on_g, on_o = _onset(hist), _onset(ho)
    assert on_g is not None and on_o is not None
    rel_onset = abs(on_g - on_o) / on_o
    rel_early = _window_rel(ho, vho, hist, vh, 0.5 * on_o)
    rel_pre = _window_rel(ho, vho, hist, vh, 0.9 * on_o)
    print("trajectory: onset %.6e s (oracle %.6e), rel %.3e; |dV/V| to 0.5 onset %.3e, to 0.9 onset %.3e; "
          "rejected %d vs %d" % (on_g, on_o, rel_onset, rel_early, rel_pre, info["rejected"], int(ho[:, 3].sum())))
    assert rel_onset <= 1e-5
    assert rel_early <= 1e-5


def test_gpu_rk_history_and_stop_criteria():
    g, fr = _c1(512)
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    rk = H.RK(Hg, fr, lod, g.ds, eps_r=1e-6, velmax=1.5e-9)
    y0 = P.init_state(fr, lod)
    y0[1::3] *= 1.01      # loaded above steady state: V grows
    rk.set_state(y0, 0.0, 1.0)
    info, hist, _ = rk.step(5000, history=True)
    assert info["status"] == 1     # stopped on max V > velmax (main_LH.f90:1438-1441)
    n = info["accepted"]
    assert 0 < n < 5000 and hist[n - 1, 4] > math.log10(1.5e-9)
    assert (np.diff(hist[:n, 0]) > 0).all()


def test_gpu_c2_full_size_structure_and_linearity():
    """32768-element planar fault, leaf 64, eps 1e-5 (the benchmark matrix):
    leaf counts of the reference partition, linearity, and rows against the
    exact kernel."""
    g = P.mesh_3dph(256, 128, 0.390625)
    Hg = H.HMatrix(g, leaf=64, eps=1e-5, bf16_margin=1.0)
    st = Hg.stats
    assert (st["nlr"], st["lr_dims"], st["ndense"], st["dense_entries"]) == (4522, 2599424, 8520, 34897920)
    assert st["ncapped"] == 0
    rng = np.random.default_rng(11)
    x1, x2 = rng.standard_normal(g.n), rng.standard_normal(g.n)
    for prec in (H.PREC_F64, H.PREC_F32, H.PREC_MIXED):
        y1, y2 = Hg.matvec(x1, prec), Hg.matvec(x2, prec)
        y12 = Hg.matvec(2.0 * x1 - 3.0 * x2, prec)
        assert _relerr(y12, 2.0 * y1 - 3.0 * y2) <= 1e-12
    # exact rows: (H x)_i against sum_j K(lod_i, lod_j) x_j for sampled rows
    lod = Hg.lod
    x = rng.standard_normal(g.n)
    y = Hg.matvec(x, H.PREC_F64)
    rows = rng.integers(0, g.n, 6)
    for r in rows:
        kr = H.entries(g, np.full(g.n, lod[r], dtype=np.int32), lod.astype(np.int32))
        assert abs(y[r] - kr @ x) <= 1e-4 * np.abs(kr).sum() * np.abs(x).max() / math.sqrt(g.n) + 1e-9 * abs(kr @ x)


def test_gpu_normal_stress_matrix_against_oracle():
    """sigmaconst F: a second H-matrix of normal tractions drives dsigma/dt
    (main_LH.f90:2100-2125, 2165-2180 with limitsigma).  60-degree dipping
    thrust fault (normal tractions change with slip), 48 x 24 elements:
    both matrices against the oracle's, derivs, an RK step, and the first
    steps of a run (identical step grids) against the oracle."""
    import dataclasses
    imax, jmax, ds0 = 48, 24, 1.0
    gs = P.mesh_3dph(imax, jmax, ds0, dipangle=60.0, crake=90.0)
    gn = dataclasses.replace(gs, vcomp=1)
    Hs, Hn = H.HMatrix(gs, leaf=32, eps=1e-6), H.HMatrix(gn, leaf=32, eps=1e-6)
    Os, On = O.HMatrix(gs, leaf=32, eps=1e-6, nthreads=8), O.HMatrix(gn, leaf=32, eps=1e-6, nthreads=8)
    assert np.array_equal(Hs.lod, Hn.lod) and np.array_equal(Hs.lod, Os.lod)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(gs.n)
    yn, yno = Hn.matvec(x, H.PREC_F64), On.matvec(x)
    assert _relerr(yn, yno) <= 1e-6 and np.abs(yn).max() > 1e-3 * np.abs(Hs.matvec(x, H.PREC_F64)).max()
    fr = P.friction_c2(gs, imax, jmax, ds0, seed=3, patch=(20.0, 10.0))
    lod = Hs.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, gs, lod, amp=0.02, depth=8.0, width=2.0)
    rk = H.RK(Hs, fr, lod, gs.ds, hdiv=ds0, eps_r=1e-6, hn=Hn)
    so = O.Sim(Os, fr, lod, hdiv=ds0, hn=On)
    dg, do = rk.derivs(y0), so.derivs(y0)
    for c in range(3):
        assert np.abs(dg[c::3] - do[c::3]).max() <= 1e-7 * np.abs(do[c::3]).max() + 1e-300, c
    assert np.abs(do[2::3]).max() > 0.0           # sigma moves
    yg, eg = rk.rkck(y0, 1e3)
    yo, eo = so.rkck(y0, 1e3)
    assert np.abs(yg - yo).max() <= 1e-9 * np.abs(yo).max()
    # first steps of a run: the step grids coincide
    rk.set_state(y0, 0.0, 1.0)
    info, hist, _ = rk.step(15, history=True)
    yr, _, _, _ = rk.get_state()
    yo = y0.copy()
    a = fr.a[lod]
    vel = 2 * fr.vref * np.exp(-yo[0::3] / a) * np.sinh(yo[1::3] / yo[2::3] / a)
    done, t_o, _, ho, _ = so.run(yo, 0.0, 1.0, 1e-6, 15, np.zeros(gs.n), vel, np.asarray(gs.ds)[lod])
    assert done == 15 and info["accepted"] == 15
    assert np.allclose(hist[:, 0], ho[:, 0], rtol=1e-9)
    for c in range(3):
        assert np.abs(yr[c::3] - yo[c::3]).max() <= 1e-9 * np.abs(yo[c::3]).max(), c


def test_gpu_rough_triangle_workload_small():
    """The rough-surface triangular workload (slip law, TD kernel) end to end
    at a small size: structure, no capped ranks, and the mixed single run
    tracking the FP64 one over the first steps."""
    from hbi_amd import workloads as W
    r = W.c3(length=12.0, width=6.0, edge=0.25, leaf=32, eps=1e-5, steps=40)
    assert r["n"] == 2 * 48 * 24 and r["hmat"]["ncapped"] == 0
    f64, mx = r["rk"]["f64"], r["rk"]["mixed"]
    assert f64["status"] == 0 and mx["status"] == 0
    assert abs(mx["t_end_s"] - f64["t_end_s"]) <= 1e-6 * f64["t_end_s"]
