"""GPU: the 64-column H product and the lockstep ensemble integrator.

* hbi_hmat_matmul against hbi_hmat_matvec column by column, same factors:
  FP64 factors accumulate in FP64 in both, in different orders -> |dY| <=
  1e-12 max|Y|; FP32 / BF16 factors run on the FP32 matrix cores (X, Z
  rounded to FP32, FP32 sums of at most 32 columns added in FP64) ->
  |dY_i| <= 64 u32 sum_j |A_ij| |x_j| with A the dense kernel matrix
* ensemble realisation r against a single-run integration of the same
  friction parameters and initial state: the products differ in summation
  order only, so the step grids coincide over the first steps (states to
  1e-9 relative) and the slip-rate histories agree in time to 1e-5
"""
import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("label,make,leaf,eps,margin", [
    ("2dvs", lambda: P.mesh_2dvs(2048, 40 / 2048), 32, 1e-6, 1.0),
    ("3dph", lambda: P.mesh_3dph(48, 24, 1.0), 32, 1e-5, 1.0),
    ("3dph-highrank", lambda: P.mesh_3dph(64, 32, 1.0), 64, 1e-10, 1.0),
    # a loose BF16 bound: blocks of the mixed plan hold BF16 factors (the
    # BF16-row kernel of phase 1 and BF16 segments of phase 2)
    ("3dph-bf16", lambda: P.mesh_3dph(48, 24, 1.0), 32, 1e-5, 1e6),
])
def test_gpu_matmul_matches_matvec_columns(label, make, leaf, eps, margin):
    g = make()
    Hg = H.HMatrix(g, leaf=leaf, eps=eps, bf16_margin=margin)
    st = Hg.stats
    print(label, g.n, "kt_max", st["kt_max"], "BF16 blocks", st["nbf16_blocks"], "of", st["nlr"], flush=True)
    if label == "3dph-bf16":
        assert st["nbf16_blocks"] >= 16
    rng = np.random.default_rng(3)
    X = rng.standard_normal((g.n, 70))      # one full 64-column pass and a partial one
    lod = Hg.lod
    A = np.abs(H.entries(g, np.repeat(lod, g.n).astype(np.int32), np.tile(lod, g.n).astype(np.int32))).reshape(g.n, g.n)
    u32 = 2.0 ** -24
    worst = {}
    for prec in (H.PREC_F64, H.PREC_F32, H.PREC_MIXED):
        Y = Hg.matmul(X, prec)
        for c in range(X.shape[1]):
            y = Hg.matvec(X[:, c], prec)
            d = np.abs(Y[:, c] - y)
            if prec == H.PREC_F64:
                assert d.max() <= 1e-12 * np.abs(y).max(), (prec, c)
            else:
                bound = 64 * u32 * (A @ np.abs(X[:, c]))
                assert (d <= bound).all(), (prec, c, (d / bound).max())
                worst[prec] = max(worst.get(prec, 0.0), (d / bound).max())
        print("  max |dY| / bound", worst, flush=True)
    if label == "3dph-highrank":
        assert st["kt_max"] > 64      # exercises the split of ranks above 64


@pytest.mark.parametrize("label,make,leaf,eps,tol", [
    ("2dvs", lambda: P.mesh_2dvs(2048, 40 / 2048), 32, 1e-6, 1e-9),
    ("3dph", lambda: P.mesh_3dph(48, 24, 1.0), 32, 1e-5, 1e-6),
])
def test_gpu_matmul_against_oracle_and_dense(label, make, leaf, eps, tol):
    """Every column of the 64-column product against the oracle's H-matrix
    product and the dense kernel matrix (cluster order): FP64 factors within
    the ACA tolerance of the dense product (1e-9 at eps 1e-6, 1e-6 at 1e-5,
    as the single-vector product) and within 2 tol of the oracle's ACA
    product; FP32 / mixed factors within 2e-6 of the oracle's product (FP32
    rounding of factors, X and the slice sums)."""
    from oracle import oracle_lib as O
    g = make()
    Hg = H.HMatrix(g, leaf=leaf, eps=eps, bf16_margin=1.0)
    Ho = O.HMatrix(g, leaf=leaf, eps=eps, nthreads=8)
    assert np.array_equal(Hg.lod, Ho.lod)
    A = Ho.dense()
    X = np.random.default_rng(5).standard_normal((g.n, 64))
    YD = A @ X
    YO = np.stack([Ho.matvec(X[:, c], nthreads=8) for c in range(64)], axis=1)
    for prec, bd, bo in ((H.PREC_F64, tol, 2 * tol), (H.PREC_F32, None, 2e-6), (H.PREC_MIXED, None, 2e-6)):
        Y = Hg.matmul(X, prec)
        ed = max(np.linalg.norm(Y[:, c] - YD[:, c]) / np.linalg.norm(YD[:, c]) for c in range(64))
        eo = max(np.linalg.norm(Y[:, c] - YO[:, c]) / np.linalg.norm(YO[:, c]) for c in range(64))
        print("%s prec %d: max column error vs dense %.3e, vs oracle %.3e" % (label, prec, ed, eo), flush=True)
        if bd is not None:
            assert ed <= bd
        assert eo <= bo


def _c1_realisations(n, nreal):
    g = P.mesh_2dvs(n, 40.0 / n)
    base = P.friction_c1(g)
    frs = []
    for r in range(nreal):
        f = P.friction_c1(g)
        f.dc = base.dc * (1.0 + 0.02 * r)
        f.b = base.b * (1.0 + 0.01 * (r % 5))
        frs.append(f)
    return g, frs


def _crossings(t, lv, levels=(-7.5, -7.0, -6.0, -5.0, -4.0, -3.0, -2.0, -1.0)):
    out = {}
    for L in levels:
        k = int(np.argmax(lv > L))
        if lv[k] > L and k > 0:
            out[L] = t[k - 1] + (L - lv[k - 1]) / (lv[k] - lv[k - 1]) * (t[k] - t[k - 1])
    return out


def _onset(t, lv, thresh=-3.0):
    k = int(np.argmax(lv > thresh))
    return t[k] if lv[k] > thresh else None


@pytest.mark.parametrize("nreal,prec", [(3, H.PREC_F64), (64, H.PREC_F64), (5, H.PREC_F32)])
def test_gpu_ensemble_matches_single_runs(nreal, prec):
    """Realisation r of the ensemble against a single-run integration of the
    same parameters: identical step grids over the first 12 steps (states to
    1e-9), then, through the first event, the event onset time to 1e-5.
    Once steps start being rejected, the error estimate of this problem moves
    by ~1% under 1e-10 relative changes of t (the single-run GPU integrator
    and the oracle part there too), so step grids are compared only before;
    after that, the times at which log10 Vmax crosses levels -7.5 .. -1
    (interpolated within each history, so independent of the other grid)
    agree to 1e-3 as a guard against gross errors (the coarse early steps,
    ~2e5 s, move the interpolated crossings of the flat early levels by up to
    2e-4 already between two FP64 integrations; FP32 matrix cores in the
    ensemble add ~2e-4)."""
    g, frs = _c1_realisations(1024, nreal)
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    ys = []
    for r, f in enumerate(frs):
        y = P.init_state(f, lod)
        P.perturb_state(y, g, lod, amp=0.02 * (1 + 0.005 * r))
        ys.append(y)
    ys = np.array(ys)
    n1, n2 = 12, 400
    ens = H.Ensemble(Hg, frs, lod, hdiv=0.0, eps_r=1e-6, prec=prec)
    ens.set_state(ys, 0.0, 1.0)
    attempts = ens.step(n1)
    es = ens.get_state()
    assert (es["accepted"] == n1).all() and (es["status"] == 0).all()
    assert attempts >= n1 + es["rejected"].max()
    _, eh = ens.step(n2, history=True)
    es2 = ens.get_state()
    print("nreal %d: rejected %s" % (nreal, es2["rejected"][:8]), flush=True)
    assert (es2["accepted"] == n1 + n2).all()
    assert np.array_equal(eh[:, :, 3].sum(axis=1), es2["rejected"] - es["rejected"])
    for r in sorted({0, nreal // 2, nreal - 1}):
        rk = H.RK(Hg, frs[r], lod, g.ds, eps_r=1e-6, prec=prec)
        rk.set_state(ys[r], 0.0, 1.0)
        rk.step(n1)
        y1, slip1, vel1, inf = rk.get_state()
        assert inf["accepted"] == n1 and inf["rejected"] == es["rejected"][r]
        assert abs(es["t"][r] - inf["t"]) <= 1e-9 * inf["t"]
        for c in range(3):
            ref = y1[c::3]
            assert np.abs(es["y"][r, c::3] - ref).max() <= 1e-9 * np.abs(ref).max(), (r, c)
        assert np.abs(es["slip"][r] - slip1).max() <= 1e-9 * np.abs(slip1).max() + 1e-15
        assert np.abs(es["vel"][r] - vel1).max() <= 1e-8 * np.abs(vel1).max()
        _, sh, _ = rk.step(n2, history=True)
        ts, lvs = sh[:, 0], sh[:, 4]
        te, lve = eh[r, :, 0], eh[r, :, 4]
        on_s, on_e = _onset(ts, lvs), _onset(te, lve)
        assert (on_s is None) == (on_e is None)
        # times at which log10 Vmax first crosses a set of levels (each
        # interpolated within its own history): grid-independent
        cs, ce = _crossings(ts, lvs), _crossings(te, lve)
        both = [k for k in cs if k in ce]
        rel = max(abs(ce[k] - cs[k]) / cs[k] for k in both)
        print("  r %d: onset %s vs %s, %d level crossings, max relative time difference %.2e"
              % (r, on_e, on_s, len(both), rel), flush=True)
        if on_s is not None:
            assert abs(on_e - on_s) <= 1e-5 * on_s
        assert len(both) >= 3 and rel <= 1e-3


def test_gpu_ensemble_stop_is_per_realisation():
    """Stops are per realisation: one leaves [velmin, velmax]
    (main_LH.f90:1438-1445), one underflows the step size (an exact steady
    state has errmax <= tiny, which rkqs treats as a failed step and halves h
    until t + h == t, main_LH.f90:2309-2352; the oracle does the same), and
    the others complete their steps, also on a second call."""
    g, frs = _c1_realisations(512, 4)
    frs[2] = frs[0]                # unmodified parameters: init_state is their exact steady state
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    ys = np.array([P.init_state(f, lod) for f in frs])
    P.perturb_state(ys[0], g, lod)
    ys[1, 1::3] *= 0.8             # max V ~ 2e-12 < velmin
    P.perturb_state(ys[3], g, lod)
    ens = H.Ensemble(Hg, frs, lod, eps_r=1e-6, velmin=1e-11)
    ens.set_state(ys, 0.0, 1.0)
    ens.step(200)
    es = ens.get_state()
    assert list(es["status"]) == [0, 1, H.E_STEP, 0]
    assert list(es["accepted"]) == [200, 1, 0, 200]
    ens.step(10)
    es2 = ens.get_state()
    assert list(es2["accepted"]) == [210, 1, 0, 210]
    # the oracle underflows on the same steady state
    from oracle import oracle_lib as O
    Ho = O.HMatrix(g, leaf=32, eps=1e-6)
    so = O.Sim(Ho, frs[2], lod)
    y = ys[2].copy()
    a = frs[2].a[lod]
    vel = 2 * frs[2].vref * np.exp(-y[0::3] / a) * np.sinh(y[1::3] / y[2::3] / a)
    done, *_ = so.run(y, 0.0, 1.0, 1e-6, 10, np.zeros(g.n), vel, np.asarray(g.ds)[lod])
    assert done == 0


def test_gpu_ensemble_continues_from_another_ensembles_clock():
    """hbi_ens_set_clock / hbi_ens_get_clock: an ensemble set to another's
    states, times and next step sizes continues exactly as that ensemble
    does (rkqs: the next step depends on the state and h only), also with
    realisations at different times."""
    g, frs = _c1_realisations(1024, 6)
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    ys = np.array([P.init_state(f, lod) for f in frs])
    for y in ys:
        P.perturb_state(y, g, lod)
    A = H.Ensemble(Hg, frs, lod, eps_r=1e-6)
    A.set_state(ys, 0.0, 1.0)
    A.step(150)
    st = A.get_state()
    t, h = A.get_clock()
    assert np.array_equal(t, st["t"]) and (h > 0).all() and np.unique(t).size > 1
    B = H.Ensemble(Hg, frs, lod, eps_r=1e-6)
    B.set_state(st["y"], 0.0, 1.0)
    B.set_clock(t, h)
    tb, hb = B.get_clock()
    assert np.array_equal(tb, t) and np.array_equal(hb, h)
    A.step(40)
    B.step(40)
    sa, sb = A.get_state(), B.get_state()
    assert np.array_equal(sa["t"], sb["t"])
    assert np.array_equal(sa["y"], sb["y"])
    assert list(sb["accepted"]) == [40] * 6 and np.array_equal(sb["rejected"], sa["rejected"] - st["rejected"])
    with pytest.raises(H.HbiError):
        B.set_clock(t, -h)
    A.close()
    B.close()
