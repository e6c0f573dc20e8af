"""GPU: the batched ensemble workload -- 64 friction realisations of the
128 x 128 planar fault (16 384 elements, a-b ~ N(-0.003, 0.001) in the
velocity-weakening patch, Dc ~ lognormal(median 15 mm, sigma_log 0.3)) on
one H-matrix, advanced in lockstep with a 64-column product per RK stage --
against the oracle and against the FP64 ensemble.

* FP64 ensemble, first K steps of three realisations against the oracle's
  single runs of the same parameters: the same accept / reject decisions,
  t to 1e-9 (h follows errmax^-1/5, and the device's and the oracle's ACA+
  factors differ at the 1e-10 level), states to 1e-9 of each component's
  size, slip to 1e-8.
* Reduced-precision ensembles (FP32 factors on the FP32 matrix cores,
  HBI_PREC_F32, or split into BF16 parts on the BF16 matrix cores,
  HBI_PREC_MIXED) over the same steps:
  the same decisions; the embedded error estimate is a small difference of
  two solutions, so FP32 product rounding moves errmax by ~1e-3 relative
  and with it h (as errmax^-1/5) and t (~1e-4): t to 1e-3, states to 1e-5
  and slip to 1e-4 of their size at the same step.
* Through the first events (1500 steps): the reduced-precision ensembles'
  event onsets agree with the FP64 ensemble's to 1e-5 (FP32 cores) and 2e-5
  (BF16 cores) relative.
* Histories at identical times: free-running runs of different precision
  take different step grids (FP32 rounding moves the error estimate), and
  interpolating coarse steps, or slip integrated by the trapezoid rule on
  different grids, differs by far more than the precision does.  So the
  mixed ensemble and the oracle replay the FP64 run's step grid
  (hbi_ens_set_step_grid, og_run_grid): the FP64 replays agree with the
  oracle's to 1e-9 over the first 120 steps, the reduced-precision ones to
  5e-5 / 1e-4 there; over the interseismic period their slip rates drift
  from the FP64 run's (figures in the test).
"""
import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import workloads as W
from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu

K = 30
CHECK = (0, 31, 63)


@pytest.fixture(scope="module")
def c4a():
    print("C4a setup: 128 x 128 fault, 64 realisations", flush=True)
    g, Hm, frs, lod, ys, probes = W.c4a_setup(64)
    return g, Hm, frs, lod, ys, probes


@pytest.fixture(scope="module")
def oracle_runs(c4a):
    g, Hm, frs, lod, ys, probes = c4a
    print("C4a oracle: H-matrix and %d steps of realisations %s" % (K, CHECK), flush=True)
    Ho = O.HMatrix(g, leaf=64, eps=1e-5, nthreads=16)
    assert np.array_equal(Ho.lod, lod)
    area = np.asarray(g.ds)[lod]
    out = {}
    for r in CHECK:
        f = frs[r]
        so = O.Sim(Ho, f, lod, hdiv=W.C4A_DS0, nthreads=16)
        y = ys[r].copy()
        a = f.a[lod]
        vel = 2 * f.vref * np.exp(-y[0::3] / a) * np.sinh(y[1::3] / y[2::3] / a)
        slip = np.zeros(g.n)
        done, t, _, hist, _ = so.run(y, 0.0, 1.0, 1e-6, K, slip, vel, area)
        assert done == K
        out[r] = (y, slip, hist)
    return out


@pytest.mark.parametrize("prec,ytol,ttol", [(H.PREC_F64, 1e-9, 1e-9), (H.PREC_F32, 1e-5, 1e-3),
                                           (H.PREC_MIXED, 1e-5, 1e-3)])
def test_gpu_c4a_first_steps_against_oracle(c4a, oracle_runs, prec, ytol, ttol):
    g, Hm, frs, lod, ys, probes = c4a
    E = H.Ensemble(Hm, frs, lod, hdiv=W.C4A_DS0, eps_r=1e-6, prec=prec)
    E.set_state(ys, 0.0, 1.0)
    _, eh = E.step(K, history=True)
    es = E.get_state()
    assert (es["accepted"] == K).all()
    for r in CHECK:
        yo, so, ho = oracle_runs[r]
        dt = (np.abs(eh[r, :, 0] - ho[:, 0]) / ho[:, 0]).max()
        de = (np.abs(eh[r, :, 2] - ho[:, 2]) / ho[:, 2]).max()
        worst = max(np.abs(es["y"][r, c::3] - yo[c::3]).max() / np.abs(yo[c::3]).max() for c in range(3))
        ds = np.abs(es["slip"][r] - so).max() / np.abs(so).max()
        print("prec %d realisation %d: t max rel %.3e, errmax max rel %.3e, state max rel %.3e, slip max rel %.3e"
              % (prec, r, dt, de, worst, ds))
        assert np.array_equal(eh[r, :, 3], ho[:, 3])      # the same accept / reject decisions
        assert dt <= ttol and worst <= ytol and ds <= 10 * ytol


@pytest.fixture(scope="module")
def free_f64(c4a):
    """The FP64 ensemble under rkqs through the first events (1500 steps)."""
    g, Hm, frs, lod, ys, probes = c4a
    print("C4a FP64 ensemble, 1500 steps", flush=True)
    hf, sf = W.run_ensemble(Hm, frs, lod, ys, probes, H.PREC_F64, 1500, W.C4A_DS0)
    assert sf["stopped"] == 0
    return hf, sf


# per reduced-precision path: (onset bound, |d ln V| bound below log10 V -8, the same below -6)
PATHS = {H.PREC_F32: (1e-5, 2e-4, 0.2), H.PREC_MIXED: (2e-5, 2e-3, 1.0)}


@pytest.mark.parametrize("prec", [H.PREC_F32, H.PREC_MIXED])
def test_gpu_c4a_reduced_precision_onsets(c4a, free_f64, prec):
    """Event onsets of the FP32-core (measured 1.0e-6) and BF16-core (9.7e-6)
    ensembles against the FP64 ensemble."""
    g, Hm, frs, lod, ys, probes = c4a
    hf, sf = free_f64
    hm, sm = W.run_ensemble(Hm, frs, lod, ys, probes, prec, 1500, W.C4A_DS0)
    em = W.ensemble_errors(hf, hm)
    print("prec %d: steps/s FP64 %.0f, reduced %.0f; vs FP64 %s" % (prec, sf["steps_per_s"], sm["steps_per_s"], em))
    assert sm["stopped"] == 0
    assert em["realisations_with_onset"] >= 32
    assert em["onset_rel_max"] <= PATHS[prec][0]


def _replay(c4a, prec, grid, steps):
    g, Hm, frs, lod, ys, probes = c4a
    E = H.Ensemble(Hm, frs, lod, hdiv=W.C4A_DS0, eps_r=1e-6, prec=prec)
    E.set_probes(probes)
    E.set_state(ys, 0.0, 1.0)
    E.set_step_grid(grid)
    att, h = E.step(steps, history=True)
    es = E.get_state()
    print("replay prec %d: %d attempts, accepted %s, status %s" % (prec, att, np.unique(es["accepted"]),
                                                                  np.unique(es["status"])), flush=True)
    E.close()
    return h


@pytest.mark.parametrize("prec", [H.PREC_F32, H.PREC_MIXED])
def test_gpu_c4a_histories_on_the_fp64_step_grid(c4a, free_f64, prec):
    """Reduced-precision and FP64 ensembles replayed on the FP64 run's step
    grid: the same times, no interpolation.  The FP64 replay equals the FP64
    run it replays.  The reduced-precision replay's slip rates (maximum and
    probes) and probe slips against the FP64 run's, per phase of each
    realisation (measured, FP32 cores / BF16 cores):

    * log10 Vmax < -8 (near plate rate): 4e-5 / 4e-4 in ln V;
    * log10 Vmax < -6 (the quasi-static phase up to the nucleation): 0.035 /
      0.42 in ln V.  V depends on tau as exp(tau / (a sigma)) with a sigma =
      0.5 MPa, so the product's rounding, integrated into tau over the
      interseismic period, moves V by 1e-5 .. 1e-3 and the nucleation
      amplifies that; the BF16 cores' truncation toward zero within each
      group of 8 columns is biased and accumulates further.
    The event onset times are the criterion past that (test above).  Past -6
    the onsets differ by hundreds of seconds and a forced grid carried on into
    the rupture, with no error control, diverges; the replay stops there."""
    hf, _ = free_f64
    grid = hf[:, :, 1]
    if prec == H.PREC_F32:
        hr = _replay(c4a, H.PREC_F64, grid, 1500)
        assert np.array_equal(hr[:, :, 0], hf[:, :, 0])
        assert np.array_equal(hr[:, :, 4:], hf[:, :, 4:])
    hm = _replay(c4a, prec, grid, 1500)
    for r in range(hf.shape[0]):
        k = int(np.argmax(hf[r, :, 4] > -6.0))
        end = k if hf[r, k, 4] > -6.0 else hf.shape[1]
        assert np.array_equal(hm[r, :end, 0], hf[r, :end, 0]), r     # replayed through the quasi-static phase
    e8, e6 = W.grid_errors(hf, hm, -8.0), W.grid_errors(hf, hm, -6.0)
    print("prec %d on the FP64 grid: %s; %s" % (prec, e8, e6))
    _, b8, b6 = PATHS[prec]
    assert e8["dlnvmax"] <= b8 and e8["dlnv_probes"] <= b8 and e8["slip_rel"] <= b8
    assert e6["dlnvmax"] <= b6 and e6["dlnv_probes"] <= b6 and e6["slip_rel"] <= 5e-2


def test_gpu_c4a_replay_against_oracle(c4a, free_f64):
    """Realisations 0 and 63 for 120 steps of the FP64 run's grid (to t ~ 4e8
    s): the oracle replaying the same grid against the FP64 device replay
    (measured 2e-13 in ln V), the FP32-core one (1.1e-5 in ln Vmax, 4e-7 in
    slip; bound 5e-5) and the BF16-core one (bound 1e-4)."""
    g, Hm, frs, lod, ys, probes = c4a
    hf, _ = free_f64
    n = 120
    grid = hf[:, :n, 1]
    hr = _replay(c4a, H.PREC_F64, grid, n)
    hs = _replay(c4a, H.PREC_F32, grid, n)
    hm = _replay(c4a, H.PREC_MIXED, grid, n)
    Ho = O.HMatrix(g, leaf=64, eps=1e-5, nthreads=16)
    area = np.asarray(g.ds)[lod]
    for r in (0, 63):
        f = frs[r]
        so = O.Sim(Ho, f, lod, hdiv=W.C4A_DS0, nthreads=16)
        y = ys[r].copy()
        a = f.a[lod]
        vel = 2 * f.vref * np.exp(-y[0::3] / a) * np.sinh(y[1::3] / y[2::3] / a)
        _, oh, ov, osl = so.run_grid(y, 0.0, grid[r], 1e-6, np.zeros(g.n), vel, area, rec=probes)
        assert np.allclose(hr[r, :, 0], oh[:, 0], rtol=1e-14, atol=0)
        for name, h, tol in (("FP64", hr, 1e-9), ("FP32 cores", hs, 5e-5), ("BF16 cores", hm, 1e-4)):
            dv = np.abs(np.log(np.abs(h[r, :, 5::2]) / np.abs(ov))).max()
            dm = np.abs(h[r, :, 4] - oh[:, 4]).max() * np.log(10)
            ds = np.abs(h[r, :, 6::2] - osl).max() / np.abs(osl).max()
            print("realisation %d, %s vs oracle over %d steps (to t = %.4e s): |d ln Vmax| %.3e, |d ln V| probes "
                  "%.3e, slip %.3e" % (r, name, n, oh[-1, 0], dm, dv, ds))
            assert dm <= tol and dv <= tol and ds <= tol
