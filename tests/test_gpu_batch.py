"""GPU: several independent runs (distinct faults, own matrices and step
control) advanced together by hbi_rk_step_many give exactly the results of
advancing each alone (same graphs, same kernels, so bit-identical)."""
import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import problems as P

pytestmark = pytest.mark.gpu


def _fault(imax, jmax, seed):
    ds0 = 50.0 / imax
    g = P.mesh_3dph(imax, jmax, ds0)
    Hm = H.HMatrix(g, leaf=32, eps=1e-5)
    fr = P.friction_c2(g, imax, jmax, ds0, seed=seed, patch=(30.0, 12.0))
    lod = Hm.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod, amp=0.02, depth=0.4 * jmax * ds0, width=3.0)
    return g, Hm, fr, lod, y0, ds0


def test_gpu_step_many_equals_individual_runs():
    faults = [_fault(im, jm, s) for im, jm, s in ((32, 16, 1), (48, 24, 2), (64, 32, 3), (40, 20, 4))]
    results = []
    for mode in ("alone", "many"):
        runs = []
        for g, Hm, fr, lod, y0, ds0 in faults:
            rk = H.RK(Hm, fr, lod, g.ds, hdiv=ds0, eps_r=1e-6, prec=H.PREC_MIXED)
            rk.set_state(y0, 0.0, 1.0)
            runs.append(rk)
        if mode == "alone":
            infos = [rk.step(150)[0] for rk in runs]
        else:
            infos = H.RK.step_many(runs, 150)
            # a second call continues every run
            infos = H.RK.step_many(runs, 50)
            for rk in runs:
                rk.step(0)
        if mode == "alone":
            infos = [rk.step(50)[0] for rk in runs]
        states = [rk.get_state() for rk in runs]
        results.append((infos, states))
    for f in range(len(faults)):
        ia, sa = results[0][0][f], results[0][1][f]
        im, sm = results[1][0][f], results[1][1][f]
        assert im["accepted"] == ia["accepted"] == 200 and im["rejected"] == ia["rejected"]
        assert im["t"] == ia["t"]
        for k in range(3):
            assert np.array_equal(sm[k], sa[k]), (f, k)
