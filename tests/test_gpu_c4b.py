"""GPU: the distinct-faults workload -- 16 planar faults of 4 232 to 29 282
elements (sizes drawn from seed 4), each with its own H-matrix, friction
realisation and step control -- advanced together by hbi_rk_step_many (one
graph launch per lockstep round, a branch per fault).

* Mixed precision, all 16 faults at full size: 60 steps together give
  bit-identical states, times and step counts to 60 steps of each fault
  alone (the same kernels on the same data), and a second call continues
  every fault.
* FP64, all 16 faults together for 20 accepted steps: the two smallest
  faults (4 232 and 6 962 elements) against the oracle's single runs -- the
  same accept / reject counts, t to 1e-9, states to 1e-9 of each
  component's size, slip to 1e-8 (the device's and the oracle's ACA+
  factors differ at the 1e-10 level).  The mixed-precision runs of the same
  faults: t to 1e-3, states to 1e-5 (FP32 product rounding moves the
  embedded error estimate by ~1e-3 and h with it as errmax^-1/5)."""
import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import workloads as W
from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu

K = 20
SMALL = (7, 15)


@pytest.fixture(scope="module")
def faults():
    print("C4b setup: 16 faults, sizes from seed 4", flush=True)
    fs = W.c4b_setup()
    assert sorted(d[0].n for d in fs)[:2] == [4232, 6962] and max(d[0].n for d in fs) == 29282
    return fs


def test_gpu_c4b_together_equals_alone(faults):
    print("C4b mixed: alone, then together", flush=True)
    res = []
    for mode in ("alone", "together"):
        rs = W.c4b_runs(faults, warm=0)
        if mode == "alone":
            infos = [rk.step(60)[0] for rk in rs]
            infos = [rk.step(20)[0] for rk in rs]
        else:
            H.RK.step_many(rs, 60)
            infos = H.RK.step_many(rs, 20)
        res.append((infos, [rk.get_state() for rk in rs]))
        for rk in rs:
            rk.close()
    for f in range(len(faults)):
        ia, sa = res[0][0][f], res[0][1][f]
        im, sm = res[1][0][f], res[1][1][f]
        assert ia["status"] == im["status"] == 0
        assert im["accepted"] == ia["accepted"] == 80 and im["rejected"] == ia["rejected"], f
        assert im["t"] == ia["t"], f
        for k in range(3):
            assert np.array_equal(sm[k], sa[k]), (f, k)


@pytest.fixture(scope="module")
def oracle_small(faults):
    out = {}
    for f in SMALL:
        g, Hm, fr, lod, y0, ds0 = faults[f]
        print("C4b oracle: fault %d (%d elements), %d steps" % (f, g.n, K), flush=True)
        Ho = O.HMatrix(g, leaf=64, eps=1e-5, nthreads=16)
        assert np.array_equal(Ho.lod, lod)
        so = O.Sim(Ho, fr, lod, hdiv=ds0, nthreads=16)
        y = y0.copy()
        a = fr.a[lod]
        vel = 2 * fr.vref * np.exp(-y[0::3] / a) * np.sinh(y[1::3] / y[2::3] / a)
        slip = np.zeros(g.n)
        done, t, _, hist, _ = so.run(y, 0.0, 1.0, 1e-6, K, slip, vel, np.asarray(g.ds)[lod])
        assert done == K
        out[f] = (y, slip, t, int(hist[:, 3].sum()))
    return out


@pytest.mark.parametrize("prec,ytol,ttol", [(H.PREC_F64, 1e-9, 1e-9), (H.PREC_MIXED, 1e-5, 1e-3)])
def test_gpu_c4b_small_faults_against_oracle(faults, oracle_small, prec, ytol, ttol):
    print("C4b prec %d: 16 faults together, %d steps" % (prec, K), flush=True)
    rs = W.c4b_runs(faults, prec=prec, warm=0)
    infos = H.RK.step_many(rs, K)
    for f in SMALL:
        yo, so, to, rej = oracle_small[f]
        y, slip, _, _ = rs[f].get_state()
        dt = abs(infos[f]["t"] - to) / to
        worst = max(np.abs(y[c::3] - yo[c::3]).max() / np.abs(yo[c::3]).max() for c in range(3))
        ds = np.abs(slip - so).max() / np.abs(so).max()
        print("fault %d: rejected %d (oracle %d), t rel %.3e, state max rel %.3e, slip max rel %.3e"
              % (f, infos[f]["rejected"], rej, dt, worst, ds))
        assert infos[f]["accepted"] == K and infos[f]["status"] == 0
        if prec == H.PREC_F64:
            assert infos[f]["rejected"] == rej
        assert dt <= ttol and worst <= ytol and ds <= 10 * ytol
    for rk in rs:
        rk.close()
