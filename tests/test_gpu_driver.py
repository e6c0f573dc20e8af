"""GPU: the Fortran driver (hbi_amd/lhbiem_gpu, the reference's main_LH
call sites over the C ABI) run end to end from .in and parameter files.

* forward mode (main_LH.f90:2586-2664): H*1 / ds0 in cluster order against
  the library's own product and sampled exact kernel rows
* cycle mode: monitor<n>.dat rows against a Python-driven run of the same
  parameters (identical initial state formula, main_LH.f90:1004)
"""
import os
import subprocess

import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import problems as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hbi_amd", "lhbiem_gpu")


def _run(tmp_path, keys):
    P.write_in_file(str(tmp_path / "run.in"), keys)
    r = subprocess.run([DRIVER, "run.in"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    return r.stdout


def test_gpu_driver_forward_mode(tmp_path):
    imax, jmax, ds0 = 48, 24, 1.0
    _run(tmp_path, {"problem": "3dph", "imax": imax, "jmax": jmax, "ds": ds0, "forward": True, "eps_h": 1e-6,
                    "leaf": 32, "filenumber": 7})
    out = np.loadtxt(tmp_path / "output" / "stress7")
    g = P.mesh_3dph(imax, jmax, ds0)
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    assert out.shape == (g.n, 6)
    # the output format (e16.4, main_LH.f90:2661) keeps 4 significant digits
    q = lambda v: 5.01e-4 * np.abs(v) + 1e-12
    # rows in cluster order: coordinates of element lod[i]
    for c, arr in ((0, g.xcol), (1, g.ycol), (2, g.zcol)):
        assert (np.abs(out[:, c] - arr[lod]) <= q(arr[lod]) + 1e-9).all()
    y = Hg.matvec(np.ones(g.n), H.PREC_F64) / ds0
    assert (np.abs(out[:, 4] - y) <= q(y)).all()
    # exact kernel rows (mesh order) for a few elements
    for i in (0, 17, g.n // 2, g.n - 1):
        k = H.entries(g, np.full(g.n, lod[i], dtype=np.int32), np.arange(g.n, dtype=np.int32))
        assert abs(out[i, 4] - k.sum() / ds0) <= q(k.sum() / ds0) + 1e-6 * np.abs(k).sum() / ds0


def test_gpu_driver_cycle_matches_library_run(tmp_path):
    n = 512
    ds0 = 40.0 / n
    g = P.mesh_2dvs(n, ds0)
    fr = P.friction_c1(g)
    # off steady state (an exact steady state has errmax <= tiny: rkqs underflows on it)
    fr.vel = fr.vel * np.where(g.xcol < 10.0, 3.0, 1.0)
    P.write_param_file(str(tmp_path / "param.dat"), fr)
    nstep = 24
    _run(tmp_path, {"problem": "2dvs", "ncellg": n, "ds": ds0, "nstep": nstep, "eps_r": 1e-6, "eps_h": 1e-6,
                    "leaf": 32, "filenumber": 3, "parameterfromfile": True, "parameter_file": "param.dat",
                    "parameter_file_ncol": 8, "outbatch": 10, "vpl": 1e-9, "tmax": 1e3})
    mon = np.loadtxt(tmp_path / "output" / "monitor3.dat")
    # k, t, log10 Vmax, mean slip, mean mu, max sigma, min sigma, errmax, dt, nrjct, wall (main_LH.f90:1491)
    assert mon.shape == (nstep + 1, 11)
    assert np.array_equal(mon[:, 0], np.arange(0, nstep + 1))
    assert mon[0, 1] == 0.0 and mon[0, 7] == 0.0 and mon[0, 8] == 0.0
    mon = mon[1:]
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    rk = H.RK(Hg, fr, lod, g.ds, eps_r=1e-6)
    rk.set_state(P.init_state(fr, lod), 0.0, 1.0)
    info, hist, _ = rk.step(nstep, history=True)
    # printed with 5 significant digits (e16.5) / 4 decimals (f19.4)
    assert np.allclose(mon[:, 1], hist[:, 0], rtol=1e-6, atol=1e-3)
    for col, hcol in ((2, 4), (3, 5), (4, 6), (5, 7), (6, 8), (7, 2), (8, 1)):
        assert np.allclose(mon[:, col], hist[:, hcol], rtol=6e-5, atol=1e-30), col
    assert np.array_equal(mon[:, 9].astype(int), hist[:, 3].astype(int))
    assert (np.diff(mon[:, 10]) >= 0).all()
    # the history's monitor quantities against the state
    y, slip, vel, _ = rk.get_state()
    area = np.asarray(g.ds)[lod]
    assert np.isclose(hist[-1, 5], np.sum(np.abs(slip) * area) / area.sum(), rtol=1e-12)
    assert np.isclose(hist[-1, 6], np.sum(y[1::3] / y[2::3] * area) / area.sum(), rtol=1e-12)
    assert hist[-1, 7] == y[2::3].max() and hist[-1, 8] == y[2::3].min()
