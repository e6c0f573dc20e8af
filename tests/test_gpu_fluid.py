"""GPU: the fluid-driven cycle (hbi_fv_*, main_fv.f90) against the oracle
(oracle/fluid.py) on the same inputs:

* derivs (4 components) with and without the normal-stress matrix, aging and
  slip laws
* the implicit pore-pressure step for every setting and boundary pair, both
  solvers, at the four grid sizes of the solve kernel's variants
* the first steps of injection runs: identical step grids, CG iteration
  counts, monitor rows and states
* at 8192 elements (the largest grid) the run holds the injected volume
"""
import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import problems as P
from oracle import fluid as F
from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu

TOL = float(np.float32(1e-4))    # the solver's residual tolerance (main_fv.f90:1785)


def _pair(g, kw, normal=False, eps=1e-8):
    Hs = H.HMatrix(g, leaf=32, eps=eps)
    Os = O.HMatrix(g, leaf=32, eps=eps, nthreads=8)
    Hn = On = None
    if normal:
        import dataclasses
        gn = dataclasses.replace(g, vcomp=1)
        Hn, On = H.HMatrix(gn, leaf=32, eps=eps), O.HMatrix(gn, leaf=32, eps=eps, nthreads=8)
    opts, arr, y, pf = P.fluid_injection(g, Hs.lod, **kw)
    dev = H.FV(Hs, opts, arr, hn=Hn)
    ora = F.FvOracle(Os, opts, arr, hn=On)
    return dev, ora, y, pf, (Hs, Hn)


@pytest.mark.parametrize("law,normal", [(0, False), (1, False), (0, True)])
def test_gpu_fluid_derivs_against_oracle(law, normal):
    g = P.mesh_2dph(384, 0.01, 45.0) if normal else P.mesh_2dp(384, 0.01)
    dev, ora, y, pf, _ = _pair(g, dict(law=law), normal=normal)
    rng = np.random.default_rng(5)
    y[1::4] *= 1 + 0.02 * rng.standard_normal(g.n)
    y[3::4] *= 1 + rng.random(g.n)
    dg, do = dev.derivs(y), ora.derivs(y)
    print("derivs law", law, "normal", normal, flush=True)
    for c in range(4):
        scale = np.abs(do[c::4]).max()
        assert np.abs(dg[c::4] - do[c::4]).max() <= 1e-9 * scale + 1e-300, c
    if normal:
        assert np.abs(do[2::4]).max() > 0


DIFFUSE_CASES = [
    # (n, setting, bcl, bcr, pressure_dependent)
    (1000, "injectionq", "Neumann", "Neumann", False),
    (1000, "injectionp", "Dirichlet", "Neumann", True),
    (2000, "steadystate", "Neumann", "Neumann", False),
    (2000, "thrust", "Neumann", "Neumann", False),
    (4000, "other", "Dirichlet", "Dirichlet", True),
    (8192, "injectionq", "Neumann", "Dirichlet", False),
]


@pytest.mark.parametrize("n,setting,bcl,bcr,pdep", DIFFUSE_CASES)
def test_gpu_fluid_pressure_step_against_oracle(n, setting, bcl, bcr, pdep):
    g = P.mesh_2dp(n, 0.01)
    kw = dict(setting=setting, bcl=bcl, bcr=bcr, pressure_dependent=pdep, pbcl=4.0, pbcr=1.0, pinj=8.0)
    dev, ora, y, pf, _ = _pair(g, kw, eps=1e-4)
    rng = np.random.default_rng(n)
    if setting == "thrust":   # sources along the fault and at qin(2000)
        ora.arr["qin"][:] = 1e-7 * rng.standard_normal(n)
        dev = H.FV(dev.hs, ora.o, ora.arr)
    pf = pf + rng.random(n)
    y[3::4] = 1e-13 * (1 + 4 * rng.random(n))
    sigma = y[2::4] + pf
    for h in (2e3, 3e5):
        print("pressure step", n, setting, bcl, bcr, pdep, h, flush=True)
        pg, ig, dtg = dev.diffuse(pf, y, sigma, h, 100.0, 1e8)
        po, io, dto = ora.diffuse(pf, y, sigma, h, 100.0, 1e8)
        # the stopping iteration may move by one under other summation orders
        assert ig > 0 and abs(ig - io) <= max(1, io // 50), (ig, io)
        # both results solve the same system to the residual bound: they are
        # 2 tol sqrt(N) / lambda_min apart at most (lambda_min >= 1/2)
        assert np.linalg.norm(pg - po) <= 4 * TOL * np.sqrt(n)
        if ig == io and io < 60:   # few iterations: rounding has not grown
            assert np.abs(pg - po).max() <= 1e-9 * np.abs(po).max(), (ig, np.abs(pg - po).max())
        if not pdep:
            # the device's pfnew is a converged solution of the reference's system
            lo, di, up, b, _ = ora.beuler_system(pf, ora.cdiff(y[3::4], sigma, pf), h, 100.0)
            r = b - ora.apply(lo, di, up, pg - ora.arr["pfhyd"])
            assert F.seqsum(r * r) < 1.001 * TOL ** 2 * n
        assert abs(dtg - dto) <= 1e-5 * abs(dto)


@pytest.mark.parametrize("setting,pdep", [("injectionq", False), ("injectionp", True)])
def test_gpu_fluid_run_against_oracle(setting, pdep):
    n = 256
    g = P.mesh_2dp(n, 0.01)
    dev, ora, y, pf, _ = _pair(g, dict(setting=setting, pressure_dependent=pdep))
    pidx = [0, 100, n // 2 - 1, n - 1]
    dev.set_probes(pidx)
    dev.set_state(y, pf, t=0.0, dt=1.0)
    st = dict(y=y.copy(), pf=pf.copy(), disp=np.zeros(n), t=0.0, h=1.0)
    rows = [ora.step(st) for _ in range(25)]
    info, hist = dev.step(25, history=True)
    print("run", setting, info, flush=True)
    assert info["status"] == 0 and info["accepted"] == 25 and len(hist) == 25
    assert info["rejected"] == sum(r["rejected"] for r in rows)
    for k, r in enumerate(rows):
        hr = hist[k]
        assert abs(hr[8] - r["dt"]) <= 1e-9 * r["dt"], k
        assert abs(hr[0] - r["t"]) <= 1e-9 * r["t"], k
        assert int(hr[9]) == r["niter"] or abs(int(hr[9]) - r["niter"]) <= 1, k
        assert int(hr[6]) == r["argmax"], k
        assert abs(hr[1] - r["log10vmax"]) <= 1e-9 * abs(r["log10vmax"]), k
        assert abs(hr[3] - r["meanmu"]) <= 1e-12 * abs(r["meanmu"]), k
        assert abs(hr[5] - r["minnorm"]) <= 1e-9 * abs(r["minnorm"]), k
    s = dev.get_state()
    assert np.abs(s["pf"] - st["pf"]).max() <= 1e-6 * max(1.0, np.abs(st["pf"]).max())
    for c in range(4):
        ref = st["y"][c::4]
        assert np.abs(s["y"][c::4] - ref).max() <= 1e-9 * np.abs(ref).max(), c
    assert np.abs(s["disp"] - st["disp"]).max() <= 1e-8 * np.abs(st["disp"]).max()
    assert np.abs(s["vel"] - st["vel"]).max() <= 1e-8 * np.abs(st["vel"]).max()
    assert np.abs(s["qflow"] - st["qflow"]).max() <= 1e-6 * np.abs(st["qflow"]).max() + 1e-300
    assert s["info"]["niter"] == int(hist[-1, 9])
    # local output rows (main_fv.f90:1312-1316) of the last step = the state
    pr = dev.get_probes()
    assert pr.shape == (25, 4, 9)
    yl = s["y"]
    last = pr[-1]
    assert np.array_equal(last[:, 0], s["vel"][pidx]) and np.array_equal(last[:, 1], s["disp"][pidx])
    assert np.array_equal(last[:, 2], yl[2::4][pidx]) and np.array_equal(last[:, 3], yl[1::4][pidx])
    assert np.array_equal(last[:, 4], s["pf"][pidx]) and np.array_equal(last[:, 7], yl[0::4][pidx])
    assert np.array_equal(last[:, 6], yl[1::4][pidx] / yl[2::4][pidx])


def test_gpu_fluid_full_grid_holds_the_injected_volume():
    """8192 elements (the largest pore-pressure grid): over 60 steps the
    half-weighted pressure sum equals the injected volume to the solver's
    residual bound, and pf peaks at the source."""
    n = H.FV_MAX_CELLS
    g = P.mesh_2dp(n, 0.01)
    Hs = H.HMatrix(g, leaf=32, eps=1e-4)
    opts, arr, y, pf = P.fluid_injection(g, Hs.lod, setting="injectionq")
    dev = H.FV(Hs, opts, arr)
    dev.set_state(y, pf, t=0.0, dt=1.0)
    info, hist = dev.step(60, history=True)
    print("full grid", info, flush=True)
    assert info["status"] == 0 and len(hist) == 60
    s = dev.get_state()
    w = np.ones(n)
    w[0] = w[-1] = 0.5
    k = n // 2 - 1
    rate = arr["qin"][k] / opts["beta"] / arr["phi"][k] * float(np.float32(1e-9)) / opts["ds0"]
    vol = rate * hist[:, 8].sum()
    assert vol > 0
    assert abs(np.dot(w, s["pf"]) - vol) <= 60 * TOL * n
    assert int(np.argmax(s["pf"])) == k and hist[:, 9].max() > 0


def test_gpu_fluid_rejects_unsupported_grids():
    g = P.mesh_2dp(H.FV_MAX_CELLS + 8, 0.01)
    Hs = H.HMatrix(g, leaf=32, eps=1e-4)
    opts, arr, y, pf = P.fluid_injection(g, Hs.lod)
    with pytest.raises(H.HbiError):
        H.FV(Hs, opts, arr)
    g = P.mesh_2dp(512, 0.01)
    Hs = H.HMatrix(g, leaf=32, eps=1e-4)
    opts, arr, y, pf = P.fluid_injection(g, Hs.lod, setting="thrust")
    with pytest.raises(H.HbiError):
        H.FV(Hs, opts, arr)


def test_gpu_fluid_driver_matches_the_library_run(tmp_path):
    """lhbiem_fv_gpu (the main_fv host) on bp6a at 1000 elements: its monitor
    rows (main_fv.f90:1307-1311) are the library run's history rows, the
    streams hold the initial and every interval-th state, and the last pf
    record is the library's final pf."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    n, nstep = 1000, 300
    keys = {"problem": "2dp", "ncellg": n, "filenumber": 3, "interval": 100, "nstep": nstep, "tmax": 2.0,
            "tinj": 0.27397, "ds": 0.01, "kpmax": 1e-13, "kpmin": 1e-13, "eta": 1e-3, "beta": 1e-8, "phi0": 0.1,
            "q0": 1.25e-6, "a": 0.007, "b": 0.005, "dc": 0.004, "f0": 0.6, "tauinit": 29.2, "sigmainit": 50.0,
            "velinit": "1d-12"}
    P.write_in_file(str(tmp_path / "bp6.in"), keys)
    r = subprocess.run([os.path.join(root, "hbi_amd", "lhbiem_fv_gpu"), "bp6.in"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "output"
    mon = np.loadtxt(out / "monitor3.dat")
    assert mon.shape == (nstep + 1, 12) and np.array_equal(mon[:, 0], np.arange(nstep + 1))
    g = P.mesh_2dp(n, 0.01)
    Hs = H.HMatrix(g, leaf=32, eps=1e-4)
    opts, arr, y, pf = P.fluid_bp6a(g, Hs.lod)
    fv = H.FV(Hs, opts, arr)
    fv.set_state(y, pf, t=0.0, dt=1.0)
    info, hist = fv.step(nstep, history=True)
    assert len(hist) == nstep
    # the two processes may tune the product to different work-item shapes
    # (another summation order), so their step grids agree until a rounding
    # difference flips a step decision: rows are compared over the first 100
    # steps, the end state to the run's own variability
    k = 100
    assert np.allclose(mon[1:k + 1, 1], hist[:k, 0], rtol=1e-9, atol=6e-5)          # f19.4
    for c_mon, c_hist in ((2, 1), (3, 2), (4, 3), (5, 4), (6, 5), (8, 7), (9, 8)):   # e16.5
        assert np.allclose(mon[1:k + 1, c_mon], hist[:k, c_hist], rtol=6e-5, atol=1e-300), c_mon
    assert np.array_equal(mon[1:k + 1, 11], hist[:k, 9])
    assert abs(mon[-1, 1] - hist[-1, 0]) <= 1e-3 * hist[-1, 0]
    # local files (main_fv.f90:862-868, 1312-1316): five points, a row per
    # step; at the output steps their values are the field streams' (e16.6)
    vel = np.fromfile(out / "vel3.dat").reshape(4, n)
    slp = np.fromfile(out / "slip3.dat").reshape(4, n)
    sgm = np.fromfile(out / "sigma3.dat").reshape(4, n)
    tau = np.fromfile(out / "tau3.dat").reshape(4, n)
    pff = np.fromfile(out / "pf3.dat").reshape(4, n)
    for i in range(1, 6):
        lid = int((i - 0.5) / 5 * n)
        loc = np.loadtxt(out / ("local3-%d.dat" % lid))
        assert loc.shape == (nstep, 11) and np.array_equal(loc[:, 0], np.arange(1, nstep + 1))
        for rec, k in ((1, 100), (2, 200), (3, 300)):
            r = loc[k - 1]
            for c, ref in ((2, np.log10(vel[rec, lid - 1])), (3, slp[rec, lid - 1]), (4, sgm[rec, lid - 1]),
                           (5, tau[rec, lid - 1]), (6, pff[rec, lid - 1])):
                assert abs(r[c] - ref) <= 6e-6 * abs(ref) + 1e-300, (lid, k, c)
    tr = np.loadtxt(out / "time3.dat", ndmin=2)
    assert list(tr[:, 0].astype(int)) == [0, 100, 200, 300]
    for name in ("vel", "slip", "sigma", "tau", "ks", "pf", "qflow"):
        d = np.fromfile(out / ("%s3.dat" % name), dtype=np.float64)
        assert d.size == 4 * n, name
    pfs = np.fromfile(out / "pf3.dat", dtype=np.float64).reshape(4, n)
    s = fv.get_state()
    assert np.abs(pfs[-1] - s["pf"]).max() <= 1e-3 * np.abs(s["pf"]).max()
    assert pfs[-1].max() > 0 and np.array_equal(pfs[0], np.zeros(n))


def test_gpu_fluid_constant_pressure_and_output_step_bound():
    """pfconst (no pressure step: pf stays, sigma_e = (sigma_e + pf) - pf) and
    outpertime (hnext <= dtout, main_fv.f90:2233-2235) against the oracle."""
    n = 256
    g = P.mesh_2dp(n, 0.01)
    dtout_yr = 2e4 / (365 * 24 * 3600)
    dev, ora, y, pf, _ = _pair(g, dict(setting="injectionq", pfconst=True, outpertime=True, dtout=dtout_yr))
    pf = pf + 0.5
    dev.set_state(y, pf, t=0.0, dt=1.0)
    st = dict(y=y.copy(), pf=pf.copy(), disp=np.zeros(n), t=0.0, h=1.0)
    rows = [ora.step(st) for _ in range(30)]
    info, hist = dev.step(30, history=True)
    assert info["status"] == 0 and len(hist) == 30
    assert np.all(hist[:, 8] <= ora.o["dtout_s"] * (1 + 1e-12)) and hist[-1, 8] == pytest.approx(ora.o["dtout_s"])
    for k, r in enumerate(rows):
        assert abs(hist[k, 8] - r["dt"]) <= 1e-9 * r["dt"], k
        assert hist[k, 9] == 0
    s = dev.get_state()
    assert np.array_equal(s["pf"], pf)
    for c in range(4):
        ref = st["y"][c::4]
        assert np.abs(s["y"][c::4] - ref).max() <= 1e-9 * np.abs(ref).max(), c


@pytest.mark.parametrize("kw,status", [(dict(velmax=1e-9), 1), (dict(velmin=1e-8), 1), (dict(tmax=1e-7), 3)])
def test_gpu_fluid_stop_controls(kw, status):
    """The stop controls of main_fv.f90:1136-1152 (V outside [velmin,
    velmax], t > tmax) stop the device run at the same step as the
    oracle's, with the step counted."""
    n = 128
    g = P.mesh_2dp(n, 0.01)
    base = dict(setting="injectionq", velinit=1e-9)
    base.update(kw)
    dev, ora, y, pf, _ = _pair(g, base)
    dev.set_state(y, pf, t=0.0, dt=1.0)
    st = dict(y=y.copy(), pf=pf.copy(), disp=np.zeros(n), t=0.0, h=1.0)
    k_ora = None
    for k in range(200):
        r = ora.step(st)
        if r["status"]:
            k_ora = k + 1
            assert r["status"] == status
            break
    assert k_ora is not None
    info, hist = dev.step(200, history=True)
    assert info["status"] == status and info["accepted"] == k_ora and len(hist) == k_ora


def test_gpu_fluid_driver_steadystate_setting(tmp_path):
    """lhbiem_fv_gpu with setting steadystate (initcond_ss, main_fv.f90:1452-1498):
    the initial pf record is the linear profile from pbcl = q0 lf eta 1e-3 / kp0
    down to 0, kTv is raised outside the central third (through kpmaxv), and the
    run advances with the steadystate flux terms of Beuler."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    n, nstep, ds0 = 512, 40, 0.01
    keys = {"problem": "2dp", "ncellg": n, "filenumber": 4, "interval": 20, "nstep": nstep, "ds": ds0,
            "setting": "steadystate", "kp0": 1e-13, "kpmin": 1e-14, "s0": 30.0, "vpl": 1e-9, "q0": 1e-7,
            "eta": 1e-3, "beta": 1e-8, "phi0": 0.1, "sigmainit": 50.0, "velinit": "1d-9", "a": 0.01, "b": 0.005,
            "dc": 0.01, "f0": 0.6}
    P.write_in_file(str(tmp_path / "ss.in"), keys)
    r = subprocess.run([os.path.join(root, "hbi_amd", "lhbiem_fv_gpu"), "ss.in"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "output"
    mon = np.loadtxt(out / "monitor4.dat")
    assert mon.shape[0] == nstep + 1 and np.isfinite(mon).all()
    pfs = np.fromfile(out / "pf4.dat").reshape(-1, n)
    assert pfs.shape[0] == 3
    lf = n * ds0
    pbcl = 1e-7 * lf / 1e-13 * 1e-3 * 1e-3
    xcol = (np.arange(n) + 0.5) * ds0
    assert np.allclose(pfs[0], pbcl + (0.0 - pbcl) * xcol / lf, rtol=1e-12, atol=1e-12 * pbcl)
    assert np.abs(pfs[-1] - pfs[0]).max() > 0


def test_gpu_fluid_break_time_ends_the_call_not_the_run():
    """hbi_fv_set_break: the call ends (status 0) at the first accepted step
    past the break time, and continuing from there is the unbroken run, row
    for row."""
    n = 256
    g = P.mesh_2dp(n, 0.01)
    Hs = H.HMatrix(g, leaf=32, eps=1e-4)
    opts, arr, y, pf = P.fluid_bp6a(g, Hs.lod)
    a = H.FV(Hs, opts, arr)
    a.set_state(y, pf, t=0.0, dt=1.0)
    _, ha = a.step(60, history=True)
    b = H.FV(Hs, opts, arr)
    b.set_state(y, pf, t=0.0, dt=1.0)
    tb = 0.5 * (ha[19, 0] + ha[20, 0])
    b.set_break(tb)
    info, hb = b.step(60, history=True)
    print("break", tb, info["accepted"], flush=True)
    assert info["status"] == 0 and info["accepted"] == 21 and len(hb) == 21
    assert hb[-2, 0] <= tb < hb[-1, 0]
    b.set_break()
    info, hb2 = b.step(39, history=True)
    assert info["accepted"] == 60
    assert np.array_equal(np.concatenate([hb, hb2]), ha)


def test_gpu_fluid_driver_stop_step_and_time_based_outputs(tmp_path):
    """lhbiem_fv_gpu: a stop control (t > tmax) ends the run before the
    stopping step's outputs, as main_fv.f90:1136-1152 precede 1156-1239 (no
    monitor, local or field row for it); field records by time (dtout) fall on
    the first step past each tout (main_fv.f90:1160-1163), not at the end of
    an output batch.  Steps and times from the library run of the same bp6a
    set-up."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    n, nstep = 1000, 300
    g = P.mesh_2dp(n, 0.01)
    Hs = H.HMatrix(g, leaf=32, eps=1e-4)
    opts, arr, y, pf = P.fluid_bp6a(g, Hs.lod)
    fv = H.FV(Hs, opts, arr)
    fv.set_state(y, pf, t=0.0, dt=1.0)
    _, hist = fv.step(80, history=True)
    t = hist[:, 0]
    tmax_s = 0.5 * (t[59] + t[60])              # the run stops at step 61
    dtout_s = 0.5 * (t[14] + t[15])             # first field record at step 16
    expect, tout = [0], dtout_s
    for k in range(1, 61):                      # step 61 stops before its outputs
        if t[k - 1] > tout:
            expect.append(k)
            tout = t[k - 1] + dtout_s
    assert len(expect) >= 3
    keys = {"problem": "2dp", "ncellg": n, "filenumber": 5, "interval": 100000, "nstep": nstep,
            "tmax": tmax_s / P.YEAR_S, "dtout": dtout_s / P.YEAR_S, "outbatch": 64,
            "tinj": 0.27397, "ds": 0.01, "kpmax": 1e-13, "kpmin": 1e-13, "eta": 1e-3, "beta": 1e-8, "phi0": 0.1,
            "q0": 1.25e-6, "a": 0.007, "b": 0.005, "dc": 0.004, "f0": 0.6, "tauinit": 29.2, "sigmainit": 50.0,
            "velinit": "1d-12"}
    P.write_in_file(str(tmp_path / "bp6.in"), keys)
    r = subprocess.run([os.path.join(root, "hbi_amd", "lhbiem_fv_gpu"), "bp6.in"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0 and "tmax" in r.stdout
    out = tmp_path / "output"
    mon = np.loadtxt(out / "monitor5.dat")
    assert np.array_equal(mon[:, 0], np.arange(61))
    loc = np.loadtxt(out / ("local5-%d.dat" % int(0.5 / 5 * n)))
    assert np.array_equal(loc[:, 0], np.arange(1, 61))
    tr = np.loadtxt(out / "time5.dat", ndmin=2)
    assert list(tr[:, 0].astype(int)) == expect
    assert np.fromfile(out / "pf5.dat").size == len(expect) * n


def test_gpu_fluid_driver_thrust_setting(tmp_path):
    """lhbiem_fv_gpu with setting thrust on a 2dph fault dipping 30 degrees
    (initcond_thrust, main_fv.f90:1378-1450; the normal-stress matrix as
    sigmaconst is F): the initial streams hold problems.fluid_thrust's state
    (sigma_e, pf, tau, ks in cluster order), and the monitor rows are those of
    the library run from that state with both matrices."""
    import dataclasses
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    n, nstep, dip = 2000, 100, 30.0
    kw = dict(a0=0.01, b0=0.015, dc=0.02, f0=0.6, vpl=1e-9, velinit=1e-9, phi0=0.1, beta=1e-8, eta=1e-3,
              kpmin=1e-16, kpmax=1e-13, kL=1e-3, s0=50.0, q0=1e-12)
    keys = {"problem": "2dph", "ncellg": n, "ds": 0.01, "dipangle": dip, "setting": "thrust", "filenumber": 5,
            "interval": 50, "nstep": nstep, "dtout": 1e3, "backslip": True, "a": kw["a0"], "b": kw["b0"],
            "dc": kw["dc"], "f0": kw["f0"], "vpl": kw["vpl"], "velinit": kw["velinit"], "phi0": kw["phi0"],
            "beta": kw["beta"], "eta": kw["eta"], "kpmin": kw["kpmin"], "kpmax": kw["kpmax"], "kL": kw["kL"],
            "s0": kw["s0"], "q0": kw["q0"]}
    P.write_in_file(str(tmp_path / "thrust.in"), keys)
    r = subprocess.run([os.path.join(root, "hbi_amd", "lhbiem_fv_gpu"), "thrust.in"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "output"
    g = P.mesh_2dph(n, 0.01, dip)
    Hs = H.HMatrix(g, leaf=32, eps=1e-4)
    Hn = H.HMatrix(dataclasses.replace(g, vcomp=1), leaf=32, eps=1e-4)
    opts, arr, y, pf = P.fluid_thrust(g, Hs.lod, dip, ds0=0.01, **kw)
    rec = {name: np.fromfile(out / ("%s5.dat" % name)).reshape(-1, n) for name in ("sigma", "pf", "tau", "ks")}
    for name, ref in (("sigma", y[2::4]), ("pf", pf), ("tau", y[1::4]), ("ks", y[3::4])):
        assert np.abs(rec[name][0] - ref).max() <= 1e-12 * np.abs(ref).max(), name
    assert rec["sigma"][0].min() > 3.0 and rec["pf"][0].max() > 90.0
    mon = np.loadtxt(out / "monitor5.dat")
    fv = H.FV(Hs, opts, arr, hn=Hn)
    fv.set_state(y, pf, t=0.0, dt=1.0)
    info, hist = fv.step(nstep, history=True)
    print("library", info, flush=True)
    k = min(len(hist), len(mon) - 1)
    assert k >= 50
    assert np.allclose(mon[1:k + 1, 1], hist[:k, 0], rtol=1e-9, atol=6e-5)
    for c_mon, c_hist in ((2, 1), (3, 2), (4, 3), (5, 4), (6, 5), (8, 7), (9, 8)):
        assert np.allclose(mon[1:k + 1, c_mon], hist[:k, c_hist], rtol=6e-5, atol=1e-300), c_mon
    assert np.array_equal(mon[1:k + 1, 11], hist[:k, 9])
