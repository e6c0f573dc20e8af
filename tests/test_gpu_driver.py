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


def _c1_nucleating(n=2048):
    """C1 with the seeded nucleation perturbation of the library tests,
    written through the parameter file's tau / sigma / vel columns (the
    driver derives psi from them, main_LH.f90:1004)."""
    ds0 = 40.0 / n
    g = P.mesh_2dvs(n, ds0)
    fr = P.friction_c1(g)
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hg.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod)
    a = fr.a[lod]
    v = 2 * fr.vref * np.exp(-y0[0::3] / a) * np.sinh(y0[1::3] / y0[2::3] / a)
    fr.vel = np.empty(n); fr.vel[lod] = v
    fr.tau = np.empty(n); fr.tau[lod] = y0[1::3]
    fr.sigma = np.empty(n); fr.sigma[lod] = y0[2::3]
    return g, fr, lod, ds0


def _keys(n, ds0, nstep, **kw):
    k = {"problem": "2dvs", "ncellg": n, "ds": ds0, "nstep": nstep, "eps_r": 1e-6, "eps_h": 1e-6, "leaf": 32,
         "filenumber": 5, "parameterfromfile": True, "parameter_file": "param.dat", "parameter_file_ncol": 8,
         "outbatch": 400, "vpl": 1e-9, "tmax": 1e4, "interval": 500, "dtout_co": 1.0, "dtout": 1.0}
    k.update(kw)
    return k


def _stream(path, n):
    d = np.fromfile(path, dtype=np.float64)
    assert d.size % n == 0
    return d.reshape(-1, n)


def test_gpu_driver_events_and_snapshots(tmp_path):
    """output/time, vel, slip, sigma, tau, event, EQslip, xyz, ind of the
    driver (main_LH.f90:1339-1425, 1518-1697): one snapshot per time row; an
    event row per finished event whose onset / end snapshots bracket the
    threshold, whose hypocentre is the first max |V| of the onset snapshot,
    whose coseismic slip is the difference of the two slip snapshots and whose
    Mw follows from their area-weighted mean slips."""
    n = 2048
    g, fr, lod, ds0 = _c1_nucleating(n)
    P.write_param_file(str(tmp_path / "param.dat"), fr)
    # outloc (main_LH.f90:322-324): 1-based mesh ids; one that is no cell gets no file
    loc_ids = [int(lod[7]) + 1, n // 2, n + 5]
    _run(tmp_path, _keys(n, ds0, 2500, outloc=loc_ids))
    out = tmp_path / "output"
    trows = np.loadtxt(out / "time5.dat", ndmin=2)
    vel, slip = _stream(out / "vel5.dat", n), _stream(out / "slip5.dat", n)
    sig, tau = _stream(out / "sigma5.dat", n), _stream(out / "tau5.dat", n)
    assert len(trows) == len(vel) == len(slip) == len(sig) == len(tau) > 0
    ks = trows[:, 0].astype(int)
    assert (np.diff(ks) > 0).all()
    assert all(k in ks for k in range(500, 2501, 500))          # every interval step
    ev = np.loadtxt(out / "event5.dat", ndmin=2)
    assert len(ev) >= 1
    eq = _stream(out / "EQslip5.dat", n)
    assert len(eq) == len(ev)
    velth = float(np.float32(1e-2))
    area = np.full(n, ds0)
    for e, row in enumerate(ev):
        count, k_end, onset, mw, hyp = int(row[0]), int(row[1]), row[2], row[3], int(row[4])
        assert count == e + 1
        j_end = int(np.where(ks == k_end)[0][0])
        # the onset snapshot: taken at the onset time (f19.4 vs f17.2 rows;
        # snapshots during an event are dtout_co = 1 s apart)
        j_on = int(np.argmin(np.abs(trows[:j_end, 1] - onset)))
        assert abs(trows[j_on, 1] - onset) <= 0.006 + 1e-9 * abs(onset)
        assert np.abs(vel[j_on]).max() > velth and np.abs(vel[j_on - 1]).max() <= velth
        assert np.abs(vel[j_end]).max() < 0.5 * velth
        assert hyp == int(np.argmax(np.abs(vel[j_on]))) + 1
        assert np.array_equal(eq[e], np.abs(slip[j_end] - slip[j_on]))
        m = (np.abs(slip[j_end]) * area).sum() / area.sum() - (np.abs(slip[j_on]) * area).sum() / area.sum()
        assert abs(mw - (np.log10(m * 32.04 * area.sum()) + 5.9) / 1.5) <= 6e-5
    # the end snapshot is the state the monitor row of that step describes
    mon = np.loadtxt(out / "monitor5.dat")
    for j, k in enumerate(ks[:5]):
        assert abs(mon[k, 6] - sig[j].min()) <= 6e-5 * abs(sig[j].min())
    xyz = np.loadtxt(out / "xyz5.dat")
    ind = np.loadtxt(out / "ind5.dat", dtype=int)
    assert np.array_equal(ind, lod + 1)
    # local<n>-<id>.dat (output_local, main_LH.f90:1495-1499): a row per step,
    # k, t, log10 V, slip, tau, sigma, mu, psi, vflow, veln, slipn; at the
    # snapshot steps the element's snapshot values (e16.5: 5 significant digits)
    assert not (out / ("local5-%d.dat" % (n + 5))).exists()
    for mid in loc_ids[:2]:
        loc = np.loadtxt(out / ("local5-%d.dat" % mid))
        i = int(np.where(lod == mid - 1)[0][0])
        assert np.array_equal(loc[:, 0], np.arange(1, 2501))
        assert np.array_equal(loc[:, 1], mon[1:2501, 1])
        assert (loc[:, 8] == 0).all() and (loc[:, 9] == 0).all() and (loc[:, 10] == 0).all()
        for j, k in enumerate(ks):
            r = loc[k - 1]
            assert abs(r[2] - np.log10(vel[j, i])) <= 5.01e-5 * abs(r[2]) + 1e-12
            for c, f in ((3, slip[j, i]), (4, tau[j, i]), (5, sig[j, i]), (6, tau[j, i] / sig[j, i])):
                assert abs(r[c] - f) <= 5.01e-5 * abs(f) + 1e-300, (k, c)
        # psi from V, tau, sigma (main_LH.f90:1004 inverted), on the rows of
        # the snapshot steps
        a = fr.a[mid - 1]
        for j, k in enumerate(ks[:20]):
            r = loc[k - 1]
            psi = a * np.log(2 * fr.vref / vel[j, i] * np.sinh(tau[j, i] / sig[j, i] / a))
            assert abs(r[7] - psi) <= 1e-4 * abs(psi) + 1e-6, (k, r[7], psi)
    assert np.allclose(xyz[:, 0], g.xcol[lod], rtol=5.01e-6, atol=1e-12)     # e15.6: 6 significant digits


def test_gpu_driver_restart_continues_the_run(tmp_path):
    """restart T (main_LH.f90:836-957): the run continues from the snapshot
    of the last interval step in its monitor file; the monitor then holds
    every step once, and the restarted state equals that snapshot."""
    n = 1024
    g, fr, lod, ds0 = _c1_nucleating(n)
    P.write_param_file(str(tmp_path / "param.dat"), fr)
    _run(tmp_path, _keys(n, ds0, 1300, interval=400))
    out = tmp_path / "output"
    vel1 = _stream(out / "vel5.dat", n)
    k1 = np.loadtxt(out / "time5.dat", ndmin=2)[:, 0].astype(int)
    snap1200 = vel1[int(np.where(k1 == 1200)[0][0])].copy()
    _run(tmp_path, _keys(n, ds0, 2000, interval=400, restart=True))
    mon = np.loadtxt(out / "monitor5.dat")
    assert np.array_equal(mon[:, 0], np.arange(0, 2001))
    assert (np.diff(mon[:, 1]) >= 0).all()
    k2 = np.loadtxt(out / "time5.dat", ndmin=2)[:, 0].astype(int)
    assert (np.diff(k2) > 0).all() and 1600 in k2 and 2000 in k2
    vel2 = _stream(out / "vel5.dat", n)
    assert len(vel2) == len(k2)
    # the first steps of the restarted file are the old file's, unchanged
    assert np.array_equal(vel2[int(np.where(k2 == 1200)[0][0])], snap1200)


def test_gpu_driver_normal_stress_and_opening(tmp_path):
    """sigmaconst F / opening T (main_LH.f90:683-716) on a rough triangular
    fault read from STL: the driver builds the normal-stress matrix and the two
    tensile-dislocation matrices; its monitor rows match a library run with the
    same four matrices, and veln / slipn<n>.dat hold each snapshot's opening
    rate and opening (post-step update, main_LH.f90:1279-1283)."""
    import dataclasses
    minsig, maxsig, nstep = 1.0, 300.0, 20
    P.write_stl(str(tmp_path / "fault.stl"),
                P.rough_surface_triangles(6.0, 3.0, 0.25, hurst=0.8, alpha=1e-2, seed=1))
    gs = P.mesh_triangles(P.read_stl(str(tmp_path / "fault.stl")))
    n = gs.n
    fr = P.friction_c3(gs)
    kw = dict(leaf=32, eps=1e-6)
    Hs = H.HMatrix(gs, **kw)
    lod = Hs.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, gs, lod, amp=0.02, depth=2.0, width=1.0)
    # a shallow patch below minsig (tau / sigma kept): it opens
    low = np.asarray(gs.zcol)[lod] > -1.0
    scale = np.where(low, 0.6 / y0[2::3], 1.0)
    y0[1::3] *= scale
    y0[2::3] *= scale
    a = fr.a[lod]
    for name, val in (("vel", 2 * fr.vref * np.exp(-y0[0::3] / a) * np.sinh(y0[1::3] / y0[2::3] / a)),
                      ("tau", y0[1::3]), ("sigma", y0[2::3])):
        arr = np.empty(n)
        arr[lod] = val
        setattr(fr, name, arr)
    P.write_param_file(str(tmp_path / "param.dat"), fr)
    _run(tmp_path, {"problem": "3dnt", "geometry_file": "fault.stl", "nstep": nstep, "eps_r": 1e-6,
                    "eps_h": 1e-6, "leaf": 32, "filenumber": 4, "parameterfromfile": True,
                    "parameter_file": "param.dat", "parameter_file_ncol": 8, "outbatch": 8, "evlaw": "slip",
                    "vref": 1e-6, "tmax": 1e4, "interval": 1, "velth": 1e3, "sigmaconst": False,
                    "opening": True, "minsig": minsig, "maxsig": maxsig})
    out = tmp_path / "output"
    mon = np.loadtxt(out / "monitor4.dat")[1:]
    Hn, Hos, Hon = (H.HMatrix(dataclasses.replace(gs, **r), **kw)
                    for r in ({"vcomp": 1}, {"md": 3}, {"md": 3, "vcomp": 1}))
    rk = H.RK(Hs, fr, lod, gs.ds, eps_r=1e-6, hn=Hn, minsig=minsig, maxsig=maxsig)
    rk.set_opening(Hos, Hon)
    rk.set_state(P.init_state(fr, lod), 0.0, 1.0)
    info, hist, _ = rk.step(nstep, history=True)
    assert info["accepted"] == nstep and len(mon) == nstep
    assert np.allclose(mon[:, 1], hist[:, 0], rtol=1e-6, atol=1e-3)
    # log10 Vmax, mean slip, mean mu, max / min sigma (sigma evolves: hn is on)
    for col, hcol in ((2, 4), (3, 5), (4, 6), (5, 7), (6, 8)):
        assert np.allclose(mon[:, col], hist[:, hcol], rtol=6e-5, atol=1e-30), col
    ks = np.loadtxt(out / "time4.dat", ndmin=2)[:, 0].astype(int)
    assert list(ks) == list(range(1, nstep + 1))
    vel, sig = _stream(out / "vel4.dat", n), _stream(out / "sigma4.dat", n)
    veln, slipn = _stream(out / "veln4.dat", n), _stream(out / "slipn4.dat", n)
    assert len(veln) == len(slipn) == len(vel) == nstep
    assert np.abs(sig[-1] - y0[2::3]).max() > 1e-9 * np.abs(y0[2::3]).max()
    for j in range(nstep):
        vn = vel[j] * (np.maximum(0.0, minsig - sig[j]) + np.minimum(0.0, maxsig - sig[j]))
        assert np.array_equal(veln[j], vn)
    assert np.abs(slipn[-1][low]).min() > 0 and np.abs(slipn[-1][~low]).max() == 0
    vn, sn = rk.get_opening()
    assert np.allclose(veln[-1], vn, rtol=1e-6, atol=1e-30) and np.allclose(slipn[-1], sn, rtol=1e-6, atol=1e-30)


def test_gpu_driver_forward_halfspace_triangles(tmp_path):
    """problem 3dht from an STL file (main_LH.f90:353-368, 479-510, 852-898):
    H*1 of the forward mode against the library's product on the same mesh
    (triangles are not divided by ds0, main_LH.f90:2645)."""
    P.write_stl(str(tmp_path / "thrust.stl"), P.plane_triangles(8.0, 4.0, 0.5, dip=30.0))
    _run(tmp_path, {"problem": "3dht", "geometry_file": "thrust.stl", "forward": True, "eps_h": 1e-6, "leaf": 32,
                    "filenumber": 8, "crake": 90.0})
    out = np.loadtxt(tmp_path / "output" / "stress8")
    g = P.mesh_triangles(P.read_stl(str(tmp_path / "thrust.stl")), crake=90.0, halfspace=True)
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    y = Hg.matvec(np.ones(g.n), H.PREC_F64)
    q = lambda v: 5.01e-4 * np.abs(v) + 1e-12
    assert out.shape == (g.n, 6)
    assert (np.abs(out[:, 4] - y) <= q(y)).all()
    # the free surface matters: the full-space kernel gives another H*1
    yf = H.HMatrix(P.mesh_triangles(P.read_stl(str(tmp_path / "thrust.stl")), crake=90.0), leaf=32,
                   eps=1e-6).matvec(np.ones(g.n), H.PREC_F64)
    assert np.abs(yf - y).max() > 1e-2 * np.abs(y).max()


def _curve(n, length, amp, y0):
    x = np.linspace(0.0, length, n + 1)
    y = y0 + amp * np.sin(6 * np.pi * x / length) + 0.3 * amp * np.sin(22 * np.pi * x / length)
    return np.stack([x[:-1], x[1:], y[:-1], y[1:]], axis=1)


def _forward_cases():
    seg = _curve(400, 20.0, 0.4, 0.0)
    segh = _curve(400, 20.0, 0.4, 1.5)
    rect = P.mesh_3dph(24, 12, 1.0, 60.0)
    rows = np.stack([rect.xcol, rect.ycol, rect.zcol, rect.ang, rect.angd, rect.dsl, rect.dsd], axis=1)
    return [
        # problem, keys, files, library geometry, divided by ds0
        ("2dp", {"ncellg": 500, "ds": 0.1}, {}, lambda: P.mesh_2dp(500, 0.1), None),
        ("2dph", {"ncellg": 800, "ds": 0.025, "dipangle": 30.0}, {}, lambda: P.mesh_2dph(800, 0.025, 30.0), None),
        ("3dp", {"imax": 32, "jmax": 16, "ds": 0.5}, {}, lambda: P.mesh_3dp(32, 16, 0.5), None),
        ("2dn", {"ncellg": 400, "geometry_file": "seg.dat"}, {"seg.dat": seg},
         lambda: P.mesh_segments(*seg.T, kind=P.KIND_2DN), None),
        ("2dnh", {"ncellg": 400, "geometry_file": "seg.dat"}, {"seg.dat": segh},
         lambda: P.mesh_segments(*segh.T, kind=P.KIND_2DPH), None),
        ("25d", {"ncellg": 400, "geometry_file": "seg.dat", "fwid": 5.0}, {"seg.dat": seg},
         lambda: P.mesh_segments(*seg.T, kind=P.KIND_25D, w=5.0), None),
        ("3dhr", {"ncellg": rect.n, "geometry_file": "rect.dat"}, {"rect.dat": rows},
         lambda: P.mesh_rectangles(rows, halfspace=True), 1.0),
        ("3dnr", {"ncellg": rect.n, "geometry_file": "rect.dat"}, {"rect.dat": rows},
         lambda: P.mesh_rectangles(rows, halfspace=False), 1.0),
    ]


@pytest.mark.parametrize("case", _forward_cases(), ids=lambda c: c[0])
def test_gpu_driver_forward_other_problems(tmp_path, case):
    """The other problems of the reference's mesh switch (main_LH.f90:432-471,
    595-612): the driver's H*1 against the library's product on the geometry
    built the same way (the 2-D segments, the 3dnr / 3dhr rectangles from
    their geometry files; only the rectangles are divided by ds0)."""
    problem, keys, files, make, div = case
    for name, data in files.items():
        np.savetxt(tmp_path / name, data, fmt="%.17e")
    _run(tmp_path, dict(problem=problem, forward=True, eps_h=1e-6, leaf=32, filenumber=9, **keys))
    out = np.loadtxt(tmp_path / "output" / "stress9")
    g = make()
    Hg = H.HMatrix(g, leaf=32, eps=1e-6)
    y = Hg.matvec(np.ones(g.n), H.PREC_F64)
    if div:
        y = y / g.ds0
    q = lambda v: 5.01e-4 * np.abs(v) + 1e-12
    assert out.shape == (g.n, 6)
    assert (np.abs(out[:, 4] - y) <= q(y) + 1e-9 * np.abs(y).max()).all()
    lod = Hg.lod
    for c, arr in ((0, g.xcol), (1, g.ycol)):
        assert (np.abs(out[:, c] - arr[lod]) <= 5.01e-4 * np.abs(arr[lod]) + 1e-9).all()


def test_gpu_driver_3dhr_file_reproduces_3dph(tmp_path):
    """3dhr reads rectangles from a file (main_LH.f90:460-471); written from
    the 3dph mesh it gives the 3dph matrix: identical forward outputs."""
    rect = P.mesh_3dph(24, 12, 1.0, 60.0)
    rows = np.stack([rect.xcol, rect.ycol, rect.zcol, rect.ang, rect.angd, rect.dsl, rect.dsd], axis=1)
    np.savetxt(tmp_path / "rect.dat", rows, fmt="%.17e")
    _run(tmp_path, {"problem": "3dhr", "ncellg": rect.n, "geometry_file": "rect.dat", "forward": True,
                    "eps_h": 1e-6, "leaf": 32, "filenumber": 1})
    _run(tmp_path, {"problem": "3dph", "imax": 24, "jmax": 12, "ds": 1.0, "dipangle": 60.0, "forward": True,
                    "eps_h": 1e-6, "leaf": 32, "filenumber": 2})
    a = np.loadtxt(tmp_path / "output" / "stress1")
    b = np.loadtxt(tmp_path / "output" / "stress2")
    assert np.array_equal(a, b)
