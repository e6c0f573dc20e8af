"""GPU: main_fv3dp's fluid-driven cycle on a 3dp fault -- the pore pressure
on the fault's imax x jmax grid, solved by the cooperative multi-workgroup
conjugate-gradient kernel -- against the oracle (oracle/fluid3dp.py):

* the pressure step for each setting and boundary set, both solvers, the
  wells of the reference's examples/wells.dat on its 100 x 100 grid, and the
  200 x 200 grid of examples/3dinjection.in (40 workgroups of 256 threads,
  four cells per thread): iteration
  counts within one, solutions within the two solvers' residual bounds, the
  step bound;
* the first steps of a run: identical step grids, monitor rows and states;
* at 200 x 200 with Neumann boundaries the run holds the injected volume.
"""
import types

import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import problems as P
from oracle import fluid3dp as F3
from oracle import oracle_lib as O

pytestmark = pytest.mark.gpu

TOL = float(np.float32(1e-6))    # Beuler's residual tolerance (main_fv3dp.f90:1434)


def _setup(im, jm, ds0=0.02, eps=1e-4, oracle_matrix=False, **kw):
    g0, _, _, _, _ = P.fluid3dp_injection(im, jm, ds0, **kw)
    Hs = H.HMatrix(g0, leaf=32, eps=eps)
    g, opts, arr, y, pf = P.fluid3dp_injection(im, jm, ds0, lod=Hs.lod, **kw)
    Os = O.HMatrix(g, leaf=32, eps=eps, nthreads=8) if oracle_matrix else types.SimpleNamespace(n=g.n)
    if oracle_matrix:
        assert np.array_equal(Os.lod, Hs.lod)
    dev = H.FV(Hs, opts, arr)
    ora = F3.Fv3Oracle(Os, opts, arr, Hs.lod)
    return g, Hs, dev, ora, opts, arr, y, pf


DIFFUSE_CASES = [
    # (imax, jmax, bc, setting, pressure_dependent, wells)
    (40, 30, "Dirichlet", 0, False, False),
    (36, 28, "Neumann", 2, False, False),      # 'ss'
    (32, 40, "Dirichlet", 3, False, False),    # 'thrust'
    (40, 30, "Neumann", 0, True, False),
    (100, 100, "Dirichlet", 0, False, True),   # examples/wells.dat
    (200, 200, "Neumann", 0, False, False),    # examples/3dinjection.in's grid
]


@pytest.mark.parametrize("im,jm,bc,setting,pdep,wells", DIFFUSE_CASES)
def test_gpu_fv3dp_pressure_step_against_oracle(im, jm, bc, setting, pdep, wells):
    g, Hs, dev, ora, opts, arr, y, pf = _setup(im, jm, bc=bc, pressure_dependent=pdep, s0=5.0)
    n = g.n
    rng = np.random.default_rng(im * jm + setting)
    if setting != 0:
        ora.o["setting"] = setting
        ora.grid.o["setting"] = setting
        ora.o["pbcl"], ora.o["pbcr"] = 2.0, 0.5
        ora.grid.o["pbcl"], ora.grid.o["pbcr"] = 2.0, 0.5
        ora.arr["pfhyd"][:] = 0.1 * rng.random(n)
        dev = H.FV(Hs, ora.o, ora.arr)
    if wells:
        dev.set_wells(P.WELLS_3DINJ2)
        ora.grid.wells = ora.wells = P.WELLS_3DINJ2
    pf = pf + rng.random(n)
    y[3::4] = 1e-14 * (1 + 2 * rng.random(n))
    sigma = y[2::4] + pf
    for h, t in ((2e3, 5e6), (3e5, 1.5e7)):
        print("pressure step", im, jm, bc, setting, pdep, wells, h, flush=True)
        pg, ig, dtg = dev.diffuse(pf, y, sigma, h, t, 1e6)
        po, io, dto = ora.diffuse(pf, y, sigma, h, t, 1e6)
        print("  iterations", ig, io, "max diff", np.abs(pg - po).max(), "dt", dtg, dto, flush=True)
        assert ig > 0 and abs(ig - io) <= max(1, io // 50), (ig, io)
        # both solve the same system to |r| < tol sqrt(N); the matrix is I + h L
        # with L positive semi-definite in the boundary-weighted norm
        assert np.linalg.norm(pg - po) <= 4 * TOL * np.sqrt(n)
        if ig == io and io < 60:
            assert np.abs(pg - po).max() <= 1e-9 * max(1.0, np.abs(po).max())
        assert abs(dtg - dto) <= 1e-6 * abs(dto)


def test_gpu_fv3dp_run_against_oracle():
    im, jm = 24, 16
    g, Hs, dev, ora, opts, arr, y, pf = _setup(im, jm, oracle_matrix=True, eps=1e-6, q0=3e-3)
    n = g.n
    dev.set_state(y, pf, t=0.0, dt=1.0)
    st = dict(y=y.copy(), pf=pf.copy(), sigma=y[2::4] + pf, disp=np.zeros(n), t=0.0, h=1.0)
    K = 25
    rows = [ora.step(st) for _ in range(K)]
    info, hist = dev.step(K, history=True)
    print("run", info, flush=True)
    assert info["status"] == 0 and info["accepted"] == K
    assert info["rejected"] == sum(r["rejected"] for r in rows)
    for k, r in enumerate(rows):
        hr = hist[k]
        assert abs(hr[8] - r["dt"]) <= 1e-9 * r["dt"], k
        assert abs(hr[0] - r["t"]) <= 1e-9 * r["t"], k
        assert abs(int(hr[9]) - r["niter"]) <= 1, k
        assert abs(hr[1] - r["log10vmax"]) <= 1e-9 * abs(r["log10vmax"]), k
        assert abs(hr[3] - r["meanmu"]) <= 1e-12 * abs(r["meanmu"]), k
        assert abs(hr[5] - r["minnorm"]) <= 1e-9 * abs(r["minnorm"]), k
    s = dev.get_state()
    assert np.abs(s["pf"] - st["pf"]).max() <= 1e-6 * max(1.0, np.abs(st["pf"]).max())
    for c in range(4):
        ref = st["y"][c::4]
        assert np.abs(s["y"][c::4] - ref).max() <= 1e-9 * np.abs(ref).max(), c
    assert np.abs(s["vel"] - st["vel"]).max() <= 1e-8 * np.abs(st["vel"]).max()
    assert st["pf"].max() > 0


def test_gpu_fv3dp_example_grid_holds_the_injected_volume():
    """200 x 200 (examples/3dinjection.in's grid, 40 000 cells), Neumann
    boundaries, the centre source: after 40 steps the boundary-halved
    pressure sum equals the injected volume to the solver's residual bound,
    and pf peaks at the source."""
    im = jm = 200
    g, Hs, dev, ora, opts, arr, y, pf = _setup(im, jm, ds0=0.01, bc="Neumann", q0=3e-4)
    dev.set_state(y, pf, t=0.0, dt=1.0)
    info, hist = dev.step(40, history=True)
    print("example grid", info, "CG iterations", hist[:, 9].min(), hist[:, 9].max(), flush=True)
    assert info["status"] == 0 and len(hist) == 40
    s = dev.get_state()
    cell = np.empty(g.n, dtype=np.int64)
    cell[Hs.lod] = np.arange(g.n)
    pfg = s["pf"][cell].reshape(im, jm)
    w = np.ones((im, jm))
    w[[0, -1], :] *= 0.5
    w[:, [0, -1]] *= 0.5
    rate = opts["q0"] / opts["beta"] / opts["phi0"] * float(np.float32(1e-12)) / opts["ds0"] ** 2
    vol = rate * hist[:, 8].sum()
    assert vol > 0
    assert abs((w * pfg).sum() - vol) <= 40 * 2 * TOL * np.sqrt(g.n)
    assert np.unravel_index(np.argmax(pfg), pfg.shape) == (im // 2 - 1, jm // 2 - 1)
    assert hist[:, 9].max() > 0
