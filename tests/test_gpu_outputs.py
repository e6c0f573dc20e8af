"""GPU: field snapshots and the event list decided on the device
(hbi_rk_set_output / hbi_rk_get_output) against the reference's per-step
output logic restated on the host (oracle/outputs.py, main_LH.f90:1339-1407),
applied to the run's own per-step history.  Step numbers, times, event rows,
hypocentres and coseismic slip are compared exactly; Mw to 1e-12 (device and
host log10)."""
import math

import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import problems as P
from oracle import outputs as OUT

pytestmark = pytest.mark.gpu

VELTH, DT_INTER, DT_CO, INTERVAL = 1e-2, 2.0e7, 1.0, 400


def _c1_run(nsteps, capacity, ev_capacity=8, drain_every=None):
    g = P.mesh_2dvs(2048, 40.0 / 2048)
    fr = P.friction_c1(g)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod)
    rk = H.RK(Hm, fr, lod, g.ds, eps_r=1e-6)
    rk.set_output(velth=VELTH, dtout_inter=DT_INTER, dtout_co=DT_CO, interval=INTERVAL, capacity=capacity,
                  ev_capacity=ev_capacity)
    rk.set_state(y0, 0.0, 1.0)
    hists, outs, calls = [], [], 0
    left = nsteps
    while left > 0:
        info, hist, _ = rk.step(left, history=True)
        calls += 1
        n = info["accepted"] - (nsteps - left)
        hists.append(hist[:n])
        left -= n
        if info["status"] == H.RK_OUTPUT_FULL or left == 0:
            outs.append(rk.get_output())
        else:
            assert info["status"] == 0, info
    hist = np.concatenate(hists)
    out = {"k": np.concatenate([o["k"] for o in outs]), "t": np.concatenate([o["t"] for o in outs]),
           "events": sum((o["events"] for o in outs), []),
           "eqslip": np.concatenate([o["eqslip"] for o in outs])}
    for f in ("vel", "slip", "sigma", "tau"):
        out[f] = np.concatenate([o[f] for o in outs])
    area = np.asarray(g.ds, dtype=np.float64)[lod]
    return rk, hist, out, calls, area


def test_gpu_outputs_match_reference_logic():
    rk, hist, out, _, area = _c1_run(3000, capacity=512)
    assert len(out["events"]) >= 1, "the run must contain a whole event"
    # per-step (t, max|V|, mean slip, argmax) from the run's own history; the
    # first max |V| at an onset is read from that step's snapshot
    snap_of = {int(k): i for i, k in enumerate(out["k"])}
    steps = []
    for j in range(len(hist)):
        k = j + 1
        loc = int(np.argmax(np.abs(out["vel"][snap_of[k]]))) if k in snap_of else -1
        steps.append((hist[j, 0], 10.0 ** hist[j, 4], hist[j, 5], loc))
    snaps, events, onsets = OUT.decisions(steps, VELTH, DT_INTER, DT_CO, INTERVAL, float(area.sum()))
    assert list(out["k"]) == snaps
    assert np.array_equal(out["t"], hist[np.array(snaps) - 1, 0])
    assert len(out["events"]) == len(events)
    for got, want in zip(out["events"], events):
        for key in ("count", "k", "onset_time", "moment", "hypoloc"):
            assert got[key] == want[key], (key, got, want)
        assert abs(got["mw"] - want["mw"]) <= 1e-12 * abs(want["mw"])
    # coseismic slip: |slip at the end - slip at the onset| from the snapshots
    for e, (ev, k_on) in enumerate(zip(out["events"], onsets)):
        s_on, s_end = out["slip"][snap_of[k_on]], out["slip"][snap_of[ev["k"]]]
        assert np.array_equal(out["eqslip"][e], np.abs(s_end - s_on))
    # snapshot fields against the history of the same step: area-weighted
    # mean |slip|, max / min sigma
    for i, k in enumerate(out["k"]):
        row = hist[k - 1]
        assert abs((np.abs(out["slip"][i]) * area).sum() / area.sum() - row[5]) <= 1e-12 * abs(row[5]) + 1e-300
        assert out["sigma"][i].max() == row[7] and out["sigma"][i].min() == row[8]
    # a snapshot at the last step is the final state
    y, slip, vel, info = rk.get_state()
    if out["k"][-1] == info["accepted"]:
        assert np.array_equal(out["vel"][-1], vel) and np.array_equal(out["slip"][-1], slip)
        assert np.array_equal(out["sigma"][-1], y[2::3]) and np.array_equal(out["tau"][-1], y[1::3])
    rk.close()


def test_gpu_outputs_drained_in_small_buffers_equal_one_buffer():
    """Buffers that fill stop the batch (status RK_OUTPUT_FULL); draining and
    continuing gives the same snapshots, events and trajectory as one large
    buffer."""
    _, h1, o1, c1, _ = _c1_run(2500, capacity=512)
    _, h2, o2, c2, _ = _c1_run(2500, capacity=3, ev_capacity=1)
    assert c2 > c1
    assert np.array_equal(h1, h2)
    assert np.array_equal(o1["k"], o2["k"]) and np.array_equal(o1["t"], o2["t"])
    for f in ("vel", "slip", "sigma", "tau", "eqslip"):
        assert np.array_equal(o1[f], o2[f]), f
    assert o1["events"] == o2["events"]


def test_gpu_outputs_must_be_drained():
    g = P.mesh_2dvs(512, 40.0 / 512)
    fr = P.friction_c1(g)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    rk = H.RK(Hm, fr, lod, g.ds, eps_r=1e-6)
    rk.set_output(velth=VELTH, dtout_inter=DT_INTER, dtout_co=DT_CO, interval=1, capacity=2)
    y0 = P.init_state(fr, lod)
    y0[1::3] *= 1.01      # off steady state (an exact one underflows the step by rkqs's rule)
    rk.set_state(y0, 0.0, 1.0)
    info, _, _ = rk.step(10)
    assert info["status"] == H.RK_OUTPUT_FULL and info["accepted"] == 2
    with pytest.raises(RuntimeError):
        rk.step(1)
    o = rk.get_output()
    assert list(o["k"]) == [1, 2]
    info, _, _ = rk.step(1)
    assert info["accepted"] == 3
    rk.close()
