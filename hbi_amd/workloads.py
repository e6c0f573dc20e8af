"""Workloads beside the headline benchmark, each printing one JSON line.

    python -m hbi_amd.workloads c3 [--length 60] [--width 30] [--edge 0.134]
        [--leaf 64] [--eps 1e-5] [--steps 200]
    python -m hbi_amd.workloads c4b [--faults 16] [--steps 200]
    python -m hbi_amd.workloads c4a [--steps 1500]

c3: a 60 x 30 km self-similar rough fault surface (Hurst 0.8, amplitude to
wavelength 1e-2, seed 0) meshed with right triangles of edge `edge` km
(0.134 km: 200 704 triangles), TDstressFS full-space kernel, slip law, the
a-b depth profile of the antiplane setup.  Reports the GPU H-matrix build,
the single-vector product of every precision, the 64-column product, and
accepted RK steps/s of single runs (FP64 and mixed) from a seeded state.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import lib as H
from . import problems as P


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def c3(length=60.0, width=30.0, edge=0.134, leaf=64, eps=1e-5, steps=200):
    t0 = time.perf_counter()
    tri = P.rough_surface_triangles(length, width, edge, hurst=0.8, alpha=1e-2, seed=0)
    g = P.mesh_triangles(tri)
    _log("c3: %d triangles, mesh %.1f s; building the H-matrix" % (g.n, time.perf_counter() - t0))
    t0 = time.perf_counter()
    Hm = H.HMatrix(g, leaf=leaf, eps=eps, bf16_margin=1.0)
    t_build = time.perf_counter() - t0
    st = Hm.stats
    _log("c3: built in %.1f s (dense fill %.1f s, ACA %.1f s)" % (t_build, st["t_dense"], st["t_aca"]))
    out = {"workload": "c3", "n": g.n, "edge_km": edge, "leaf": leaf, "eps_aca": eps, "build_s": t_build,
           "hmat": {k: st[k] for k in ("nlr", "ndense", "kt_mean", "kt_max", "dense_entries", "lr_entries",
                                       "nbf16_blocks", "ncapped", "t_partition", "t_dense", "t_aca", "t_pack")},
           "stored_bytes": st["stored_bytes"], "matvec": {}, "matmul64": {}, "rk": {}}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_F32, "f32"), (H.PREC_MIXED, "mixed")):
        t = Hm.time_matvec(prec, 20)
        out["matvec"][name] = {"ms": t["total_ms"], "GBps": st["alg_bytes"][prec] / (t["total_ms"] * 1e-3) / 1e9,
                               "a_GBps": st["alg_bytes_p1"][prec] / (t["a_ms"] * 1e-3) / 1e9,
                               "b_GBps": st["alg_bytes_p2"][prec] / (t["b_ms"] * 1e-3) / 1e9}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_MIXED, "mixed")):
        t = Hm.time_matmul(prec, 5)
        out["matmul64"][name] = {"ms": t["total_ms"], "per_rhs_vs_matvec": out["matvec"][name]["ms"] * 64 / t["total_ms"]}
    _log("c3: products timed; RK runs of %d steps" % steps)
    fr = P.friction_c3(g)
    lod = Hm.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod, amp=0.02, depth=12.0, width=2.0)
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_MIXED, "mixed")):
        rk = H.RK(Hm, fr, lod, g.ds, eps_r=1e-6, prec=prec)
        rk.set_state(y0, 0.0, 1.0)
        t0 = time.perf_counter()
        info, hist, _ = rk.step(steps, history=True)
        wall = time.perf_counter() - t0
        out["rk"][name] = {"steps_per_s": steps / wall, "attempts": info["attempts"], "t_end_s": info["t"],
                           "log10_vmax_end": float(hist[-1, 4]), "status": info["status"]}
        rk.close()
    return out


C4A_IMAX, C4A_JMAX, C4A_DS0 = 128, 128, 0.390625


def c4a_setup(nreal=64, seed=7):
    """The 128 x 128 planar fault (50 x 50 km, 3dph half-space, leaf 64, ACA
    eps 1e-5) with nreal friction realisations (a-b ~ N(-0.003, 0.001) in a
    40 x 20 km patch, Dc ~ lognormal(median 15 mm, sigma_log 0.3)), each
    seeded with the same nucleation bump at 25 km depth; 16 probe elements
    along the patch's mid-depth.  Returns (g, Hm, frs, lod, ys, probes)."""
    g = P.mesh_3dph(C4A_IMAX, C4A_JMAX, C4A_DS0)
    Hm = H.HMatrix(g, leaf=64, eps=1e-5, bf16_margin=1.0)
    frs = P.friction_ensemble(g, C4A_IMAX, C4A_JMAX, C4A_DS0, nreal, seed=seed)
    lod = Hm.lod
    ys = []
    for f in frs:
        y = P.init_state(f, lod)
        P.perturb_state(y, g, lod, amp=0.02, depth=25.0, width=3.0)
        ys.append(y)
    # probes: 16 elements at 25 km depth, x from -18 to +18 km (cluster order)
    inv = np.empty(g.n, dtype=np.int64)
    inv[lod] = np.arange(g.n)
    xs = np.linspace(-18.0, 18.0, 16)
    mesh_ids = [int(np.argmin((g.xcol - x) ** 2 + (g.zcol + 25.0) ** 2)) for x in xs]
    probes = inv[mesh_ids].astype(np.int32)
    return g, Hm, frs, lod, np.array(ys), probes


def _onset(t, lv, thresh=-3.0):
    k = int(np.argmax(lv > thresh))
    return float(t[k]) if lv[k] > thresh else None


def ensemble_errors(hf, hm):
    """Errors of ensemble histories hm against reference histories hf (both
    (nreal, steps, 5 + 2 p) rows of hbi_ens_step with p probes): over the
    realisations whose log10 Vmax reaches -3 in both, the relative event
    onset difference, and over the quasi-static phase -- up to half the
    onset or the first time log10 Vmax exceeds -6, whichever is earlier (a
    precursor may reach fast slip rates long before the onset) -- the max
    |d ln Vmax| and the max relative difference of the probe slips on hf's
    step times."""
    on_err, lv_err, sl_err = [], [], []
    for r in range(hf.shape[0]):
        of, om = _onset(hf[r, :, 0], hf[r, :, 4]), _onset(hm[r, :, 0], hm[r, :, 4])
        if of is None or om is None:
            continue
        on_err.append(abs(om - of) / of)
        k6 = int(np.argmax(hf[r, :, 4] > -6.0))
        t_end = min(0.5 * of, hf[r, k6, 0]) if hf[r, k6, 4] > -6.0 else 0.5 * of
        sel = (hf[r, :, 0] <= t_end) & (hf[r, :, 0] >= hm[r, 0, 0])
        tg = hf[r, sel, 0]
        lv_err.append(float(np.abs(np.interp(tg, hm[r, :, 0], hm[r, :, 4]) - hf[r, sel, 4]).max() * np.log(10)))
        sf, sm = hf[r, :, 6::2], hm[r, :, 6::2]
        if sf.shape[1]:
            sl = max(np.abs(np.interp(tg, hm[r, :, 0], sm[:, q]) - sf[sel, q]).max() for q in range(sf.shape[1]))
            sl_err.append(float(sl / max(np.abs(sf[sel]).max(), 1e-300)))
    return {"realisations_with_onset": len(on_err), "onset_rel_max": max(on_err) if on_err else None,
            "dlnv_max_quasistatic": max(lv_err) if lv_err else None,
            "probe_slip_rel_max_quasistatic": max(sl_err) if sl_err else None}


def run_ensemble(Hm, frs, lod, ys, probes, prec, steps, hdiv):
    """One lockstep ensemble from the states ys for `steps` accepted steps:
    (history rows, summary)."""
    E = H.Ensemble(Hm, frs, lod, hdiv=hdiv, eps_r=1e-6, prec=prec)
    E.set_probes(probes)
    E.set_state(ys, 0.0, 1.0)
    E.step(5)                       # graph capture outside the timing
    E.set_state(ys, 0.0, 1.0)
    t1 = time.perf_counter()
    att, hist = E.step(steps, history=True)
    wall = time.perf_counter() - t1
    es = E.get_state()
    summ = {"attempts": int(att), "wall_s": wall, "device_ms": E.last_ms(),
            "steps_per_s": float(es["accepted"].sum()) / wall,
            "sim_time_per_wall_s": float(es["t"].sum()) / wall,
            "stopped": int((es["status"] != 0).sum())}
    E.close()
    return hist, summ


def grid_errors(hf, hm, thr=-6.0):
    """Errors of histories hm replayed on the step grid of hf (identical
    times), per realisation up to the first row where hf's log10 Vmax
    exceeds thr (the whole history when it never does): max |d ln Vmax|,
    max |d ln V| and max relative slip difference at the probes (relative to
    the largest probe slip), and the rows compared (those the replay reached).
    thr = -6: the quasi-static phase."""
    lv, lp, sl, rows = 0.0, 0.0, 0.0, 0
    for r in range(hf.shape[0]):
        k = int(np.argmax(hf[r, :, 4] > thr))
        end = k if hf[r, k, 4] > thr else hf.shape[1]
        if end == 0:
            continue
        alive = hm[r, :end, 0] == hf[r, :end, 0]     # rows the replay reached (it stops if it diverges)
        end = int(np.argmin(alive)) if not alive.all() else end
        if end == 0:
            continue
        a, b = hf[r, :end], hm[r, :end]
        lv = max(lv, float(np.abs(a[:, 4] - b[:, 4]).max() * np.log(10)))
        va, vb = a[:, 5::2], b[:, 5::2]
        if va.shape[1]:
            lp = max(lp, float(np.abs(np.log(np.abs(vb) / np.abs(va))).max()))
            sa, sb = a[:, 6::2], b[:, 6::2]
            sl = max(sl, float(np.abs(sb - sa).max() / max(np.abs(sa).max(), 1e-300)))
        rows += end
    return {"log10v_below": thr, "rows": rows, "dlnvmax": lv, "dlnv_probes": lp, "slip_rel": sl}


def c4a(steps=1500, nreal=64, save=None):
    """FP64 and mixed (FP32 factors on the matrix cores) lockstep ensembles
    of the C4a workload from the same states, plus an FP64 ensemble whose
    initial tau is perturbed by 1e-12 relative (the integration's own
    sensitivity to last-bit changes: a different step grid).  Reports
    accepted steps/s of each and ensemble_errors of the mixed and perturbed
    runs against the FP64 run."""
    t0 = time.perf_counter()
    g, Hm, frs, lod, ys, probes = c4a_setup(nreal)
    _log("c4a: %d elements, %d realisations, set up in %.1f s" % (g.n, nreal, time.perf_counter() - t0))
    out = {"workload": "c4a", "n": g.n, "nreal": nreal, "steps": steps, "runs": {}}
    yp = ys.copy()
    yp[:, 1::3] *= 1 + 1e-12
    hists = {}
    for name, prec, y in (("f64", H.PREC_F64, ys), ("f64_perturbed", H.PREC_F64, yp), ("mixed", H.PREC_MIXED, ys)):
        hists[name], out["runs"][name] = run_ensemble(Hm, frs, lod, y, probes, prec, steps, C4A_DS0)
        _log("c4a %s: %s" % (name, out["runs"][name]))
    if save:
        np.savez_compressed(save, probes=probes, **hists)
    out["mixed_vs_f64"] = ensemble_errors(hists["f64"], hists["mixed"])
    out["f64_perturbed_vs_f64"] = ensemble_errors(hists["f64"], hists["f64_perturbed"])
    return out


def c4b_setup(nfaults=16, seed=4, nmin=2000, nmax=30000):
    """Distinct planar faults (100 x 50 km, imax x imax/2 elements, sizes
    drawn uniformly from [nmin, nmax] with a seed), each with its own
    H-matrix (leaf 64, ACA eps 1e-5) and friction realisation: a list of
    (geometry, H-matrix, friction, lod, initial state, element size)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, nfaults)
    out = []
    for f, n in enumerate(sizes):
        imax = max(4, int(round(np.sqrt(2.0 * n) / 2.0)) * 2)
        jmax = imax // 2
        ds0 = 100.0 / imax
        g = P.mesh_3dph(imax, jmax, ds0)
        Hm = H.HMatrix(g, leaf=64, eps=1e-5, bf16_margin=1.0)
        fr = P.friction_c2(g, imax, jmax, ds0, seed=100 + f)
        lod = Hm.lod
        y0 = P.init_state(fr, lod)
        P.perturb_state(y0, g, lod, amp=0.02, depth=15.0, width=3.0)
        out.append((g, Hm, fr, lod, y0, ds0))
    return out


def c4b_runs(faults, prec=H.PREC_MIXED, warm=5):
    """One run per fault (own step control, eps_r 1e-6), warm steps taken so
    that graph capture falls outside any timing."""
    rs = []
    for g, Hm, fr, lod, y0, ds0 in faults:
        rk = H.RK(Hm, fr, lod, g.ds, hdiv=ds0, eps_r=1e-6, prec=prec)
        rk.set_state(y0, 0.0, 1.0)
        if warm:
            rk.step(warm)
        rs.append(rk)
    return rs


def c4b(nfaults=16, steps=200, seed=4, nmin=2000, nmax=30000):
    """The faults of c4b_setup, mixed precision: steps accepted steps of every
    fault, one run after another and all together (hbi_rk_step_many, one graph
    launch per lockstep round); the two give bit-identical states."""
    t0 = time.perf_counter()
    faults = c4b_setup(nfaults, seed, nmin, nmax)
    t_build = time.perf_counter() - t0
    ns = [d[0].n for d in faults]
    _log("c4b: %d faults, %d..%d elements (%d total), built in %.1f s" % (nfaults, min(ns), max(ns), sum(ns), t_build))
    rs = c4b_runs(faults)
    t0 = time.perf_counter()
    for rk in rs:
        rk.step(steps)
    wall_seq = time.perf_counter() - t0
    seq = [rk.get_state() for rk in rs]
    for rk in rs:
        rk.close()
    rs = c4b_runs(faults)
    t0 = time.perf_counter()
    infos = H.RK.step_many(rs, steps)
    wall_many = time.perf_counter() - t0
    many = [rk.get_state() for rk in rs]
    for rk in rs:
        rk.close()
    same = all(all(np.array_equal(a[k], b[k]) for k in range(3)) for a, b in zip(seq, many))
    total = nfaults * steps
    return {"workload": "c4b", "faults": nfaults, "elements": ns, "steps_per_fault": steps, "build_s": t_build,
            "sequential": {"wall_s": wall_seq, "steps_per_s": total / wall_seq},
            "concurrent": {"wall_s": wall_many, "steps_per_s": total / wall_many,
                           "all_ok": all(i["status"] == 0 and i["accepted"] >= steps + 5 for i in infos)},
            "identical": same, "speedup": wall_seq / wall_many}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("c3", "c4a", "c4b"))
    ap.add_argument("--faults", type=int, default=16)
    ap.add_argument("--length", type=float, default=60.0)
    ap.add_argument("--width", type=float, default=30.0)
    ap.add_argument("--edge", type=float, default=0.134)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--eps", type=float, default=1e-5)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--save", default=None, help="c4a: write the ensemble histories to this .npz")
    a = ap.parse_args()
    if a.workload == "c3":
        print(json.dumps(c3(a.length, a.width, a.edge, a.leaf, a.eps, a.steps)), flush=True)
    elif a.workload == "c4a":
        print(json.dumps(c4a(a.steps, save=a.save)), flush=True)
    else:
        print(json.dumps(c4b(a.faults, a.steps)), flush=True)


if __name__ == "__main__":
    main()
