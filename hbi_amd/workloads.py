"""Workloads beside the headline benchmark, each printing one JSON line.

    python -m hbi_amd.workloads c3 [--length 60] [--width 30] [--edge 0.134]
        [--leaf 64] [--eps 1e-5] [--steps 200]
    python -m hbi_amd.workloads c4b [--faults 16] [--steps 200]

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


def c4b(nfaults=16, steps=200, seed=4, nmin=2000, nmax=30000):
    """Distinct planar faults (100 x 50 km, imax x imax/2 elements, sizes
    drawn uniformly from [nmin, nmax] with a seed), each with its own
    H-matrix (leaf 64, ACA eps 1e-5), friction realisation and step control,
    mixed precision: nsteps accepted steps of every fault, one after another
    and concurrently (hbi_rk_step_many)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, nfaults)
    runs_data = []
    t0 = time.perf_counter()
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
        runs_data.append((g, Hm, fr, lod, y0, ds0))
    t_build = time.perf_counter() - t0
    ns = [d[0].n for d in runs_data]
    _log("c4b: %d faults, %d..%d elements (%d total), built in %.1f s" % (nfaults, min(ns), max(ns), sum(ns), t_build))

    def make_runs():
        rs = []
        for g, Hm, fr, lod, y0, ds0 in runs_data:
            rk = H.RK(Hm, fr, lod, g.ds, hdiv=ds0, eps_r=1e-6, prec=H.PREC_MIXED)
            rk.set_state(y0, 0.0, 1.0)
            rk.step(5)                   # graph capture outside the timing
            rs.append(rk)
        return rs

    rs = make_runs()
    t0 = time.perf_counter()
    for rk in rs:
        rk.step(steps)
    wall_seq = time.perf_counter() - t0
    for rk in rs:
        rk.close()
    rs = make_runs()
    t0 = time.perf_counter()
    infos = H.RK.step_many(rs, steps)
    wall_many = time.perf_counter() - t0
    for rk in rs:
        rk.close()
    total = nfaults * steps
    return {"workload": "c4b", "faults": nfaults, "elements": ns, "steps_per_fault": steps, "build_s": t_build,
            "sequential": {"wall_s": wall_seq, "steps_per_s": total / wall_seq},
            "concurrent": {"wall_s": wall_many, "steps_per_s": total / wall_many,
                           "all_ok": all(i["status"] == 0 and i["accepted"] >= steps + 5 for i in infos)},
            "speedup": wall_seq / wall_many}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("c3", "c4b"))
    ap.add_argument("--faults", type=int, default=16)
    ap.add_argument("--length", type=float, default=60.0)
    ap.add_argument("--width", type=float, default=30.0)
    ap.add_argument("--edge", type=float, default=0.134)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--eps", type=float, default=1e-5)
    ap.add_argument("--steps", type=int, default=200)
    a = ap.parse_args()
    if a.workload == "c3":
        print(json.dumps(c3(a.length, a.width, a.edge, a.leaf, a.eps, a.steps)), flush=True)
    else:
        print(json.dumps(c4b(a.faults, a.steps)), flush=True)


if __name__ == "__main__":
    main()
