"""Benchmark: accepted RK45 steps/s of the earthquake-cycle hot path on one GPU.

Workload: the 32768-element 3-D planar strike-slip fault (256 x 128 elements,
ds 0.390625 km, vertical, half-space Okada kernel), lattice H-matrix with
leaf 64 and ACA eps 1e-5, velocity-weakening patch 60 x 30 km, lognormal a, Dc
from a seed, aging law, sigma 50 MPa, V0 = Vpl = 1e-9 m/s initial state.  A
step is one accepted rkqs step (Cash-Karp, eps_r 1e-6, 6 H-matvecs per
attempt) plus the post-step update, all on the device.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  ``value`` = accepted steps/s of the mixed-precision
form (FP32 factors, BF16 where the ACA tolerance allows, FP64 ODE state);
the plain FP64 form, the H-matvec bandwidth of every precision, the
roofline of the dominant kernel and the CPU oracle baseline ride along.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
IMAX, JMAX, DS0, LEAF, EPS_H, EPS_R = 256, 128, 0.390625, 64, 1e-5, 1e-6


def gpu_part(steps: int, warmup: int, cpu_steps: int):
    from hbi_amd import lib as H
    from hbi_amd import problems as P

    g = P.mesh_3dph(IMAX, JMAX, DS0, 90.0)
    fr = P.friction_c2(g, IMAX, JMAX, DS0, seed=1)
    t0 = time.perf_counter()
    Hm = H.HMatrix(g, leaf=LEAF, eps=EPS_H, bf16_margin=1.0)
    t_build = time.perf_counter() - t0
    st = Hm.stats
    lod = Hm.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod, amp=0.02, depth=15.0, width=3.0)

    out = {"build_s": round(t_build, 3), "hmat": {k: st[k] for k in ("nlr", "ndense", "kt_mean", "kt_max",
                                                                        "dense_entries", "lr_entries",
                                                                        "nbf16_blocks")}}
    # matvec kernels (HIP events on the library stream)
    mv = {}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_F32, "f32"), (H.PREC_MIXED, "mixed")):
        t = Hm.time_matvec(prec, 50)
        mv[name] = {"ms": t["total_ms"], "lrvt_ms": t["lrvt_ms"], "tile_ms": t["tile_ms"],
                    "alg_bytes": st["alg_bytes"][prec],
                    "GBps": st["alg_bytes"][prec] / (t["total_ms"] * 1e-3) / 1e9,
                    "tile_GBps": st["alg_bytes_p2"][prec] / (t["tile_ms"] * 1e-3) / 1e9,
                    "lrvt_GBps": st["alg_bytes_p1"][prec] / (t["lrvt_ms"] * 1e-3) / 1e9}
        mv[name]["pct_hbm"] = 100.0 * mv[name]["GBps"] / HBM_PEAK_GBS
    out["matvec"] = mv

    res = {}
    finals = {}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_MIXED, "mixed")):
        rk = H.RK(Hm, fr, lod, g.ds, hdiv=DS0, eps_r=EPS_R, prec=prec)
        rk.set_state(y0, 0.0, 1.0)
        if warmup:
            rk.step(warmup)
        t0 = time.perf_counter()
        info, hist, _ = rk.step(steps, history=True)
        wall = time.perf_counter() - t0
        dev_ms = rk.last_ms()
        y, slip, vel, _ = rk.get_state()
        finals[name] = (hist, vel)
        res[name] = {"steps_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps, "device_ms": dev_ms,
                     "attempts": info["attempts"], "rejected": info["rejected"], "t_end_s": info["t"],
                     "vmax": info["vmax"]}
        rk.close()
    h64, v64 = finals["f64"]
    hmx, vmx = finals["mixed"]
    res["mixed_vs_f64"] = {
        "max_rel_slip_rate": float(np.max(np.abs(vmx - v64) / np.abs(v64))),
        "max_rel_time": float(np.max(np.abs(hmx[:, 0] - h64[:, 0]) / h64[:, 0])),
        "speedup": res["mixed"]["steps_per_s"] / res["f64"]["steps_per_s"]}
    out["rk"] = res
    Hm.close()
    return out


def cpu_part(nsteps: int):
    """CPU oracle (FP64, OpenMP over all host cores) on a bounded sample:
    the same fault and matrix settings, nsteps accepted steps."""
    from hbi_amd import problems as P
    from oracle import oracle_lib as O

    O.build()
    ncores = os.cpu_count() or 1
    try:
        ncores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    g = P.mesh_3dph(IMAX, JMAX, DS0, 90.0)
    fr = P.friction_c2(g, IMAX, JMAX, DS0, seed=1)
    t0 = time.perf_counter()
    Ho = O.HMatrix(g, leaf=LEAF, eps=EPS_H, nthreads=ncores)
    t_build = time.perf_counter() - t0
    lod = Ho.lod
    y = P.init_state(fr, lod)
    P.perturb_state(y, g, lod, amp=0.02, depth=15.0, width=3.0)
    sim = O.Sim(Ho, fr, lod, hdiv=DS0, nthreads=ncores)
    a = fr.a[lod]
    vel = 2 * fr.vref * np.exp(-y[0::3] / a) * np.sinh(y[1::3] / y[2::3] / a)
    slip = np.zeros(g.n)
    t0 = time.perf_counter()
    done, *_ = sim.run(y, 0.0, 1.0, EPS_R, nsteps, slip, vel, np.asarray(g.ds)[lod])
    wall = time.perf_counter() - t0
    return {"value": done / wall, "unit": "RK steps/s", "cores": ncores, "kind": "port",
            "sample": "%d accepted RK45 steps (FP64 oracle, OpenMP) of the 32768-element fault after a %.1f s "
                      "H-matrix build" % (done, t_build)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one process, one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    r = gpu_part(args.steps, args.warmup, args.cpu_steps)
    mv = r["matvec"]["mixed"]
    tile_bytes = None
    from hbi_amd import lib as H  # noqa: F401
    line = {
        "metric": "RK steps/s & H-matvec GB/s (%HBM peak), mixed+learned vs FP64, 1e-5 rel err",
        "value": round(r["rk"]["mixed"]["steps_per_s"], 3),
        "unit": "accepted RK steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["rk"]["mixed"]["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state / f32+bf16 H factors",
        "data": "synthetic",
        "config": {"workload": "3dph 256x128 planar fault, N=32768, leaf 64, ACA eps 1e-5, RK45 eps_r 1e-6, "
                               "mixed-precision H (FP32/BF16 factors), hipGraph RK attempts",
                   "n_elements": IMAX * JMAX},
        "roofline": {"bound": "hbm", "kernel": "tile_kernel (row-tile H-matvec, mixed)",
                     "achieved": round(mv["tile_GBps"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(mv["tile_GBps"] / HBM_PEAK_GBS, 4), "traffic": tile_bytes},
        "fp64_steps_per_s": round(r["rk"]["f64"]["steps_per_s"], 3),
        "speedup_vs_fp64": round(r["rk"]["mixed_vs_f64"]["speedup"], 3),
        "mixed_vs_fp64_max_rel_slip_rate": r["rk"]["mixed_vs_f64"]["max_rel_slip_rate"],
        "matvec": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                   for k, v in r["matvec"].items()},
        "rk": r["rk"],
        "build_s": r["build_s"],
        "hmat": r["hmat"],
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_part(args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
