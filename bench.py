"""Benchmark: accepted RK45 steps/s of the earthquake-cycle hot path on one GPU.

Fault: the 32768-element 3-D planar strike-slip fault (256 x 128 elements,
ds 0.390625 km, vertical, half-space Okada kernel), lattice H-matrix with
leaf 64 and ACA eps 1e-5; velocity-weakening patch 60 x 30 km, lognormal a
and Dc from a seed, aging law, sigma 50 MPa, V0 = Vpl = 1e-9 m/s initial
state with a nucleation seed.  A step is one accepted rkqs step (Cash-Karp,
eps_r 1e-6, 6 H products per attempt) plus the post-step update, all on the
device.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline (``value``): the mixed-precision, batched, learned-step form -- 64
realisations of the fault's friction (seeds 1..64) advanced in lockstep with
one 64-column H product per RK stage, FP32 factors split into three BF16
parts on the BF16 matrix cores (FP64 state), each
realisation's next step proposed by the step-size MLP trained on this fault
(hbi_amd/data/step_mlp_c2.npz; within [1/8, 1] x the rkqs proposal; the
embedded error check judges every step) -- as accepted steps/s summed over the realisations, over K lockstep steps
after W warm-up steps.  The same ensemble under the reference's controller
(rkqs) is measured beside it.
Alongside: single-run FP64 (the plain GPU form) and mixed, the single-vector
H-matvec bandwidth of every precision, the 64-column product on the matrix
cores, trajectory agreement mixed vs FP64, the roofline of the dominant
kernel and the CPU oracle baseline.  Prints one JSON line.
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
FP64_MFMA_PEAK_TFS = 78.6
FP32_MFMA_PEAK_TFS = 157.3
BF16_MFMA_PEAK_TFS = 2500.0
# HBM bytes per launch of the roofline kernels (mixed precision), from
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE passes over
# hbi_amd/profile_kernels.py, scaled by the counters of a 256 MiB device copy
# (fetch x 2.00, write x 1.00): mr_bf16_kernel<4,true,2,2> (141292.7 KiB
# fetched, 1016 x 32 KiB partial tiles written at the tuner's usual 1016 items
# of up to 1024 columns); the single-vector product's second launch with the
# dense blocks in it (lowrank_kernel, --dense-launch 2: 110648.8 KiB fetched,
# 256 KiB written) and its first launch when it holds them (stream_kernel,
# --dense-launch 1: 109129.8 KiB fetched, 2887.5 KiB written)
TRAFFIC_MR_BF16_MIXED = 322.6e6
TRAFFIC_LOWRANK_DENSE_MIXED = 226.9e6
TRAFFIC_STREAM_DENSE_MIXED = 226.5e6
# rocprofv3 --pmc MfmaUtil over the same program: fraction of cycles the
# matrix cores of mr_bf16_kernel<4,true,2,2> are busy
MFMA_UTIL_MR_BF16 = 0.235
IMAX, JMAX, DS0, LEAF, EPS_H, EPS_R = 256, 128, 0.390625, 64, 1e-5, 1e-6
NREAL = 64
FLUID_STEPS = 500      # timed accepted steps of the bp6a fluid-driven cycle
FLUID_CPU_STEPS = 200
C1_STEPS = 20000       # accepted steps of each GPU run of the 2-D antiplane fault (its configured run length)
C1_CPU_STEPS = 20000   # the CPU oracle runs the same run length


def _onset(t, lv, thr=-3.0):
    k = int(np.argmax(lv > thr))
    return float(t[k]) if lv[k] > thr else None


def gpu_part(steps: int, warmup: int, single_steps: int):
    from hbi_amd import lib as H
    from hbi_amd import problems as P

    print("bench: building the 32768-element H-matrix", file=sys.stderr, flush=True)
    g = P.mesh_3dph(IMAX, JMAX, DS0, 90.0)
    t0 = time.perf_counter()
    Hm = H.HMatrix(g, leaf=LEAF, eps=EPS_H, bf16_margin=1.0)
    t_build = time.perf_counter() - t0
    st = Hm.stats
    lod = Hm.lod
    lv = Hm.leaves
    dense, lr = lv[:, 0] != 1, lv[:, 0] == 1
    p1_entries = int((lv[lr, 4] * lv[lr, 5]).sum())
    p2_entries = int((lv[dense, 2] * lv[dense, 4]).sum() + (lv[lr, 2] * lv[lr, 5]).sum())
    out = {"build_s": round(t_build, 3),
           "hmat": {k: st[k] for k in ("nlr", "ndense", "kt_mean", "kt_max", "dense_entries", "lr_entries",
                                       "nbf16_blocks", "dense_launch")}}

    # single-vector products (HIP events on the library stream)
    mv = {}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_F32, "f32"), (H.PREC_MIXED, "mixed")):
        t = Hm.time_matvec(prec, 50)
        mv[name] = {"ms": t["total_ms"], "a_ms": t["a_ms"], "b_ms": t["b_ms"],
                    "alg_bytes": st["alg_bytes"][prec],
                    "GBps": st["alg_bytes"][prec] / (t["total_ms"] * 1e-3) / 1e9,
                    "a_GBps": st["alg_bytes_p1"][prec] / (t["a_ms"] * 1e-3) / 1e9,
                    "b_GBps": st["alg_bytes_p2"][prec] / (t["b_ms"] * 1e-3) / 1e9}
        mv[name]["pct_hbm"] = 100.0 * mv[name]["GBps"] / HBM_PEAK_GBS
    out["matvec"] = mv

    # 64-column products on the matrix cores
    mm = {}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_MIXED, "mixed")):
        t = Hm.time_matmul(prec, 20)
        mm[name] = {"ms": t["total_ms"], "p1_ms": t["p1_ms"], "p2_ms": t["p2_ms"],
                    "p1_TFs": p1_entries * 64 * 2 / (t["p1_ms"] * 1e-3) / 1e12,
                    "p2_TFs": p2_entries * 64 * 2 / (t["p2_ms"] * 1e-3) / 1e12,
                    "per_rhs_vs_matvec": mv[name]["ms"] * 64 / t["total_ms"]}
    out["matmul64"] = mm
    out["matmul64_plan"] = {name: Hm.matmul_plan(prec) for prec, name in ((H.PREC_F64, "f64"), (H.PREC_MIXED, "mixed"))}
    # algorithmic bytes of one phase-2 launch (FP32 factors): every factor value once, every X row and Z row
    # (sum kt of them) once in FP32, and the 64-row FP64 partial block of every item written once
    plan = out["matmul64_plan"]["mixed"]
    nz = int(round(st["kt_mean"] * st["nlr"]))
    out["p2_alg_bytes"] = int(p2_entries * 4 + (g.n + nz) * 64 * 4 + plan["items2"] * 64 * 64 * 8)
    # what the matrix cores sustain here for these instructions (no memory traffic)
    out["mfma_f64_loop_tfs"] = {str(w): round(H.mfma_f64_rate(w), 2) for w in (1, 2, 4)}
    out["mfma_f32_loop_tfs"] = {str(w): round(H.mfma_f32_rate(w), 2) for w in (1, 2, 4)}

    # single runs: FP64 (the plain GPU form) and mixed, with slip-rate histories
    fr = P.friction_c2(g, IMAX, JMAX, DS0, seed=1)
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod, amp=0.02, depth=15.0, width=3.0)
    rec = np.arange(0, g.n, 997, dtype=np.int32)
    res, traj = {}, {}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_MIXED, "mixed")):
        print("bench: single run %s, %d steps" % (name, single_steps), file=sys.stderr, flush=True)
        rk = H.RK(Hm, fr, lod, g.ds, hdiv=DS0, eps_r=EPS_R, prec=prec)
        rk.set_state(y0, 0.0, 1.0)
        rk.step(20)
        rk.set_state(y0, 0.0, 1.0)
        t0 = time.perf_counter()
        info, hist, vh = rk.step(single_steps, history=True, rec=rec)
        wall = time.perf_counter() - t0
        traj[name] = (hist, vh)
        res[name] = {"steps_per_s": single_steps / wall, "ms_per_step": 1e3 * wall / single_steps,
                     "device_ms": rk.last_ms(), "attempts": info["attempts"], "rejected": info["rejected"],
                     "t_end_s": info["t"], "onset_s": _onset(hist[:, 0], hist[:, 4])}
        rk.close()
    h64, v64 = traj["f64"]
    hmx, vmx = traj["mixed"]
    on64, onmx = res["f64"]["onset_s"], res["mixed"]["onset_s"]
    t_end = 0.5 * on64 if on64 else min(h64[-1, 0], hmx[-1, 0])
    tg = np.linspace(max(h64[0, 0], hmx[0, 0]), t_end, 4000)
    dlnv = max(float(np.abs(np.interp(tg, h64[:, 0], np.log(np.abs(v64[:, r])))
                            - np.interp(tg, hmx[:, 0], np.log(np.abs(vmx[:, r])))).max()) for r in range(len(rec)))
    out["single"] = res
    out["accuracy_mixed_vs_f64"] = {
        "onset_rel": abs(onmx - on64) / on64 if (on64 and onmx) else None,
        "max_abs_dlnV_to_half_onset": dlnv,
        "note": "slip-rate histories at %d elements compared on a common time grid (step grids differ)" % len(rec)}

    # the fluid-driven cycle (main_fv.f90) on the reference's bp6a example
    print("bench: fluid-driven cycle bp6a, 4000 elements", file=sys.stderr, flush=True)
    gf = P.mesh_2dp(P.BP6A_N, P.BP6A_DS0)
    Hf = H.HMatrix(gf, leaf=32, eps=1e-4)
    fo, fa, fy, fpf = P.fluid_bp6a(gf, Hf.lod)
    fv = H.FV(Hf, fo, fa)
    fv.set_state(fy, fpf, t=0.0, dt=1.0)
    fv.step(20)
    t0 = time.perf_counter()
    finfo, fhist = fv.step(FLUID_STEPS, history=True)
    wall = time.perf_counter() - t0
    out["fluid_bp6a"] = {"steps_per_s": round(FLUID_STEPS / wall, 1), "ms_per_step": round(1e3 * wall / FLUID_STEPS, 4),
                         "device_ms": round(fv.last_ms(), 3), "attempts": finfo["attempts"],
                         "cg_iter_mean": round(float(fhist[:, 9].mean()), 1), "t_end_s": finfo["t"],
                         "log10_vmax_end": round(float(fhist[-1, 1]), 4)}
    fv.close()

    out["c1"] = c1_gpu(H, P, C1_STEPS)
    out["fluid3dp"] = fluid3dp_gpu(H, P)

    # batched form: 64 realisations in lockstep, mixed precision, rkqs and learned step size
    from hbi_amd import learned as L
    mlp = L.load(os.path.join(ROOT, "hbi_amd", "data", "step_mlp_c2.npz"))
    frs = [P.friction_c2(g, IMAX, JMAX, DS0, seed=1 + r) for r in range(NREAL)]
    ys = []
    for f in frs:
        y = P.init_state(f, lod)
        P.perturb_state(y, g, lod, amp=0.02, depth=15.0, width=3.0)
        ys.append(y)
    ys = np.array(ys)
    hists = {}
    for name, learned in (("ensemble_rkqs", False), ("ensemble", True)):
        print("bench: ensemble of %d realisations (%s), %d + %d steps" % (NREAL, name, warmup, steps),
              file=sys.stderr, flush=True)
        E = H.Ensemble(Hm, frs, lod, hdiv=DS0, eps_r=EPS_R, prec=H.PREC_MIXED)
        if learned:
            E.set_controller(1, mlp, **L.C2_BOUNDS)
        E.set_state(ys, 0.0, 1.0)
        hw = E.step(warmup, history=True)[1] if warmup else np.zeros((NREAL, 0, 5))
        t0 = time.perf_counter()
        att, ht = E.step(steps, history=True)
        wall = time.perf_counter() - t0
        es = E.get_state()
        hists[name] = np.concatenate([hw, ht], axis=1)
        t_start = hw[:, -1, 0] if warmup else np.zeros(NREAL)
        out[name] = {"realisations": NREAL, "learned_step": learned, "steps": steps, "wall_s": wall,
                     "device_ms": E.last_ms(), "lockstep_attempts": att, "ms_per_attempt": 1e3 * wall / max(att, 1),
                     "steps_per_s": NREAL * steps / wall,
                     "sim_time_per_wall_s": float((es["t"] - t_start).sum()) / wall,
                     "max_rejected": int(es["rejected"].max()),
                     "status_ok": bool((es["status"] == 0).all())}
        E.close()
    # the two controllers from the same start to a common simulated time per realisation
    h0, h1 = hists["ensemble_rkqs"], hists["ensemble"]
    common = np.minimum(h0[:, -1, 0], h1[:, -1, 0])
    a0 = np.array([L.attempts_to_time(h0[r], common[r]) for r in range(NREAL)])
    a1 = np.array([L.attempts_to_time(h1[r], common[r]) for r in range(NREAL)])
    w0 = a0.max() * out["ensemble_rkqs"]["ms_per_attempt"]
    w1 = a1.max() * out["ensemble"]["ms_per_attempt"]
    out["learned_vs_rkqs"] = {"attempt_ratio_median": float(np.median(a0 / a1)),
                              "attempt_ratio_min": float((a0 / a1).min()),
                              "lockstep_attempts": [int(a0.max()), int(a1.max())],
                              "time_to_common_sim_time_ms": [w0, w1], "speedup": w0 / w1,
                              "note": "attempts (accepted + rejected) of each realisation from t = 0 to the smaller "
                                      "of the two controllers' final times; lockstep wall = max attempts x ms/attempt"}
    out["controllers_by_phase"] = controllers_by_phase(H, Hm, frs, lod, ys, mlp, warmup, steps)
    out["accuracy_ensemble_vs_f64"] = ensemble_accuracy(H, Hm, frs, lod, ys, g)
    out["p2_entries"], out["p1_entries"] = p2_entries, p1_entries
    Hm.close()
    return out


FV3_IM = FV3_JM = 200      # examples/3dinjection.in's grid (40 000 cells of 10 m)
FV3_STEPS = 300
FV3_LEAF = 32


def fluid3dp_gpu(H, P):
    """main_fv3dp's cycle on the 200 x 200 3dp fault of examples/3dinjection.in
    (uniform friction of examples/3dinjection2.in, the centre source q0 3e-4,
    Neumann boundaries, ACA eps 1e-4): FV3_STEPS accepted steps after 20,
    the 2-D pressure solve by the cooperative conjugate-gradient kernel."""
    print("bench: main_fv3dp 200 x 200 grid, %d steps" % FV3_STEPS, file=sys.stderr, flush=True)
    g0, _, _, _, _ = P.fluid3dp_injection(FV3_IM, FV3_JM, 0.01)
    Hm = H.HMatrix(g0, leaf=FV3_LEAF, eps=1e-4)
    g, opts, arr, y, pf = P.fluid3dp_injection(FV3_IM, FV3_JM, 0.01, bc="Neumann", q0=3e-4, lod=Hm.lod)
    fv = H.FV(Hm, opts, arr)
    fv.set_state(y, pf, t=0.0, dt=1.0)
    fv.step(20)
    t0 = time.perf_counter()
    info, hist = fv.step(FV3_STEPS, history=True)
    wall = time.perf_counter() - t0
    res = {"cells": g.n, "steps_per_s": round(FV3_STEPS / wall, 1), "ms_per_step": round(1e3 * wall / FV3_STEPS, 4),
           "device_ms": round(fv.last_ms(), 3), "attempts": info["attempts"],
           "cg_iter_mean": round(float(hist[:, 9].mean()), 2), "cg_iter_max": int(hist[:, 9].max()),
           "t_end_s": info["t"], "log10_vmax_end": round(float(hist[-1, 1]), 4), "status": info["status"]}
    fv.close()
    Hm.close()
    return res


def _events(t, lv, thr=-3.0):
    """onset times: upward crossings of log10 Vmax = thr"""
    up = np.nonzero((lv[1:] > thr) & (lv[:-1] <= thr))[0] + 1
    return [float(t[k]) for k in up]


def c1_gpu(H, P, steps):
    """The 2-D antiplane fault (2048 elements over 40 km, leaf 32, ACA eps
    1e-6, aging law) from the steady state at V = Vpl with a seeded
    nucleation bump (the exact steady state has no dynamics: rkqs underflows
    on it, as the reference's does), `steps` accepted steps of single GPU runs: FP64, mixed, and mixed with the learned step size
    (hbi_amd/data/step_mlp_c1.npz); the single-vector product of each
    precision; event onsets of the mixed runs against the FP64 run's."""
    from hbi_amd import learned as L
    print("bench: C1, 2048-element antiplane fault, %d steps per run" % steps, file=sys.stderr, flush=True)
    g = P.mesh_2dvs(2048, 40.0 / 2048)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    st = Hm.stats
    lod = Hm.lod
    fr = P.friction_c1(g)
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod)
    out = {"n": g.n, "steps": steps, "matvec_ms": {}, "runs": {}}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_F32, "f32"), (H.PREC_MIXED, "mixed")):
        t = Hm.time_matvec(prec, 50)
        out["matvec_ms"][name] = round(t["total_ms"], 5)
    out["matvec_alg_bytes_f64"] = st["alg_bytes"][H.PREC_F64]
    mlp = L.load(os.path.join(ROOT, "hbi_amd", "data", "step_mlp_c1.npz"))
    hists = {}
    for name, prec, learned in (("f64", H.PREC_F64, False), ("mixed", H.PREC_MIXED, False),
                                ("mixed_learned", H.PREC_MIXED, True)):
        rk = H.RK(Hm, fr, lod, g.ds, eps_r=EPS_R, prec=prec)
        if learned:
            rk.set_controller(1, mlp, 0.9)
        rk.set_state(y0, 0.0, 1.0)
        rk.step(20)
        rk.set_state(y0, 0.0, 1.0)
        t0 = time.perf_counter()
        info, hist, _ = rk.step(steps, history=True)
        wall = time.perf_counter() - t0
        hists[name] = hist
        out["runs"][name] = {"steps_per_s": round(steps / wall, 1), "attempts": info["attempts"],
                             "rejected": info["rejected"], "t_end_s": info["t"],
                             "sim_time_per_wall_s": info["t"] / wall,
                             "events": len(_events(hist[:, 0], hist[:, 4])), "status": info["status"]}
        rk.close()
    e64 = _events(hists["f64"][:, 0], hists["f64"][:, 4])
    for name in ("mixed", "mixed_learned"):
        em = _events(hists[name][:, 0], hists[name][:, 4])
        k = min(len(e64), len(em))
        out["runs"][name]["onset_rel_vs_f64"] = (max(abs(a - b) / b for a, b in zip(em[:k], e64[:k]))
                                                 if k else None)
    t_common = min(hists["f64"][-1, 0], hists["mixed_learned"][-1, 0])
    a_p = L.attempts_to_time(hists["mixed"], t_common)
    a_l = L.attempts_to_time(hists["mixed_learned"], t_common)
    out["learned_attempt_ratio_to_common_time"] = a_p / a_l
    out["speedup_mixed_learned_vs_f64_sim_time"] = (out["runs"]["mixed_learned"]["sim_time_per_wall_s"]
                                                    / out["runs"]["f64"]["sim_time_per_wall_s"])
    Hm.close()
    return out


PHASES = ((300, "nucleation"), (1000, "first event"), (2000, "interseismic"), (3000, "second event"))


def controllers_by_phase(H, Hm, frs, lod, ys, mlp, warmup, steps):
    """The two controllers over the cycle: an rkqs ensemble from t = 0 is
    spun up to each of the PHASES' lockstep step counts (first event
    nucleating at ~300, rupturing at ~600-1200, interseismic by 2000, second
    event at 3000); from that state (per-realisation t and next step) each
    controller takes `warmup` untimed and `steps` timed lockstep steps:
    accepted steps/s, lockstep attempts and simulated time per wall-second."""
    from hbi_amd import learned as L
    S = H.Ensemble(Hm, frs, lod, hdiv=DS0, eps_r=EPS_R, prec=H.PREC_MIXED)
    S.set_state(ys, 0.0, 1.0)
    done, rows = 0, []
    for spin, name in PHASES:
        S.step(spin - done)
        done = spin
        st = S.get_state()
        t, h = S.get_clock()
        print("bench: controllers from the %s state (%d steps)" % (name, spin), file=sys.stderr, flush=True)
        row = {"phase": name, "spin_steps": spin, "t_median_s": float(np.median(t)),
               "log10_vmax_median": float(np.median(np.log10(np.abs(st["vel"]).max(axis=1))))}
        for key, learned in (("rkqs", False), ("learned", True)):
            E = H.Ensemble(Hm, frs, lod, hdiv=DS0, eps_r=EPS_R, prec=H.PREC_MIXED)
            if learned:
                E.set_controller(1, mlp, **L.C2_BOUNDS)
            E.set_state(st["y"], 0.0, 1.0)
            E.set_clock(t, h)
            if warmup:
                E.step(warmup)
            t1 = E.get_clock()[0]
            w0 = time.perf_counter()
            att = E.step(steps)
            wall = time.perf_counter() - w0
            es = E.get_state()
            row[key] = {"lockstep_attempts": int(att), "steps_per_s": round(NREAL * steps / wall, 1),
                        "sim_time_per_wall_s": float((es["t"] - t1).sum()) / wall,
                        "status_ok": bool((es["status"] == 0).all())}
            E.close()
        rows.append(row)
    S.close()
    mean = {k: float(np.mean([r[k]["steps_per_s"] for r in rows])) for k in ("rkqs", "learned")}
    return {"phases": rows, "mean_steps_per_s": mean,
            "note": "both controllers from the same spun-up states (rkqs spin-up from t = 0); the headline window "
                    "starts at t = 0 (the ensemble and ensemble_rkqs entries)"}


def ensemble_accuracy(H, Hm, frs, lod, ys, g, steps=1500):
    """The headline ensemble's precision (mixed: FP32 factors on the BF16
    matrix cores) against an FP64 ensemble of the same realisations, both
    under rkqs from t = 0: event onset times; and, replayed on the FP64 run's
    step grid (identical times), the slip-rate and probe-slip differences over
    the near-plate-rate (log10 Vmax < -8) and quasi-static (< -6) phases."""
    from hbi_amd import workloads as W
    print("bench: ensemble accuracy, FP64 and mixed ensembles of %d steps" % steps, file=sys.stderr, flush=True)
    inv = np.empty(g.n, dtype=np.int64)
    inv[lod] = np.arange(g.n)
    xs = np.linspace(-25.0, 25.0, 16)
    probes = inv[[int(np.argmin((g.xcol - x) ** 2 + (g.zcol + 15.0) ** 2)) for x in xs]].astype(np.int32)
    hf, sf = W.run_ensemble(Hm, frs, lod, ys, probes, H.PREC_F64, steps, DS0)
    hm, sm = W.run_ensemble(Hm, frs, lod, ys, probes, H.PREC_MIXED, steps, DS0)
    E = H.Ensemble(Hm, frs, lod, hdiv=DS0, eps_r=EPS_R, prec=H.PREC_MIXED)
    E.set_probes(probes)
    E.set_state(ys, 0.0, 1.0)
    E.set_step_grid(hf[:, :, 1])
    _, hr = E.step(steps, history=True)
    E.close()
    return {"steps": steps, "f64_ensemble_steps_per_s": sf["steps_per_s"], "onsets": W.ensemble_errors(hf, hm),
            "on_f64_grid_below_1e-8": W.grid_errors(hf, hr, -8.0), "on_f64_grid_below_1e-6": W.grid_errors(hf, hr, -6.0),
            "note": "mixed vs FP64 lockstep ensembles from the same states; onset = first log10 Vmax > -3; "
                    "grid errors: mixed replayed on the FP64 run's accepted step sizes"}


def cpu_part(nsteps: int):
    """CPU oracle (FP64, OpenMP over all host cores) on a bounded sample:
    the same fault and matrix settings, nsteps accepted single-run steps."""
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
    out = {"value": done / wall, "unit": "accepted RK steps/s (one realisation)", "cores": ncores, "kind": "port",
           "sample": "%d accepted RK45 steps (FP64 oracle, OpenMP) of one realisation of the 32768-element fault "
                     "after a %.1f s H-matrix build" % (done, t_build)}
    Ho = None
    # C1: the 2-D antiplane fault from the steady state at Vpl with the seeded nucleation bump
    g1 = P.mesh_2dvs(2048, 40.0 / 2048)
    fr1 = P.friction_c1(g1)
    H1 = O.HMatrix(g1, leaf=32, eps=1e-6, nthreads=ncores)
    x1 = np.random.default_rng(0).standard_normal(g1.n)
    H1.matvec(x1, nthreads=ncores)
    t0 = time.perf_counter()
    for _ in range(50):
        H1.matvec(x1, nthreads=ncores)
    mv1 = (time.perf_counter() - t0) / 50
    y1 = P.init_state(fr1, H1.lod)
    P.perturb_state(y1, g1, H1.lod)
    a1 = fr1.a[H1.lod]
    vel1 = 2 * fr1.vref * np.exp(-y1[0::3] / a1) * np.sinh(y1[1::3] / y1[2::3] / a1)
    t0 = time.perf_counter()
    d1, *_ = O.Sim(H1, fr1, H1.lod, nthreads=ncores).run(y1, 0.0, 1.0, EPS_R, C1_CPU_STEPS, np.zeros(g1.n), vel1,
                                                          np.asarray(g1.ds)[H1.lod])
    wall = time.perf_counter() - t0
    out["c1"] = {"steps_per_s": round(d1 / wall, 1), "matvec_ms": round(1e3 * mv1, 4), "cores": ncores,
                 "sample": "%d accepted steps of the 2048-element antiplane fault from the seeded steady "
                           "state (FP64 oracle, OpenMP products)" % d1}
    H1 = None
    # the fluid-driven cycle's oracle (numpy, C oracle products) on bp6a
    from oracle import fluid as F
    gf = P.mesh_2dp(P.BP6A_N, P.BP6A_DS0)
    Hf = O.HMatrix(gf, leaf=32, eps=1e-4, nthreads=ncores)
    fo, fa, fy, fpf = P.fluid_bp6a(gf, Hf.lod)
    ora = F.FvOracle(Hf, fo, fa)
    st = dict(y=fy.copy(), pf=fpf.copy(), disp=np.zeros(gf.n), t=0.0, h=1.0)
    for _ in range(20):
        ora.step(st)
    t0 = time.perf_counter()
    for _ in range(FLUID_CPU_STEPS):
        ora.step(st)
    wall = time.perf_counter() - t0
    out["fluid_bp6a"] = {"steps_per_s": round(FLUID_CPU_STEPS / wall, 2), "cores": ncores,
                         "sample": "%d accepted steps after 20 warm-up steps (numpy oracle, OpenMP products)"
                                   % FLUID_CPU_STEPS}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1500, help="timed lockstep steps of the ensemble")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--single-steps", type=int, default=2000)
    ap.add_argument("--cpu-steps", type=int, default=600)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one process, one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    r = gpu_part(args.steps, args.warmup, args.single_steps)
    ens = r["ensemble"]
    ens0 = r["ensemble_rkqs"]
    mmx = r["matmul64"]["mixed"]
    mvx = r["matvec"]["mixed"]
    # the single-vector product's longer launch: the second when it streams the dense blocks too
    if r["hmat"]["dense_launch"][2] == 2:
        mv_kernel = ("lowrank_kernel<64> (single-vector product, second launch: per row tile the dense and a2 "
                     "columns as one run, one x/z gather, mixed)")
        mv_gbps, mv_traffic = mvx["b_GBps"], TRAFFIC_LOWRANK_DENSE_MIXED
    else:
        mv_kernel = "stream_kernel<64> (single-vector product, first launch: stacked a1^T tiles + dense tiles, mixed)"
        mv_gbps, mv_traffic = mvx["a_GBps"], TRAFFIC_STREAM_DENSE_MIXED
    f64 = r["single"]["f64"]["steps_per_s"]
    line = {
        "metric": "RK steps/s & H-matvec GB/s (%HBM peak), mixed+learned vs FP64, 1e-5 rel err",
        "value": round(ens["steps_per_s"], 1),
        "unit": "accepted RK steps/s (sum over %d realisations)" % NREAL,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * ens["wall_s"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state / f32 H factors (split into 3 bf16 parts on the BF16 MFMA, FP64 sums of 64-column slices)",
        "data": "synthetic",
        "config": {"workload": "3dph 256x128 planar fault, N=32768, leaf 64, ACA eps 1e-5, RK45 eps_r 1e-6; "
                               "64 friction realisations in lockstep, one 64-column H product per RK stage, "
                               "FP32 factors on the BF16 matrix cores (3-part split), learned step-size proposals "
                               "(rkqs check), hipGraph attempts",
                   "n_elements": IMAX * JMAX, "realisations": NREAL},
        "roofline": {"bound": "hbm",
                     "kernel": "mr_bf16_kernel<4,true,2,2> (64-RHS row tiles on the BF16 matrix cores, split operands)",
                     "achieved": round(r["p2_alg_bytes"] / (mmx["p2_ms"] * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r["p2_alg_bytes"] / (mmx["p2_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC_MR_BF16_MIXED,
                     "alg_bytes": r["p2_alg_bytes"],
                     "useful_tflops": round(mmx["p2_TFs"], 2),
                     "useful_frac_of_bf16_peak": round(mmx["p2_TFs"] / BF16_MFMA_PEAK_TFS, 4),
                     "mfma_executed_frac_of_bf16_peak": round(6 * mmx["p2_TFs"] / BF16_MFMA_PEAK_TFS, 4),
                     "mfma_util_counter": MFMA_UTIL_MR_BF16,
                     "note": "arithmetic intensity 2*64 flop per 4-byte factor (32 flop/B) is below the BF16 ridge "
                             "(312 flop/B): HBM-bound; six BF16 MFMA terms per FP32 product"},
        "roofline_matvec": {"bound": "hbm", "kernel": mv_kernel,
                            "achieved": round(mv_gbps, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(mv_gbps / HBM_PEAK_GBS, 4), "traffic": mv_traffic,
                            "whole_product_frac": {k: round(v["GBps"] / HBM_PEAK_GBS, 4)
                                                   for k, v in r["matvec"].items()}},
        "fp64_single_steps_per_s": round(f64, 2),
        "mixed_single_steps_per_s": round(r["single"]["mixed"]["steps_per_s"], 2),
        "speedup_vs_fp64_single": round(ens["steps_per_s"] / f64, 2),
        "batched_rkqs_steps_per_s": round(ens0["steps_per_s"], 1),
        "sim_time_per_wall_s": {"learned": round(ens["sim_time_per_wall_s"], 1),
                                "rkqs": round(ens0["sim_time_per_wall_s"], 1)},
        "learned_vs_rkqs": r["learned_vs_rkqs"],
        "controllers_by_phase": r["controllers_by_phase"],
        "c1": r["c1"],
        "fluid3dp": r["fluid3dp"],
        "ensemble_rkqs": ens0,
        "accuracy_mixed_vs_f64": r["accuracy_mixed_vs_f64"],
        "accuracy_ensemble_vs_f64": r["accuracy_ensemble_vs_f64"],
        "ensemble": ens,
        "matvec": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                   for k, v in r["matvec"].items()},
        "matmul64": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in r["matmul64"].items()},
        "matmul64_plan": r["matmul64_plan"],
        "mfma_loop_tfs": {"f64": r["mfma_f64_loop_tfs"], "f32": r["mfma_f32_loop_tfs"]},
        "single": r["single"],
        "fluid_bp6a": r["fluid_bp6a"],
        "build_s": r["build_s"],
        "hmat": r["hmat"],
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_part(args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
