"""Kernel timings of the H products on the 32768-element planar fault, for
profilers (rocprofv3 --kernel-trace / --pmc) and for the roofline figures.

    python -m hbi_amd.profile_kernels [--iters N] [--copy-mb M] [--dense-launch 0|1|2]

Builds the 256 x 128 half-space fault's lattice H-matrix (leaf 64, ACA eps
1e-5) and times, with HIP events on the library stream, the single-vector
product of every precision and the 64-column product (FP64 on the FP64
matrix cores, FP32 on the FP32 cores, mixed on the BF16 cores) over
`iters` calls each; the same kernels and work items the RK graphs launch.
With --copy-mb > 0 it also runs a plain device-to-device copy of that many
MiB (HIP runtime blit kernel) `iters` times, the reference for scaling the
profiler's byte counters.  With --ensemble K it instead runs the bench's
ensemble (64 friction realisations, learned step size, mixed precision) for 20
warm-up and K timed lockstep steps, for a kernel trace of the attempt; with
--fluid K the fluid-driven cycle of the bp6a example (4000 elements) for 20
warm-up and K timed steps; with --fluid3dp K main_fv3dp's cycle on the 200 x
200 grid of examples/3dinjection.in likewise.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import lib as H
from . import problems as P

IMAX, JMAX, DS0, LEAF, EPS_H = 256, 128, 0.390625, 64, 1e-5


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--copy-mb", type=int, default=0)
    ap.add_argument("--ensemble", type=int, default=0)
    ap.add_argument("--ensemble-warmup", type=int, default=20,
                    help="lockstep steps from t = 0 before the timed ones (the bench's window: 5 + 20)")
    ap.add_argument("--fluid", type=int, default=0)
    ap.add_argument("--fluid3dp", type=int, default=0)
    ap.add_argument("--dense-launch", type=int, default=0,
                    help="single-vector product layout: 0 autotuned, 1 dense blocks in the first launch, 2 in the second")
    a = ap.parse_args()
    if a.fluid > 0:
        print(json.dumps(fluid(a.fluid)), flush=True)
        return
    if a.fluid3dp > 0:
        print(json.dumps(fluid3dp(a.fluid3dp)), flush=True)
        return
    g = P.mesh_3dph(IMAX, JMAX, DS0, 90.0)
    Hm = H.HMatrix(g, leaf=LEAF, eps=EPS_H, bf16_margin=1.0, dense_launch=a.dense_launch)
    if a.ensemble > 0:
        print(json.dumps(ensemble(g, Hm, a.ensemble, a.ensemble_warmup)), flush=True)
        return
    st = Hm.stats
    lv = Hm.leaves
    dense, lr = lv[:, 0] != 1, lv[:, 0] == 1
    out = {"n": g.n, "iters": a.iters,
           "p1_entries": int((lv[lr, 4] * lv[lr, 5]).sum()),
           "p2_entries": int((lv[dense, 2] * lv[dense, 4]).sum() + (lv[lr, 2] * lv[lr, 5]).sum()),
           "alg_bytes_p1": st["alg_bytes_p1"], "alg_bytes_p2": st["alg_bytes_p2"], "dense_launch": st["dense_launch"],
           "matvec": {}, "matmul64": {}}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_F32, "f32"), (H.PREC_MIXED, "mixed")):
        out["matvec"][name] = Hm.time_matvec(prec, a.iters)
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_F32, "f32"), (H.PREC_MIXED, "mixed")):
        out["matmul64"][name] = Hm.time_matmul(prec, a.iters)
        out["matmul64"][name]["plan"] = Hm.matmul_plan(prec)
    if a.copy_mb > 0:
        out["copy"] = device_copy(a.copy_mb * (1 << 20), a.iters)
    print(json.dumps(out), flush=True)


def ensemble(g, Hm, steps: int, warmup: int = 20) -> dict:
    """The bench's learned-step ensemble (its proposal bounds): `warmup` steps from t = 0, then
    `steps` lockstep steps; attempts counts the graph launches of the timed call, including those
    launched past the last realisation's target (gated off: each kernel returns at once)."""
    import os
    from . import learned as L
    lod = Hm.lod
    frs = [P.friction_c2(g, IMAX, JMAX, DS0, seed=1 + r) for r in range(64)]
    ys = []
    for f in frs:
        y = P.init_state(f, lod)
        P.perturb_state(y, g, lod, amp=0.02, depth=15.0, width=3.0)
        ys.append(y)
    E = H.Ensemble(Hm, frs, lod, hdiv=DS0, eps_r=1e-6, prec=H.PREC_MIXED)
    E.set_controller(1, L.load(os.path.join(os.path.dirname(__file__), "data", "step_mlp_c2.npz")), **L.C2_BOUNDS)
    E.set_state(np.array(ys), 0.0, 1.0)
    E.step(warmup)
    att = E.step(steps)
    out = {"steps": steps, "attempts": int(att), "device_ms": E.last_ms(), "plan": Hm.matmul_plan(H.PREC_MIXED)}
    E.close()
    return out


def fluid(steps: int) -> dict:
    """bp6a (main_fv): 20 warm-up and `steps` timed accepted steps."""
    g = P.mesh_2dp(P.BP6A_N, P.BP6A_DS0)
    Hm = H.HMatrix(g, leaf=32, eps=1e-4)
    opts, arr, y, pf = P.fluid_bp6a(g, Hm.lod)
    fv = H.FV(Hm, opts, arr)
    fv.set_state(y, pf, t=0.0, dt=1.0)
    fv.step(20)
    info, hist = fv.step(steps, history=True)
    return {"steps": steps, "device_ms": fv.last_ms(), "ms_per_step": fv.last_ms() / max(1, len(hist)),
            "cg_iter_mean": float(hist[:, 9].mean()), "attempts": info["attempts"], "t_end_s": info["t"]}


def fluid3dp(steps: int) -> dict:
    """main_fv3dp on the 200 x 200 3dp fault (bench.py's fluid3dp): 20
    warm-up and `steps` timed accepted steps."""
    g0, _, _, _, _ = P.fluid3dp_injection(200, 200, 0.01)
    Hm = H.HMatrix(g0, leaf=32, eps=1e-4)
    g, opts, arr, y, pf = P.fluid3dp_injection(200, 200, 0.01, bc="Neumann", q0=3e-4, lod=Hm.lod)
    fv = H.FV(Hm, opts, arr)
    fv.set_state(y, pf, t=0.0, dt=1.0)
    fv.step(20)
    info, hist = fv.step(steps, history=True)
    return {"steps": steps, "device_ms": fv.last_ms(), "ms_per_step": fv.last_ms() / max(1, len(hist)),
            "cg_iter_mean": float(hist[:, 9].mean()), "attempts": info["attempts"], "t_end_s": info["t"]}


def device_copy(nbytes: int, iters: int) -> dict:
    """iters device-to-device copies of nbytes through the HIP runtime (its
    blit kernel), timed with HIP events: the reference for scaling byte counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    vp = C.c_void_p

    def ok(e):
        if e != 0:
            raise RuntimeError("HIP error %d" % e)
    src, dst, e0, e1 = vp(), vp(), vp(), vp()
    ok(hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)))
    ok(hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)))
    ok(hip.hipMemset(src, 1, C.c_size_t(nbytes)))
    ok(hip.hipEventCreate(C.byref(e0)))
    ok(hip.hipEventCreate(C.byref(e1)))
    kind_d2d = 3
    ok(hip.hipMemcpy(dst, src, C.c_size_t(nbytes), kind_d2d))
    ok(hip.hipDeviceSynchronize())
    ok(hip.hipEventRecord(e0, None))
    for _ in range(iters):
        ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), kind_d2d, None))
    ok(hip.hipEventRecord(e1, None))
    ok(hip.hipEventSynchronize(e1))
    ms = C.c_float()
    ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1))
    for b in (src, dst):
        ok(hip.hipFree(b))
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms.value / iters}


if __name__ == "__main__":
    main()
