"""Kernel timings of the H products on the 32768-element planar fault, for
profilers (rocprofv3 --kernel-trace / --pmc) and for the roofline figures.

    python -m hbi_amd.profile_kernels [--iters N] [--copy-mb M]

Builds the 256 x 128 half-space fault's lattice H-matrix (leaf 64, ACA eps
1e-5) and times, with HIP events on the library stream, the single-vector
product of every precision and the 64-column product (FP64 and mixed) over
`iters` calls each; the same kernels and work items the RK graphs launch.
With --copy-mb > 0 it also runs a plain device copy of that many MiB (torch)
`iters` times, the reference for scaling the profiler's byte counters.
Prints one JSON line.
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
    a = ap.parse_args()
    g = P.mesh_3dph(IMAX, JMAX, DS0, 90.0)
    Hm = H.HMatrix(g, leaf=LEAF, eps=EPS_H, bf16_margin=1.0)
    st = Hm.stats
    lv = Hm.leaves
    dense, lr = lv[:, 0] != 1, lv[:, 0] == 1
    out = {"n": g.n, "iters": a.iters,
           "p1_entries": int((lv[lr, 4] * lv[lr, 5]).sum()),
           "p2_entries": int((lv[dense, 2] * lv[dense, 4]).sum() + (lv[lr, 2] * lv[lr, 5]).sum()),
           "alg_bytes_p1": st["alg_bytes_p1"], "alg_bytes_p2": st["alg_bytes_p2"], "matvec": {}, "matmul64": {}}
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_F32, "f32"), (H.PREC_MIXED, "mixed")):
        out["matvec"][name] = Hm.time_matvec(prec, a.iters)
    for prec, name in ((H.PREC_F64, "f64"), (H.PREC_MIXED, "mixed")):
        out["matmul64"][name] = Hm.time_matmul(prec, a.iters)
    if a.copy_mb > 0:
        import torch
        n = a.copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(a.iters):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        out["copy"] = {"bytes_read": n * 4, "bytes_written": n * 4, "ms": e0.elapsed_time(e1) / a.iters}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
