"""Train the step-size MLP and write it as an .npz.

    python -m hbi_amd.train_step_mlp [--problem c1|c2] [--out PATH]
        [--seeds N] [--steps K] [--rounds R] [--epochs E]

c1: the 2-D antiplane fault (2048 elements, 40 km, leaf 32, ACA eps 1e-6),
    single runs of `seeds` perturbed realisations (default 8 x 8000 steps);
c2: the 32768-element 3-D planar fault (256 x 128, leaf 64, ACA eps 1e-5,
    FP32 factors), `seeds` realisations of its lognormal friction (default
    64) advanced as one lockstep ensemble (default 1500 steps);
RK eps_r 1e-6 for both.  Round 1 records traces of accepted steps under the
reference's controller (rkqs); every further round adds traces recorded under
the model of the previous round (the states the learned controller itself
visits) and retrains on all of them.  Runs on the GPU: traces come from the
device RK graphs, training from the device kernels.
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import learned as L
from . import lib as H
from . import problems as P

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_OUT = os.path.join(DATA, "step_mlp_c1.npz")
C2_OUT = os.path.join(DATA, "step_mlp_c2.npz")
C2 = dict(imax=256, jmax=128, ds0=0.390625, leaf=64, eps=1e-5)


def c2_setup(seeds, first_seed=1):
    """The 3-D fault, its H-matrix and `seeds` friction realisations with
    perturbed initial states (seeds first_seed ...)."""
    g = P.mesh_3dph(C2["imax"], C2["jmax"], C2["ds0"], 90.0)
    Hm = H.HMatrix(g, leaf=C2["leaf"], eps=C2["eps"], bf16_margin=1.0)
    lod = Hm.lod
    frs, ys = [], []
    for s in range(first_seed, first_seed + seeds):
        fr = P.friction_c2(g, C2["imax"], C2["jmax"], C2["ds0"], seed=s)
        rng = np.random.default_rng(s + 1000)
        y = P.init_state(fr, lod)
        P.perturb_state(y, g, lod, amp=float(rng.uniform(0.01, 0.03)), depth=float(rng.uniform(10.0, 20.0)),
                        width=3.0)
        frs.append(fr)
        ys.append(y)
    return g, Hm, lod, frs, np.array(ys)


def train_c2(seeds=64, steps=1500, rounds=3, epochs=40, log=print):
    g, Hm, lod, frs, ys = c2_setup(seeds)
    traces = []
    m = None
    t0 = time.time()
    for r in range(rounds):
        traces += L.ensemble_trace_run(Hm, g, frs, lod, ys, steps, hdiv=C2["ds0"], prec=H.PREC_MIXED, mlp=m)
        X, T = L.dataset(traces)
        m, loss = L.fit(X, T, epochs=epochs)
        log("round %d: %d samples, loss %.3e (target variance %.3e), %.1f s" % (r + 1, len(X), loss[-1], T.var(),
                                                                              time.time() - t0))
    return m


def train(seeds=8, steps=8000, rounds=3, epochs=40, n=2048, log=print):
    g = P.mesh_2dvs(n, 40.0 / n)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    setups = []
    for s in range(1, seeds + 1):
        fr = L.c1_realisation(g, s)
        setups.append((fr, L.initial_state(g, fr, lod, s)))
    traces = []
    m = None
    t0 = time.time()
    for r in range(rounds):
        traces += [L.trace_run(Hm, g, fr, lod, y0, steps, mlp=m) for fr, y0 in setups]
        X, T = L.dataset(traces)
        m, loss = L.fit(X, T, epochs=epochs)
        log("round %d: %d samples, loss %.3e (target variance %.3e), %.1f s" % (r + 1, len(X), loss[-1], T.var(),
                                                                              time.time() - t0))
    return m


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problem", choices=("c1", "c2"), default="c1")
    ap.add_argument("--out", default=None)
    ap.add_argument("--seeds", type=int, default=None)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=40)
    a = ap.parse_args()
    log = lambda s: print(s, flush=True)
    if a.problem == "c1":
        out = a.out or DEFAULT_OUT
        m = train(a.seeds or 8, a.steps or 8000, a.rounds, a.epochs, log=log)
    else:
        out = a.out or C2_OUT
        m = train_c2(a.seeds or 64, a.steps or 1500, a.rounds, a.epochs, log=log)
    a.out = out
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    L.save(m, a.out)
    print("wrote", a.out, flush=True)


if __name__ == "__main__":
    main()
