"""Train the step-size MLP on the 2-D antiplane fault (2048 elements, 40 km,
leaf 32, ACA eps 1e-6, RK eps_r 1e-6) and write it as an .npz.

    python -m hbi_amd.train_step_mlp [--out hbi_amd/data/step_mlp_c1.npz]
        [--seeds 8] [--steps 8000] [--rounds 3] [--epochs 40]

Round 1 records traces of accepted steps under the reference's controller
(rkqs) on `seeds` perturbed realisations; every further round adds traces
recorded under the model of the previous round (the states the learned
controller itself visits) and retrains on all of them.  Runs on the GPU:
traces come from the device RK graph, training from the device kernels.
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import learned as L
from . import lib as H
from . import problems as P

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "step_mlp_c1.npz")


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
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=40)
    a = ap.parse_args()
    m = train(a.seeds, a.steps, a.rounds, a.epochs, log=lambda s: print(s, flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    L.save(m, a.out)
    print("wrote", a.out, flush=True)


if __name__ == "__main__":
    main()
