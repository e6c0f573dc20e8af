"""Train the step-size MLP and write it as an .npz.

    python -m hbi_amd.train_step_mlp [--problem c1|c2] [--out PATH]
        [--seeds N] [--steps K] [--rounds R] [--epochs E]

c1: the 2-D antiplane fault (2048 elements, 40 km, leaf 32, ACA eps 1e-6),
    `seeds` perturbed realisations (default 44 x 8000 steps) advanced as one
    lockstep ensemble, FP64;
c2: the 32768-element 3-D planar fault (256 x 128, leaf 64, ACA eps 1e-5,
    mixed precision), realisations of its lognormal friction advanced 64 at
    a time as lockstep ensembles: seeds 1-64 for `steps` steps (default
    5300: the first event cycle and the second event), seeds 201-264 and
    301-364 (clear of the held-out 101-164) for 1750 steps each (the first
    event); --seeds N takes the first N of these 192;
the defaults give 1.06e6 (c1) and 1.69e6 (c2) (features, step) pairs over
three rounds;
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


def c2_setup(seeds, first_seed=1, hmat=None):
    """The 3-D fault, its H-matrix (built, or hmat = (g, Hm) reused) and
    `seeds` friction realisations with perturbed initial states (seeds
    first_seed ...)."""
    if hmat is None:
        g = P.mesh_3dph(C2["imax"], C2["jmax"], C2["ds0"], 90.0)
        Hm = H.HMatrix(g, leaf=C2["leaf"], eps=C2["eps"], bf16_margin=1.0)
    else:
        g, Hm = hmat
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


C2_TRAIN_FIRST_SEEDS = (1, 201, 301)
C2_SHORT_STEPS = 1750


def train_c2(seeds=192, steps=5300, rounds=3, epochs=40, log=print):
    g, Hm, lod, frs, ys = c2_setup(min(seeds, 64))
    batches = [(frs, ys, steps)]
    for f0 in C2_TRAIN_FIRST_SEEDS[1:]:
        k = min(64, seeds - 64 * len(batches))
        if k <= 0:
            break
        _, _, _, fb, yb = c2_setup(k, first_seed=f0, hmat=(g, Hm))
        batches.append((fb, yb, min(steps, C2_SHORT_STEPS)))
    traces = []
    m = None
    t0 = time.time()
    for r in range(rounds):
        for fb, yb, k in batches:
            traces += L.ensemble_trace_run(Hm, g, fb, lod, yb, k, hdiv=C2["ds0"], prec=H.PREC_MIXED, mlp=m)
        X, T = L.dataset(traces)
        m, loss = L.fit(X, T, epochs=epochs)
        log("round %d: %d samples, loss %.3e (target variance %.3e), %.1f s" % (r + 1, len(X), loss[-1], T.var(),
                                                                              time.time() - t0))
    return m


def train(seeds=44, steps=8000, rounds=3, epochs=40, n=2048, log=print):
    g = P.mesh_2dvs(n, 40.0 / n)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    frs = [L.c1_realisation(g, s) for s in range(1, seeds + 1)]
    ys = np.array([L.initial_state(g, fr, lod, s) for s, fr in zip(range(1, seeds + 1), frs)])
    traces = []
    m = None
    t0 = time.time()
    for r in range(rounds):
        for b in range(0, seeds, 64):
            traces += L.ensemble_trace_run(Hm, g, frs[b:b + 64], lod, ys[b:b + 64], steps, mlp=m)
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
        m = train(a.seeds or 44, a.steps or 8000, a.rounds, a.epochs, log=log)
    else:
        out = a.out or C2_OUT
        m = train_c2(a.seeds or 192, a.steps or 5300, a.rounds, a.epochs, log=log)
    a.out = out
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    L.save(m, a.out)
    print("wrote", a.out, flush=True)


if __name__ == "__main__":
    main()
