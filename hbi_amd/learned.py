"""Learned step-size proposals for the device RK integrator.

Data: traces of accepted steps recorded on the device under the reference's
controller (rkqs, main_LH.f90:2309-2357) over perturbed realisations of the
2-D antiplane fault.  Each accepted step n gives the 12 features the RK graph
computes (log10 max / median / p99 slip rate, their changes since the last
accepted step, log10 h and errmax of this and the previous step, the
rejection flag, f11 = log10 of the error-optimal step h e^(-1/5)).  The
target is the change of that optimal step to the next accepted step,
f11[n+1] - f11[n]; the device then proposes safety 10^(f11 + mlp(features)).
The embedded error estimate still judges every step; rejected attempts
shrink h by the rkqs rule.
"""
from __future__ import annotations

import numpy as np

from . import lib as H
from . import problems as P

N_FEAT = 12

# Proposal bounds of the 3-D ensemble controller: safety 0.9 and the learned
# step within [1/8, 1] x the rkqs proposal, i.e. the network may shorten the
# reference's proposal (to step short of a predicted rejection) but never
# lengthen it.  On the 32768-element fault (64 realisations, mixed precision)
# this never loses to rkqs: equal attempts in the 2x-growth start-up from
# t = 0, where rkqs never rejects, and fewer rejections from every spun-up
# state; an upper bound of 8 x rkqs buys longer steps that are rejected more
# often (start-up: 25 lockstep attempts for 20 steps against 20).
C2_BOUNDS = dict(safety=0.9, ratio_lo=0.125, ratio_hi=1.0)


def c1_realisation(g, seed: int):
    """Friction of the 2-D antiplane fault with lognormal perturbations of Dc
    and b (sigma_log 0.1, per element, smoothed over 8 elements) from a seed."""
    rng = np.random.default_rng(seed)
    fr = P.friction_c1(g)
    k = np.ones(8) / 8.0
    for name in ("dc", "b"):
        z = np.convolve(rng.standard_normal(g.n), k, mode="same") * np.sqrt(8.0)
        setattr(fr, name, getattr(fr, name) * np.exp(0.1 * z))
    return fr


def initial_state(g, fr, lod, seed: int):
    rng = np.random.default_rng(seed + 1000)
    y = P.init_state(fr, lod)
    P.perturb_state(y, g, lod, amp=float(rng.uniform(0.01, 0.03)), depth=float(rng.uniform(5.0, 12.0)))
    return y


def trace_run(Hm, g, fr, lod, y0, nsteps: int, eps_r=1e-6, prec=H.PREC_F64, mlp=None, safety=0.9):
    """Features and (h, errmax) of nsteps accepted steps under rkqs, or under
    the learned controller when mlp is given (data for a second round of
    training on the states the learned controller itself visits)."""
    rk = H.RK(Hm, fr, lod, g.ds, eps_r=eps_r, prec=prec)
    if mlp is not None:
        rk.set_controller(1, mlp, safety)
    rk.set_trace(nsteps)
    rk.set_state(y0, 0.0, 1.0)
    rk.step(nsteps)
    feat, raw = rk.get_trace()
    rk.close()
    return feat, raw


def ensemble_trace_run(Hm, g, frs, lod, ys, nsteps: int, hdiv=0.0, eps_r=1e-6, prec=H.PREC_F64, mlp=None,
                       safety=0.9):
    """Traces of nsteps accepted steps of up to 64 realisations advanced in
    lockstep (one 64-column H product per stage), under rkqs or the learned
    controller; one (features, raw) pair per realisation."""
    E = H.Ensemble(Hm, frs, lod, hdiv=hdiv, eps_r=eps_r, prec=prec)
    if mlp is not None:
        E.set_controller(1, mlp, safety)
    E.set_trace(nsteps)
    E.set_state(np.asarray(ys), 0.0, 1.0)
    E.step(nsteps)
    out = E.get_trace()
    E.close()
    return out


def dataset(traces):
    """(X, T) from a list of (features, raw) traces: X = features of step n,
    T = f11[n+1] - f11[n]."""
    Xs, Ts = [], []
    for feat, _ in traces:
        if len(feat) < 2:
            continue
        Xs.append(feat[:-1])
        Ts.append(feat[1:, 11] - feat[:-1, 11])
    X = np.concatenate(Xs).astype(np.float32)
    T = np.concatenate(Ts).astype(np.float32)
    ok = np.isfinite(X).all(axis=1) & np.isfinite(T)
    return X[ok], T[ok]


def fit(X, T, epochs=30, batch=1024, lr=1e-3, seed=0):
    """Train an MLP on the device; returns (model, per-epoch loss)."""
    mu = X.mean(axis=0)
    sd = X.std(axis=0)
    sd = np.where(sd > 1e-6, sd, 1.0)
    m = H.MLP(seed)
    m.set_params(mu=mu, sd=sd)
    loss = m.train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return m, loss


def attempts_to_time(hist, t_end):
    """RK attempts (accepted + rejected) until simulated time t_end, from
    history rows (t, dt, errmax, nrjct, ...)."""
    k = int(np.searchsorted(hist[:, 0], t_end))
    k = min(k, len(hist) - 1)
    return int(k + 1 + hist[: k + 1, 3].sum())


def save(m: H.MLP, path: str):
    p, mu, sd = m.get_params()
    np.savez(path, params=p, mu=mu, sd=sd)


def load(path: str, seed=0) -> H.MLP:
    d = np.load(path)
    m = H.MLP(seed)
    m.set_params(d["params"], d["mu"], d["sd"])
    return m
