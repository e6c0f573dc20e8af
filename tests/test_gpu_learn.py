"""GPU: the step-size MLP's forward, gradient and Adam kernels against the
CPU oracle (oracle/mlp_oracle.py, FP64 with the same BF16 weight rounding),
its training, and the step features the RK graph computes.

Tolerances: forward 2e-5 x max|y| (FP32 accumulation over 128 terms);
gradient 1e-4 relative in norm per parameter block; Adam updates to 1e-6
where the gradient is not negligible."""
import numpy as np
import pytest

from hbi_amd import lib as H
from hbi_amd import learned as L
from hbi_amd import problems as P
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _model(seed=1):
    m = H.MLP(seed)
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal(12).astype(np.float32)
    sd = rng.uniform(0.5, 2.0, 12).astype(np.float32)
    p, _, _ = m.get_params()
    p = p.copy()
    p[M.OFF_B1:M.OFF_W2] = 0.1 * rng.standard_normal(128)
    p[M.OFF_B2:M.OFF_W3] = 0.1 * rng.standard_normal(128)
    p[M.OFF_B3] = 0.05
    m.set_params(p, mu, sd)
    return m


def _data(n, seed=2):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, 12)) * 1.5).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.2 * X[:, 11]).astype(np.float32)
    return X, T


def test_gpu_mlp_forward_matches_oracle():
    m = _model()
    p, mu, sd = m.get_params()
    X, _ = _data(333)
    y = m.forward(X)
    yo, _ = M.forward(p, mu, sd, X)
    assert np.abs(y - yo).max() <= 2e-5 * np.abs(yo).max()


def test_gpu_mlp_gradient_matches_oracle():
    m = _model(3)
    p, mu, sd = m.get_params()
    X, T = _data(257, 4)     # not a multiple of the 16-sample groups
    Lg, g = m.loss_grad(X, T)
    Lo, go = M.loss_grad(p, mu, sd, X, T)
    assert abs(Lg - Lo) <= 1e-5 * Lo
    blocks = [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
              (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAM)]
    for a, b in blocks:
        assert np.linalg.norm(g[a:b] - go[a:b]) <= 1e-4 * np.linalg.norm(go[a:b]) + 1e-9, (a, b)


def test_gpu_mlp_adam_steps_match_oracle():
    m = _model(5)
    p0, mu, sd = m.get_params()
    X, T = _data(200, 6)
    lr = 1e-3
    p = p0.astype(np.float64)
    m1 = np.zeros(M.N_PARAM)
    m2 = np.zeros(M.N_PARAM)
    for t in (1, 2):
        _, g = M.loss_grad(p.astype(np.float32), mu, sd, X, T)
        p, m1, m2 = M.adam(p, m1, m2, g, t, lr)
        m.train(X, T, epochs=1, batch=len(X), lr=lr)       # one full-batch step
        pg, _, _ = m.get_params()
        big = np.abs(g) > 1e-5
        assert np.abs(pg[big] - p[big]).max() <= 1e-6, t
        assert np.abs(pg - p).max() <= 2.5 * lr          # the rest moves by at most an Adam step


def test_gpu_mlp_training_reduces_loss():
    X, T = _data(20000, 7)
    m = H.MLP(0)
    m.set_params(mu=X.mean(0), sd=X.std(0))
    loss = m.train(X, T, epochs=30, batch=512, lr=2e-3, seed=1)
    print("loss", loss[0], loss[-1], flush=True)
    assert loss[-1] < 0.1 * loss[0]
    Xv, Tv = _data(2000, 8)
    assert np.mean((m.forward(Xv) - Tv) ** 2) < 0.2 * np.var(Tv)


def test_gpu_step_features_of_the_rk_graph():
    g = P.mesh_2dvs(512, 40.0 / 512)
    fr = P.friction_c1(g)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    y0 = P.init_state(fr, lod)
    P.perturb_state(y0, g, lod)
    rk = H.RK(Hm, fr, lod, g.ds, eps_r=1e-6)
    rk.set_trace(40)
    rk.set_state(y0, 0.0, 1.0)
    info, hist, _ = rk.step(40, history=True)
    feat, raw = rk.get_trace()
    assert feat.shape == (40, 12)
    assert np.allclose(raw[:, 0], hist[:, 1]) and np.allclose(raw[:, 1], hist[:, 2])
    f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    assert np.allclose(feat[:, 6], f32(np.log10(raw[:, 0])), rtol=0, atol=1e-6)
    assert np.allclose(feat[:, 7], f32(np.log10(raw[:, 1])), rtol=0, atol=1e-6)
    assert np.allclose(feat[:, 11], f32(np.log10(raw[:, 0]) - 0.2 * np.log10(raw[:, 1])), atol=1e-6)
    assert np.allclose(feat[1:, 9], feat[:-1, 6], atol=1e-6) and np.allclose(feat[1:, 10], feat[:-1, 7], atol=1e-6)
    assert np.allclose(feat[1:, 3:6], f32(feat[1:, 0:3] - feat[:-1, 0:3]), atol=1e-5)
    assert np.array_equal(feat[:, 8] > 0, hist[:, 3] > 0)
    # quantiles of the last accepted step against the final slip rates (bins of 24/1024 in log10 V)
    _, _, vel, _ = rk.get_state()
    lv = np.sort(np.log10(np.abs(vel)))
    n = len(lv)
    w = 24.0 / 1024
    bin_center = lambda v: -20.0 + (np.floor((v + 20.0) / w) + 0.5) * w
    for col, k in ((0, n - 1), (1, (n + 1) // 2 - 1), (2, int(np.ceil(0.99 * n)) - 1)):
        assert abs(feat[-1, col] - bin_center(lv[k])) <= w * 1.01, col


def test_gpu_training_pipeline_short():
    """Traces -> dataset -> device training -> learned controller in the RK
    graph, on a short run: the model learns (loss below the target variance)
    and the learned run reaches the same event onset as the rkqs run."""
    m = __import__("hbi_amd.train_step_mlp", fromlist=["train"]).train(seeds=2, steps=1500, rounds=1, epochs=10,
                                                                       n=1024)
    g = P.mesh_2dvs(1024, 40.0 / 1024)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    fr = L.c1_realisation(g, 5)
    y0 = L.initial_state(g, fr, lod, 5)
    runs = []
    for mode in (0, 1):
        rk = H.RK(Hm, fr, lod, g.ds, eps_r=1e-6)
        if mode:
            rk.set_controller(1, m, 0.9)
        rk.set_state(y0, 0.0, 1.0)
        info, hist, _ = rk.step(1500, history=True)
        assert info["accepted"] == 1500 and info["status"] == 0
        runs.append(hist)
    on = [h[np.argmax(h[:, 4] > -3.0), 0] for h in runs]
    assert all(h[:, 4].max() > -3.0 for h in runs)
    assert abs(on[1] - on[0]) <= 1e-5 * on[0]


def test_gpu_learned_step_gain_on_held_out_seeds():
    """The committed model (python -m hbi_amd.train_step_mlp: 44 seeds x 8000
    steps, 3 rounds, 1.06e6 pairs) against rkqs on three realisations it was
    not trained on, 2048 elements, 8000 rkqs steps (several seismic cycles):
    RK attempts to reach the same simulated time (measured 1.07, 1.10, 1.10
    times fewer), event onset times."""
    from hbi_amd import train_step_mlp as TS
    m = L.load(TS.DEFAULT_OUT)
    g = P.mesh_2dvs(2048, 40.0 / 2048)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    ratios, rej = [], [0, 0]
    for s in (101, 102, 103):
        fr = L.c1_realisation(g, s)
        y0 = L.initial_state(g, fr, lod, s)
        rk = H.RK(Hm, fr, lod, g.ds, eps_r=1e-6)
        rk.set_state(y0, 0.0, 1.0)
        _, hp, _ = rk.step(8000, history=True)
        t_end = hp[-1, 0]
        rk2 = H.RK(Hm, fr, lod, g.ds, eps_r=1e-6)
        rk2.set_controller(1, m, 0.9)
        rk2.set_state(y0, 0.0, 1.0)
        hs = []
        while not hs or (hs[-1][-1, 0] < t_end and len(hs) < 20):
            hs.append(rk2.step(1000, history=True)[1])
        hl = np.concatenate(hs)
        assert hl[-1, 0] >= t_end
        a_pi, a_l = L.attempts_to_time(hp, t_end), L.attempts_to_time(hl, t_end)
        ratios.append(a_pi / a_l)
        rej[0] += int(hp[:, 3].sum())
        rej[1] += int(hl[hl[:, 0] <= t_end, 3].sum())
        on_p = hp[np.argmax(hp[:, 4] > -3.0), 0]
        on_l = hl[np.argmax(hl[:, 4] > -3.0), 0]
        print("seed %d: rkqs %d attempts, learned %d (x%.3f); onset %.9e vs %.9e" % (s, a_pi, a_l, a_pi / a_l, on_p,
                                                                                  on_l), flush=True)
        assert abs(on_l - on_p) <= 1e-5 * on_p
    print("rejections: rkqs %d, learned %d" % tuple(rej), flush=True)
    assert np.mean(ratios) >= 1.05 and min(ratios) >= 1.0
    assert rej[1] <= 0.6 * rej[0]


def test_gpu_ensemble_traces_and_learned_steps_match_single_runs():
    """Features recorded by the ensemble graph equal those of single runs of
    the same realisations while the step grids coincide (first 10 steps), and
    the learned controller takes the same steps in both."""
    g = P.mesh_2dvs(1024, 40.0 / 1024)
    Hm = H.HMatrix(g, leaf=32, eps=1e-6)
    lod = Hm.lod
    frs = [L.c1_realisation(g, s) for s in (11, 12, 13)]
    ys = np.array([L.initial_state(g, f, lod, s) for f, s in zip(frs, (11, 12, 13))])
    from hbi_amd import train_step_mlp as TS
    m = L.load(TS.DEFAULT_OUT)
    w = 24.0 / 1024
    for mlp in (None, m):
        et = L.ensemble_trace_run(Hm, g, frs, lod, ys, 10, mlp=mlp)
        for r, f in enumerate(frs):
            st = L.trace_run(Hm, g, f, lod, ys[r], 10, mlp=mlp)
            fe, re_ = et[r]
            fs, rs = st
            assert fe.shape == fs.shape == (10, 12)
            assert np.allclose(re_[:, 0], rs[:, 0], rtol=1e-9), (mlp is None, r)
            assert np.abs(fe[:, :6] - fs[:, :6]).max() <= 1.01 * w
            assert np.abs(fe[:, 6:] - fs[:, 6:]).max() <= 1e-5


def test_gpu_learned_step_gain_on_the_3d_fault_ensemble():
    """The committed 3-D model (python -m hbi_amd.train_step_mlp --problem c2:
    192 realisations x 1750 lockstep steps, 3 rounds, 1.01e6 pairs) on 64
    realisations it was not trained on (seeds 101-164), mixed precision, in one ensemble: RK
    attempts of each realisation to reach the simulated time of 1500 rkqs
    steps, and the lockstep attempts until every realisation got there."""
    from hbi_amd import train_step_mlp as TS
    m = L.load(TS.C2_OUT)
    g, Hm, lod, frs, ys = TS.c2_setup(64, first_seed=101)
    K = 1500
    E = H.Ensemble(Hm, frs, lod, hdiv=TS.C2["ds0"], eps_r=1e-6, prec=H.PREC_MIXED)
    E.set_state(ys, 0.0, 1.0)
    _, h0 = E.step(K, history=True)
    E.close()
    t_end = h0[:, -1, 0]
    a0 = np.array([L.attempts_to_time(h0[r], t_end[r]) for r in range(64)])
    E = H.Ensemble(Hm, frs, lod, hdiv=TS.C2["ds0"], eps_r=1e-6, prec=H.PREC_MIXED)
    E.set_controller(1, m, 0.9)
    E.set_state(ys, 0.0, 1.0)
    hs = []
    while not hs or (np.any(np.concatenate(hs, axis=1)[:, -1, 0] < t_end) and len(hs) < 8):
        hs.append(E.step(500, history=True)[1])
    E.close()
    hl = np.concatenate(hs, axis=1)
    assert (hl[:, -1, 0] >= t_end).all()
    a1 = np.array([L.attempts_to_time(hl[r], t_end[r]) for r in range(64)])
    ratio = a0 / a1
    print("per-realisation attempt ratio: mean %.3f min %.3f max %.3f; lockstep attempts %d vs %d"
          % (ratio.mean(), ratio.min(), ratio.max(), a1.max(), a0.max()), flush=True)
    assert ratio.mean() >= 1.05 and ratio.min() >= 1.0
    assert a1.max() < a0.max()
