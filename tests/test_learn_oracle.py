"""CPU: the step-size MLP oracle pinned by closed forms, and the dataset
construction of the learned controller."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 3.0e38, np.inf],
                 dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even -> 1; 1 + 3 2^-8 ties up to 1 + 2^-6;
    # 1 + 2^-7 + 2^-8 ties to the even neighbour 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -6, -2.5, 2.9975e38, np.inf], dtype=np.float32)
    got = M.bf16_round(x)
    assert np.array_equal(got[:5], want[:5])
    assert got[6] == np.inf and np.isfinite(got[5]) and abs(got[5] / 3.0e38 - 1) < 2 ** -8
    assert np.isnan(M.bf16_round(np.array([np.nan], dtype=np.float32)))[0]


def test_gelu_and_derivative():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.abs(fd - M.gelu_grad(x)).max() < 1e-8
    assert M.gelu(0.0) == 0.0 and abs(M.gelu(10.0) - 10.0) < 1e-12 and abs(M.gelu(-10.0)) < 1e-12


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(M.N_PARAM) * scale / 4).astype(np.float32)


def test_loss_gradient_against_finite_differences():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((9, M.N_IN))
    T = rng.standard_normal(9)
    mu, sd = np.zeros(M.N_IN), np.ones(M.N_IN)
    p = _params(3).astype(np.float64)
    L, g = M.loss_grad(p, mu, sd, X, T, quantize=False)
    idx = list(range(M.OFF_W1, M.OFF_W1 + 5)) + [M.OFF_B1 + 7, M.OFF_W2 + 300, M.OFF_W2 + 9000,
                                                    M.OFF_B2 + 11, M.OFF_W3 + 17, M.OFF_B3]
    for i in idx:
        h = 1e-4
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        # the oracle rounds parameters to FP32; evaluate in FP64 by going through its formulas directly
        Lp = _loss64(pp, mu, sd, X, T)
        Lm = _loss64(pm, mu, sd, X, T)
        assert abs((Lp - Lm) / (2 * h) - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def _loss64(p, mu, sd, X, T):
    W1 = p[M.OFF_W1:M.OFF_B1].reshape(M.N_HID, M.N_IN)
    b1 = p[M.OFF_B1:M.OFF_W2]
    W2 = p[M.OFF_W2:M.OFF_B2].reshape(M.N_HID, M.N_HID)
    b2 = p[M.OFF_B2:M.OFF_W3]
    W3 = p[M.OFF_W3:M.OFF_B3]
    b3 = p[M.OFF_B3]
    x = (X - mu) / sd
    out = M.gelu(M.gelu(x @ W1.T + b1) @ W2.T + b2) @ W3 + b3
    return np.mean((out - T) ** 2)


def test_oracle_forward_matches_plain_formula():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((5, M.N_IN))
    p = _params(5)
    out, _ = M.forward(p, np.zeros(M.N_IN), np.ones(M.N_IN), X, quantize=False)
    T = np.zeros(5)
    assert abs(np.mean(out ** 2) - _loss64(p.astype(np.float64), 0.0, 1.0, X, T)) < 1e-12


def test_adam_first_step_is_signed_lr():
    p = np.ones(4)
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p1, m1, m2 = M.adam(p, np.zeros(4), np.zeros(4), g, 1, 0.01)
    # bias-corrected first step: m/v^(1/2) = sign(g) (up to eps)
    assert np.allclose(p1[:3], 1.0 - 0.01 * np.sign(g[:3]), atol=1e-7)
    assert p1[3] == 1.0


def test_dataset_target_is_change_of_optimal_step():
    from hbi_amd import learned as L
    f = np.zeros((5, 12))
    f[:, 11] = [0.0, 0.1, 0.3, 0.2, 0.25]
    X, T = L.dataset([(f, np.zeros((5, 2)))])
    assert X.shape == (4, 12) and np.allclose(T, [0.1, 0.2, -0.1, 0.05])


def test_attempts_to_time():
    from hbi_amd import learned as L
    hist = np.array([[1.0, 1, 0.1, 0, 0, 0], [3.0, 2, 0.1, 1, 0, 0], [7.0, 4, 0.1, 2, 0, 0]])
    assert L.attempts_to_time(hist, 3.0) == 2 + 1
    assert L.attempts_to_time(hist, 5.0) == 3 + 3
