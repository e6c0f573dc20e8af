"""CPU oracle of the step-size MLP (test infrastructure only: used by tests/
as the checker of the device kernels, never by the product path).

FP64 numpy restatement of the network the device trains and runs:
12 -> 128 -> 128 -> 1, exact GELU (x Phi(x)), weights rounded to BF16
(round-to-nearest-even of their FP32 value) in the forward and backward
passes, biases FP32, mean-squared-error loss, Adam (beta1 0.9, beta2 0.999,
eps 1e-8).  The network has no counterpart in the reference: the reference's
only step controller is rkqs (main_LH.f90:2309-2357), which the device keeps
as the judge of each step.

Pinned by closed forms in tests/test_learn_oracle.py: GELU' against central
differences, loss gradient against finite differences of the loss (BF16
rounding off), BF16 rounding against hand-worked values.
"""
import numpy as np
from scipy.special import erf

N_IN, N_HID = 12, 128
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_IN
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAM = OFF_B3 + 1


def bf16_round(x):
    """FP32 -> nearest BF16 (ties to even), returned as float32; NaN stays NaN."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    r = np.where(nan, u, r) & 0xFFFFFFFF
    return r.astype(np.uint32).view(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def unpack(p, quantize=True):
    p = np.asarray(p, dtype=np.float32)
    q = bf16_round(p) if quantize else p
    W1 = q[OFF_W1:OFF_B1].astype(np.float64).reshape(N_HID, N_IN)
    b1 = p[OFF_B1:OFF_W2].astype(np.float64)
    W2 = q[OFF_W2:OFF_B2].astype(np.float64).reshape(N_HID, N_HID)
    b2 = p[OFF_B2:OFF_W3].astype(np.float64)
    W3 = q[OFF_W3:OFF_B3].astype(np.float64)
    b3 = float(p[OFF_B3])
    return W1, b1, W2, b2, W3, b3


def forward(p, mu, sd, X, quantize=True):
    W1, b1, W2, b2, W3, b3 = unpack(p, quantize)
    x = (np.asarray(X, dtype=np.float64) - np.asarray(mu, dtype=np.float64)) / np.asarray(sd, dtype=np.float64)
    z1 = x @ W1.T + b1
    a1 = gelu(z1)
    z2 = a1 @ W2.T + b2
    a2 = gelu(z2)
    out = a2 @ W3 + b3
    return out, (x, z1, a1, z2, a2)


def loss_grad(p, mu, sd, X, T, quantize=True):
    """mean squared error and its gradient with respect to the parameter vector
    (weights: with respect to their BF16 values when quantize)"""
    W1, b1, W2, b2, W3, b3 = unpack(p, quantize)
    out, (x, z1, a1, z2, a2) = forward(p, mu, sd, X, quantize)
    n = len(out)
    e = out - np.asarray(T, dtype=np.float64)
    L = float(np.mean(e * e))
    go = 2.0 * e / n
    g = np.zeros(N_PARAM)
    g[OFF_W3:OFF_B3] = go @ a2
    g[OFF_B3] = go.sum()
    dz2 = np.outer(go, W3) * gelu_grad(z2)
    g[OFF_W2:OFF_B2] = (dz2.T @ a1).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * gelu_grad(z1)
    g[OFF_W1:OFF_B1] = (dz1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    return L, g


def adam(p, m1, m2, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam step (t = step number after this update, 1-based)"""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m1, m2
