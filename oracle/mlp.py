"""CPU (numpy) restatement of the candidate-depth predictor network, for
tests only: the GPU forward/backward kernels (voro_amd/csrc/learn.hip) are
checked against it, and it is itself pinned against central finite
differences in float64 (tests/test_learned_cpu.py).  The network has no
counterpart in the reference.

Layout of the parameter vector (as in learn.h): W1 (32 x 28), b1 (32),
W2 (32 x 32), b2 (32), W3 (2 x 32), b3 (2).  Loss: mean over samples of the
summed squared output error, L = (1/n) sum_s sum_o (y_so - t_so)^2.
"""
import numpy as np

IN, H, OUT = 28, 32, 2
oW1 = 0
oB1 = oW1 + H * IN
oW2 = oB1 + H
oB2 = oW2 + H * H
oW3 = oB2 + H
oB3 = oW3 + OUT * H
NP = oB3 + OUT


def unpack(p):
    p = np.asarray(p)
    return (p[oW1:oB1].reshape(H, IN), p[oB1:oW2], p[oW2:oB2].reshape(H, H), p[oB2:oW3],
            p[oW3:oB3].reshape(OUT, H), p[oB3:NP])


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3.T + b3
    return h1, h2, y


def loss(p, X, T):
    _, _, y = forward(p, X)
    return float(np.sum((y - T) ** 2) / len(X))


def loss_grad(p, X, T):
    """Loss and analytic gradient (same dtype as the inputs)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = len(X)
    h1, h2, y = forward(p, X)
    dy = 2.0 * (y - T) / n
    dz2 = (dy @ W3) * (h2 > 0)
    dz1 = (dz2 @ W2) * (h1 > 0)
    g = np.zeros(NP, dtype=np.result_type(p, X))
    g[oW1:oB1] = (dz1.T @ X).ravel()
    g[oB1:oW2] = dz1.sum(0)
    g[oW2:oB2] = (dz2.T @ h1).ravel()
    g[oB2:oW3] = dz2.sum(0)
    g[oW3:oB3] = (dy.T @ h2).ravel()
    g[oB3:NP] = dy.sum(0)
    return float(np.sum((y - T) ** 2) / n), g


def finite_difference(p, X, T, idx, h=1e-6):
    """Central differences of the loss at parameters idx (float64)."""
    p = np.asarray(p, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    out = np.zeros(len(idx))
    for k, i in enumerate(idx):
        a = p.copy(); a[i] += h
        b = p.copy(); b[i] -= h
        out[k] = (loss(a, X, T) - loss(b, X, T)) / (2 * h)
    return out
