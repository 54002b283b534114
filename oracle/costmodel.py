"""CPU oracle of the autotuner's cost model (float64 numpy).

TEST INFRASTRUCTURE ONLY: nothing in the product imports this; tests use it
as the checker of the GPU kernels behind sba_cm_forward / sba_cm_backward /
sba_cm_adam. No reference counterpart: the reference has no autotuner or
learned part. The network and its gradient are restated from their
definitions:

  f(x) = w3 . relu(W2 relu(W1 x + b1) + b2) + b3,   relu'(0) = 0
  loss = (1/n) sum_s (f(x_s) - t_s)^2

with the flat parameter layout W1 [32][16], b1 [32], W2 [32][32], b2 [32],
w3 [32], b3 (1633 floats). Pinned by central finite differences of the
loss (tests/test_costmodel_cpu.py) and by a network whose output is known
in closed form.
"""
from __future__ import annotations

import numpy as np

N_IN, H1, H2 = 16, 32, 32
NPARAM = H1 * N_IN + H1 + H2 * H1 + H2 + H2 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    assert p.size == NPARAM
    o = 0
    W1 = p[o:o + H1 * N_IN].reshape(H1, N_IN); o += H1 * N_IN
    b1 = p[o:o + H1]; o += H1
    W2 = p[o:o + H2 * H1].reshape(H2, H1); o += H2 * H1
    b2 = p[o:o + H2]; o += H2
    w3 = p[o:o + H2]; o += H2
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    """Predictions y [n] and the cache of the backward pass."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1.T + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0.0)
    y = a2 @ w3 + b3
    return y, (X, z1, a1, z2, a2)


def loss(p, X, t):
    y, _ = forward(p, X)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def grad(p, X, t):
    """Gradient of the loss, flat in the parameter layout (float64)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (X, z1, a1, z2, a2) = forward(p, X)
    n = X.shape[0]
    e = 2.0 * (y - np.asarray(t, dtype=np.float64)) / n       # dL/dy
    d2 = np.outer(e, w3) * (z2 > 0)                             # dL/dz2
    d1 = (d2 @ W2) * (z1 > 0)                                   # dL/dz1
    gW1 = d1.T @ X
    gb1 = d1.sum(axis=0)
    gW2 = d2.T @ a1
    gb2 = d2.sum(axis=0)
    gw3 = a2.T @ e
    gb3 = e.sum()
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
