"""CPU oracle for the as-of router MLP (test infrastructure only).

The learned as-of index has no counterpart in the reference; this numpy
restatement of its router network is the checker for the device forward pass
and the MFMA gradient kernel. It is itself pinned by central finite
differences (tests/test_mlp_oracle.py) and by closed forms (the identity
initialisation computes y(x) = x).

Network (parameter layout of the device, 321 floats):
    h1 = ramp(w1 * x + b1)          w1[16], b1[16]
    h2 = ramp(W2 @ h1 + b2)         W2[16][16] (row = output unit), b2[16]
    y  = w3 . h2 + b3               w3[16], b3
with ramp(z) = min(max(z, 0), 1). Loss over a batch: mean((y - c)^2).
"""
from __future__ import annotations

import numpy as np

H = 16
N_PARAMS = 321
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 16, 32, 288, 304, 320


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1], p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def identity_params():
    p = np.zeros(N_PARAMS)
    p[OFF_W1:OFF_B1] = H
    p[OFF_B1:OFF_W2] = -np.arange(H)
    p[OFF_W2:OFF_B2] = np.eye(H).ravel()
    p[OFF_W3:OFF_B3] = 1.0 / H
    return p


def ramp(z):
    return np.minimum(np.maximum(z, 0.0), 1.0)


def forward(p, x):
    w1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = np.outer(x, w1) + b1
    h1 = ramp(z1)
    z2 = h1 @ W2.T + b2
    h2 = ramp(z2)
    return h2 @ w3 + b3


def loss_and_grad(p, x, c):
    """Mean squared error and its gradient (ramp' = 1 on the open interval)."""
    w1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    n = len(x)
    z1 = np.outer(x, w1) + b1
    h1 = ramp(z1)
    z2 = h1 @ W2.T + b2
    h2 = ramp(z2)
    y = h2 @ w3 + b3
    d = y - c
    loss = float(np.mean(d * d))
    dy = 2.0 * d / n
    g = np.zeros(N_PARAMS)
    g[OFF_B3] = dy.sum()
    g[OFF_W3:OFF_B3] = h2.T @ dy
    dz2 = np.outer(dy, w3) * ((z2 > 0) & (z2 < 1))
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (dz2.T @ h1).ravel()
    dh1 = dz2 @ W2
    dz1 = dh1 * ((z1 > 0) & (z1 < 1))
    g[OFF_W1:OFF_B1] = (dz1 * x[:, None]).sum(0)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g
