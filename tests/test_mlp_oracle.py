"""Pins the CPU MLP oracle (oracle/mlp_oracle.py) by closed forms and central
finite differences, so it can serve as the checker of the device kernels."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import mlp_oracle as M  # noqa: E402


def test_identity_initialisation_is_identity():
    x = np.linspace(0, 1, 1001)
    assert np.allclose(M.forward(M.identity_params(), x), x, atol=1e-12)


def test_monotone_for_nonnegative_weights():
    r = np.random.default_rng(1)
    p = np.abs(r.normal(size=M.N_PARAMS))
    p[M.OFF_B1:M.OFF_W2] = r.normal(size=16)
    p[M.OFF_B2:M.OFF_W3] = r.normal(size=16)
    y = M.forward(p, np.linspace(0, 1, 5000))
    assert (np.diff(y) >= -1e-12).all()


def test_gradient_matches_finite_differences():
    r = np.random.default_rng(2)
    p = M.identity_params() + r.normal(scale=0.05, size=M.N_PARAMS)
    x = r.uniform(0, 1, 257)
    c = np.sqrt(x)
    loss, g = M.loss_and_grad(p, x, c)
    eps = 1e-6
    num = np.zeros_like(g)
    for i in range(M.N_PARAMS):
        d = np.zeros_like(p)
        d[i] = eps
        lp, _ = M.loss_and_grad(p + d, x, c)
        lm, _ = M.loss_and_grad(p - d, x, c)
        num[i] = (lp - lm) / (2 * eps)
    # Kinks of the ramp can sit within eps of a sample; allow a few entries off.
    close = np.isclose(g, num, rtol=1e-4, atol=1e-7)
    assert close.mean() > 0.99, np.nonzero(~close)
    assert abs(loss - np.mean((M.forward(p, x) - c) ** 2)) < 1e-15


def test_planted_gradient_mistake_is_caught():
    # A dropped ramp' mask on layer 2 changes the gradient: the check above bites.
    r = np.random.default_rng(3)
    p = M.identity_params() + r.normal(scale=0.3, size=M.N_PARAMS)
    x = r.uniform(0, 1, 300)
    c = x ** 2
    _, g = M.loss_and_grad(p, x, c)
    w1, b1, W2, b2, w3, b3 = M.unpack(p)
    z1 = np.outer(x, w1) + b1
    h1 = M.ramp(z1)
    z2 = h1 @ W2.T + b2
    y = M.ramp(z2) @ w3 + b3
    dy = 2 * (y - c) / len(x)
    bad_dz2 = np.outer(dy, w3)  # mask omitted
    assert not np.allclose((bad_dz2.T @ h1).ravel(), g[M.OFF_W2:M.OFF_B2], atol=1e-9)
