"""Pins the NumPy MLP restatement: its analytic gradient agrees with central
finite differences (a dropped term, wrong transpose or missing ReLU mask
fails this)."""
import numpy as np

from oracle import mlp_ref


def test_analytic_grad_matches_finite_differences():
    rng = np.random.default_rng(0)
    F, H, B = 7, 16, 23
    n = H * F + H + H * H + H + H + 1
    p = rng.normal(0, 0.4, n)
    X = rng.normal(0, 1, (B, F))
    T = rng.normal(0, 1, B)
    g = mlp_ref.grad(p, X, T, F, H)
    idx = rng.choice(n, 60, replace=False)
    num = mlp_ref.numeric_grad(p, X, T, F, H, idx)
    assert np.allclose(g[idx], num, rtol=1e-5, atol=1e-7)


def test_forward_shapes_and_relu():
    F, H = 3, 4
    n = H * F + H + H * H + H + H + 1
    p = np.zeros(n)
    p[-1] = 2.5  # b3 only: constant output
    Y, (X, A1, A2) = mlp_ref.forward(p, np.ones((5, F)), F, H)
    assert Y.shape == (5,) and np.allclose(Y, 2.5) and (A1 == 0).all()
