"""HIP cost-model MLP: forward and gradients against the NumPy float64
restatement (oracle/mlp_ref.py), and training reduces the loss."""
import numpy as np
import pytest

from impeller_amd import costmodel as CM
from oracle import mlp_ref

pytestmark = pytest.mark.gpu

# FP32 forward / backward vs float64: relative tolerance 1e-4 (sums of <= 64
# products in fp32), absolute 1e-5 for near-zero gradient entries
RTOL, ATOL = 1e-4, 1e-5


def test_forward_and_grad_vs_numpy(stream):
    rng = np.random.default_rng(1)
    F, H, B = 12, 64, 37
    m = CM.Mlp(F, H, seed=3)
    X = rng.normal(0, 1, (B, F)).astype(np.float32)
    T = rng.normal(0, 1, B).astype(np.float32)
    p = m.params_host()
    Y = m.predict(X, stream)
    Yref, _ = mlp_ref.forward(p.astype(np.float64), X, F, H)
    assert np.allclose(Y, Yref, rtol=RTOL, atol=ATOL)
    g, loss_sum = m.grad(X, T, stream)
    gref = mlp_ref.grad(p.astype(np.float64), X, T, F, H)
    assert np.allclose(g, gref, rtol=RTOL, atol=ATOL)
    assert np.isclose(loss_sum / B, mlp_ref.loss(p.astype(np.float64), X, T, F, H), rtol=1e-4)


def test_training_reduces_loss(stream):
    rng = np.random.default_rng(2)
    F, H, B = 6, 64, 256
    X = rng.normal(0, 1, (B, F)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    m = CM.Mlp(F, H, seed=4)
    first = m.train(X, T, stream, steps=1, lr=1e-3)
    last = m.train(X, T, stream, steps=400, lr=3e-3)
    assert last < 0.2 * first
