"""NumPy float64 restatement of the cost-model MLP (checker for the HIP kernels).

TEST INFRASTRUCTURE ONLY.  Architecture and parameter layout are those of
impeller_amd/csrc/mlp.hip: F -> H -> H -> 1 with ReLU, parameters packed as
W1[H][F] b1[H] W2[H][H] b2[H] W3[H] b3, loss = mean squared error.
"""
import numpy as np


def unpack(p, F, H):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]; o += 1
    return W1, b1, W2, b2, W3, b3


def forward(p, X, F, H):
    W1, b1, W2, b2, W3, b3 = unpack(p, F, H)
    X = np.asarray(X, dtype=np.float64)
    A1 = np.maximum(X @ W1.T + b1, 0.0)
    A2 = np.maximum(A1 @ W2.T + b2, 0.0)
    return A2 @ W3 + b3, (X, A1, A2)


def loss(p, X, T, F, H):
    Y, _ = forward(p, X, F, H)
    return float(np.mean((Y - np.asarray(T, dtype=np.float64)) ** 2))


def grad(p, X, T, F, H):
    """Analytic gradient of the MSE (backpropagation), packed like p."""
    W1, b1, W2, b2, W3, b3 = unpack(p, F, H)
    Y, (X, A1, A2) = forward(p, X, F, H)
    B = X.shape[0]
    dY = 2.0 * (Y - np.asarray(T, dtype=np.float64)) / B
    dW3 = A2.T @ dY
    db3 = dY.sum()
    dA2 = np.outer(dY, W3) * (A2 > 0)
    dW2 = dA2.T @ A1
    db2 = dA2.sum(axis=0)
    dA1 = (dA2 @ W2) * (A1 > 0)
    dW1 = dA1.T @ X
    db1 = dA1.sum(axis=0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]])


def numeric_grad(p, X, T, F, H, idx, eps=1e-6):
    """Central finite differences of the loss for the parameters in idx."""
    p = np.asarray(p, dtype=np.float64).copy()
    out = []
    for i in idx:
        old = p[i]
        p[i] = old + eps
        lp = loss(p, X, T, F, H)
        p[i] = old - eps
        lm = loss(p, X, T, F, H)
        p[i] = old
        out.append((lp - lm) / (2 * eps))
    return np.array(out)
