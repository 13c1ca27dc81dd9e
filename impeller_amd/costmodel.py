"""Learned cost model of the autotuner: an FP32 MLP (3 layers, 64 hidden)
trained on the GPU by the library's own HIP kernels (csrc/mlp.hip) on the
timings the exhaustive autotuner measures.  It predicts log-latency from
configuration features only; it never touches field values.
"""
import ctypes

import numpy as np

from . import _lib


def _f32(t):
    return ctypes.c_void_p(t.data_ptr())


class Mlp:
    """Parameters, Adam state and workspace live in device memory."""

    def __init__(self, F, H=64, seed=0):
        import torch
        self.F, self.H = F, H
        self.n = int(_lib.call("impeller_mlp_param_count", F, H))
        rng = np.random.default_rng(seed)
        p = np.zeros(self.n, np.float32)
        o = 0
        # He initialisation for the ReLU layers, small output layer
        p[o:o + H * F] = rng.normal(0, np.sqrt(2.0 / F), H * F); o += H * F
        o += H
        p[o:o + H * H] = rng.normal(0, np.sqrt(2.0 / H), H * H); o += H * H
        o += H
        p[o:o + H] = rng.normal(0, np.sqrt(1.0 / H), H); o += H
        self.dev = torch.device("cuda:0")
        self.params = torch.from_numpy(p).to(self.dev)
        self.m = torch.zeros(self.n, dtype=torch.float32, device=self.dev)
        self.v = torch.zeros(self.n, dtype=torch.float32, device=self.dev)
        self.step = 0
        self._ws = None
        self._ws_b = 0

    def _workspace(self, B):
        import torch
        if self._ws is None or self._ws_b < B:
            nf = int(_lib.call("impeller_mlp_workspace_floats", self.F, self.H, B))
            self._ws = torch.empty(nf, dtype=torch.float32, device=self.dev)
            self._ws_b = B
        return self._ws

    def params_host(self):
        return self.params.cpu().numpy().copy()

    def load(self, p):
        import torch
        self.params = torch.from_numpy(np.asarray(p, np.float32).copy()).to(self.dev)

    def predict(self, X, stream):
        import torch
        X = torch.from_numpy(np.ascontiguousarray(X, np.float32)).to(self.dev)
        B = X.shape[0]
        Y = torch.empty(B, dtype=torch.float32, device=self.dev)
        _lib.call("impeller_mlp_forward", _f32(self.params), _f32(X), B, self.F, self.H, _f32(Y),
                  _f32(self._workspace(B)), stream.raw())
        stream.sync()
        return Y.cpu().numpy()

    def grad(self, X, T, stream):
        import torch
        X = torch.from_numpy(np.ascontiguousarray(X, np.float32)).to(self.dev)
        T = torch.from_numpy(np.ascontiguousarray(T, np.float32)).to(self.dev)
        B = X.shape[0]
        g = torch.empty(self.n, dtype=torch.float32, device=self.dev)
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        _lib.call("impeller_mlp_grad", _f32(self.params), _f32(X), _f32(T), B, self.F, self.H,
                  _f32(g), _f32(loss), _f32(self._workspace(B)), stream.raw())
        stream.sync()
        return g.cpu().numpy(), float(loss.item())

    def train(self, X, T, stream, steps=200, lr=1e-3):
        """Full-batch Adam steps; returns the mean squared error of the last step."""
        import torch
        X = torch.from_numpy(np.ascontiguousarray(X, np.float32)).to(self.dev)
        T = torch.from_numpy(np.ascontiguousarray(T, np.float32)).to(self.dev)
        B = X.shape[0]
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        ws = self._workspace(B)
        for _ in range(steps):
            self.step += 1
            _lib.call("impeller_mlp_train_step", _f32(self.params), _f32(self.m), _f32(self.v),
                      _f32(X), _f32(T), B, self.F, self.H, ctypes.c_float(lr), self.step,
                      _f32(loss), _f32(ws), stream.raw())
        stream.sync()
        return float(loss.item()) / B


class CostModel:
    """Standardised features -> MLP -> predicted log(ms)."""

    def __init__(self, F, H=64, seed=0):
        self.mlp = Mlp(F, H, seed)
        self.mu = np.zeros(F, np.float32)
        self.sd = np.ones(F, np.float32)
        self.ymu, self.ysd = 0.0, 1.0

    def fit(self, feats, ms, stream, steps=1500, lr=2e-3, groups=None):
        """groups: optional problem label per row; log-times are then centred
        per group (predictions rank configurations within a problem)."""
        X = np.asarray(feats, np.float32)
        y = np.log(np.asarray(ms, np.float64))
        if groups is not None:
            g = np.asarray(groups)
            for label in set(groups):
                m = g == label
                y[m] -= y[m].mean()
        self.mu = X.mean(axis=0)
        self.sd = X.std(axis=0) + 1e-6
        self.ymu, self.ysd = float(y.mean()), float(y.std() + 1e-9)
        Xs = (X - self.mu) / self.sd
        ys = ((y - self.ymu) / self.ysd).astype(np.float32)
        return self.mlp.train(Xs, ys, stream, steps=steps, lr=lr)

    def predict_ms(self, feats, stream):
        X = (np.asarray(feats, np.float32) - self.mu) / self.sd
        return np.exp(self.mlp.predict(X, stream) * self.ysd + self.ymu)
