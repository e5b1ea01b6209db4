"""Learned warm start for a parametric LP family (the family parameter theta
changes the cost).

Pipeline: PDHG solutions of training parameters -> snapshots in the engine's
scaled space (x / (C beta), y / (R gamma), where every column and row has a
comparable magnitude) -> POD basis of `rank` modes (host FP64 Jacobi on the
snapshot Gram matrix) -> an MLP theta -> normalised POD coefficients trained on
the GPU (BF16 MFMA forward / backward, FP32 accumulation, Adam) -> at solve
time the BF16-MFMA decode z = mean + U a gives (x0, y0), passed to the batched
solve as its starting point.  Every warm-started answer still has to reach the
same KKT tolerance and pass the same certifier as a cold one; a cold start
stays available (x0 = y0 = None).

The reference solver has no warm start: it hard-wires x0 = clamp(0) and
y0 = 0 (src/pdhg/pdhg.cpp:509,519).
"""
from __future__ import annotations

import numpy as np

import pdhg_amd as P


class WarmStart:
    def __init__(self, lp: P.LP, rank: int = 32, hidden: int = 128, theta_scale: float = 0.3, seed: int = 0):
        self.lp = lp
        self.rank = rank
        self.hidden = hidden
        self.theta_scale = theta_scale
        self.seed = seed
        pr = P.prepare(lp)
        self.xs = pr["C"] * pr["beta"]   # x = x_hat * xs
        self.ys = pr["R"] * pr["gamma"]  # y = y_hat * ys
        self.mlp = None

    def _to_scaled(self, xs, ys):
        return np.hstack([xs / self.xs, ys / self.ys])

    def fit(self, thetas: np.ndarray, xs: np.ndarray, ys: np.ndarray, steps: int = 3000, lr: float = 1e-3):
        """POD of the solutions (xs k x n, ys k x m) and GPU training of theta -> coefficients."""
        Z = self._to_scaled(xs, ys)
        self.mean, self.basis, A, self.energy, used = P.pod_basis(Z, self.rank)
        self.used = used
        self.coef_scale = np.maximum(A.std(0), 1e-30)
        T = A / self.coef_scale
        X = np.asarray(thetas, dtype=np.float64) / self.theta_scale
        self.mlp = P.Mlp(X.shape[1], self.hidden, self.rank, seed=self.seed)
        self.losses = self.mlp.train(X, T, steps=steps, lr=lr)
        return self

    def coefficients(self, thetas: np.ndarray) -> np.ndarray:
        X = np.asarray(thetas, dtype=np.float64) / self.theta_scale
        return self.mlp.forward(X) * self.coef_scale

    def predict(self, thetas: np.ndarray, coeffs: np.ndarray | None = None):
        """(x0 k x n, y0 k x m) in the original space; coeffs=0 gives the snapshot mean."""
        a = self.coefficients(thetas) if coeffs is None else coeffs
        z = P.decode_bf16(self.basis, self.mean, a)
        n = self.lp.n
        return z[:, :n] * self.xs, z[:, n:] * self.ys
