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

Primal weight: restarted PDHG adapts the primal weight omega along its run
(src/pdhg/pdhg.cpp:640-650).  A start near the optimum with the cold-start
weight ||c|| / ||b|| gains little, because near the optimum the weight updates
see only small, noisy movements.  When the training solves' final weights are
given, a second small MLP (same kernels) maps theta to log omega, and the
prediction is passed as the cases' initial primal weight.

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

    def fit(self, thetas: np.ndarray, xs: np.ndarray, ys: np.ndarray, steps: int = 3000, lr: float = 1e-3,
            omegas: np.ndarray | None = None):
        """POD of the solutions (xs k x n, ys k x m) and GPU training of theta -> coefficients;
        with omegas (the training solves' final primal weights) also theta -> log omega."""
        Z = self._to_scaled(xs, ys)
        self.mean, self.basis, A, self.energy, used = P.pod_basis(Z, self.rank)
        self.used = used
        self.coef_scale = np.maximum(A.std(0), 1e-30)
        T = A / self.coef_scale
        X = np.asarray(thetas, dtype=np.float64) / self.theta_scale
        self.mlp = P.Mlp(X.shape[1], self.hidden, self.rank, seed=self.seed)
        self.losses = self.mlp.train(X, T, steps=steps, lr=lr)
        self.wnet = None
        if omegas is not None:
            lw = np.log(np.asarray(omegas, dtype=np.float64))
            self.lw_mean, self.lw_scale = float(lw.mean()), float(max(lw.std(), 1e-12))
            self.wnet = P.Mlp(X.shape[1], 64, 1, seed=self.seed + 1)
            self.weight_losses = self.wnet.train(X, ((lw - self.lw_mean) / self.lw_scale)[:, None], steps=2000,
                                                 lr=3e-3)
        return self

    def primal_weights(self, thetas: np.ndarray) -> np.ndarray | None:
        """Predicted initial primal weights (None when fitted without omegas)."""
        if self.wnet is None:
            return None
        X = np.asarray(thetas, dtype=np.float64) / self.theta_scale
        return np.exp(self.wnet.forward(X)[:, 0] * self.lw_scale + self.lw_mean)

    def coefficients(self, thetas: np.ndarray) -> np.ndarray:
        X = np.asarray(thetas, dtype=np.float64) / self.theta_scale
        return self.mlp.forward(X) * self.coef_scale

    def predict(self, thetas: np.ndarray, coeffs: np.ndarray | None = None):
        """(x0 k x n, y0 k x m) in the original space; coeffs=0 gives the snapshot mean."""
        a = self.coefficients(thetas) if coeffs is None else coeffs
        z = P.decode_bf16(self.basis, self.mean, a)
        n = self.lp.n
        return z[:, :n] * self.xs, z[:, n:] * self.ys
