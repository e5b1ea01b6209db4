"""Re-fit of the engine router's CPU/GPU PDHG crossover for this GPU.

The reference router predicts each engine's solve time as
log t = w . phi(features) (src/router/router.cpp:96-104, features at
router.cpp:60-95, phi as in bench/fit_router.py:27-30) and calibrates the
PDHG engines with a fixed-iteration sweep of random LPs
(calibrateBandwidth, src/parametric/parametric.cpp:537-590: m = s rows,
n = 2s columns, s in {1e3 ... 1e6}, 512 iterations, seconds per iteration).

CPU and GPU PDHG run the same algorithm, so on one LP they take the same
number of iterations and differ only in the time per iteration.  The re-fit
therefore keeps the router's CPU-PDHG weights and models the measured
per-iteration ratio on the sweep,
    log(t_gpu / t_cpu) = a + c log(nnz),
which turns into the GPU-PDHG weights w_gpu = w_cpu + a e_1 + c e_nnz.
(On the sweep log(m+n) and log(nnz) are collinear, nnz ~ 8/3 (m+n), so only
one size term is identifiable.)  The crossover is where the ratio is 1.
"""
from __future__ import annotations

import math
import time

import numpy as np

import pdhg_amd as P

SIZES = (1000, 4000, 16000, 64000, 256000, 1000000)
ITERATIONS = 512
ENGINES = ("dual", "ipm", "pdhg-cpu", "pdhg-gpu")
# The reference router's built-in weights (src/router/router.cpp:99-104):
# 1, log(m+n), log(nnz), log(nnz/row), log(max row), log(range), eq, free.
PRIOR = {
    "dual": [-13.0, 1.20, 0.35, 0.30, 0.00, 0.05, 0.20, 0.30],
    "ipm": [-12.0, 0.40, 1.00, 0.40, 0.10, 0.03, 0.00, 0.00],
    "pdhg-cpu": [-11.5, 0.60, 0.80, 0.00, 0.10, 0.15, 0.30, 0.00],
    "pdhg-gpu": [-6.0, 0.20, 0.35, 0.00, 0.05, 0.15, 0.30, 0.00],
}
NNZ_TERM = 2  # index of log(nnz) in phi


def sweep_lp(s: int, seed: int = 7) -> P.LP:
    return P.gen_calibration(seed, s)


def effective_gbs(lp: P.LP, seconds_per_iteration: float) -> float:
    """The reference's effective bandwidth of one iteration
    (parametric.cpp:574): 2 x 12 bytes per nonzero + 10 x 8 bytes per row and column."""
    bytes_ = 2.0 * 12.0 * lp.nnz + 10.0 * 8.0 * (lp.n + lp.m)
    return bytes_ / seconds_per_iteration / 1e9


def gpu_timing(lp: P.LP, precision: int, iterations: int = ITERATIONS) -> dict:
    """Seconds per iteration of this engine on one LP (fixed iteration count)."""
    opts = P.default_options(tolerance=1e-30, max_iterations=iterations, precision=precision)
    t = time.perf_counter()
    _, st, _ = P.solve(lp, opts)
    total = time.perf_counter() - t
    per = st["solve_seconds"] / max(1, st["iterations"])
    return {"seconds_per_iteration": per, "effective_GBs": effective_gbs(lp, per),
            "setup_seconds": st["setup_seconds"], "total_seconds": total, "iterations": st["iterations"]}


def fit_ratio(nnz, t_cpu, t_gpu) -> tuple[float, float]:
    """Least squares of log(t_gpu / t_cpu) = a + c log(nnz)."""
    x = np.log(np.asarray(nnz, dtype=np.float64))
    y = np.log(np.asarray(t_gpu, dtype=np.float64) / np.asarray(t_cpu, dtype=np.float64))
    A = np.stack([np.ones_like(x), x], 1)
    a, c = np.linalg.lstsq(A, y, rcond=None)[0]
    return float(a), float(c)


def measured_crossover(nnz, t_cpu, t_gpu) -> float | None:
    """Smallest nnz at which the GPU is faster per iteration, interpolated in
    log-log between neighbouring sweep points (None: never, or always)."""
    nnz = np.asarray(nnz, dtype=np.float64)
    r = np.log(np.asarray(t_gpu, dtype=np.float64) / np.asarray(t_cpu, dtype=np.float64))
    if r[0] <= 0:
        return float(nnz[0]) if np.all(r <= 0) else None
    for i in range(1, len(r)):
        if r[i] <= 0 < r[i - 1]:
            f = r[i - 1] / (r[i - 1] - r[i])
            return float(math.exp(math.log(nnz[i - 1]) + f * (math.log(nnz[i]) - math.log(nnz[i - 1]))))
    return None


def refit(rows: list[dict], key: str = "gpu") -> dict:
    """Router model from sweep rows ({nnz, cpu: {...}, <key>: {...}}): the
    ratio fit, the crossovers and the four engines' weights."""
    nnz = [r["nnz"] for r in rows]
    tc = [r["cpu"]["seconds_per_iteration"] for r in rows]
    tg = [r[key]["seconds_per_iteration"] for r in rows]
    a, c = fit_ratio(nnz, tc, tg)
    w = {e: list(v) for e, v in PRIOR.items()}
    w["pdhg-gpu"] = list(PRIOR["pdhg-cpu"])
    w["pdhg-gpu"][0] += a
    w["pdhg-gpu"][NNZ_TERM] += c
    fit_cross = math.exp(-a / c) if c < 0 else None
    return {"ratio_fit": {"a": a, "c": c, "model": "log(t_gpu/t_cpu) = a + c log(nnz)"},
            "crossover_nnz_fit": fit_cross,
            "crossover_nnz_measured": measured_crossover(nnz, tc, tg),
            "weights": w}
