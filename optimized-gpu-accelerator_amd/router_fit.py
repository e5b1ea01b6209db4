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


# ---------------------------------------------------------------------------
# Full router re-fit from measured solve times (bench/fit_router.py:30-126)
# ---------------------------------------------------------------------------
FEATURES = ["1", "log(m+n)", "log(nnz)", "log(nnz/row)", "log(max row nnz)", "log(coef range)", "eq fraction",
            "free fraction"]


def phi(f: dict) -> list[float]:
    """Feature vector of the router (RouterFeatures::phi, src/router/router.cpp:24-28;
    bench/fit_router.py:30-33): logs floored at log 1 = 0."""
    lg = lambda v: math.log(max(float(v or 0), 1.0))  # noqa: E731
    return [1.0, lg(f["rows"] + f["cols"]), lg(f["nnz"]), lg(f["avg_row_nnz"]), lg(f["max_row_nnz"]),
            lg(f["coefficient_range"]), float(f["equality_row_fraction"]), float(f["free_column_fraction"])]


def ridge(X, y, lam: float = 1e-2) -> np.ndarray:
    """argmin_w ||X w - y||^2 + lam ||w||^2 (bench/fit_router.py:68-72)."""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    return np.linalg.solve(X.T @ X + lam * np.eye(X.shape[1]), X.T @ y)


def predicted_seconds(w, f: dict) -> float:
    return math.exp(float(np.dot(w, phi(f))))


def decide(weights: dict, f: dict, gpu: bool = True, tolerance: float = 1e-6) -> str:
    """The reference router's LP choice (Router::decide, src/router/router.cpp:157-203):
    minimum predicted time with a crossover polish of 0.15 x the dual simplex
    prediction on the first-order engines, x4 for those below tolerance 1e-7,
    x1.1 on the IPM (its crossover)."""
    t = {e: predicted_seconds(weights[e], f) for e in ENGINES}
    if not gpu:
        t["pdhg-gpu"] = math.inf
    polish = 0.15 * t["dual"]
    tight = 4.0 if tolerance < 1e-7 else 1.0
    opts = [("dual", t["dual"]), ("ipm", t["ipm"] * 1.1), ("pdhg-cpu", t["pdhg-cpu"] * tight + polish),
            ("pdhg-gpu", t["pdhg-gpu"] * tight + polish)]
    return min(opts, key=lambda o: o[1])[0]


def censored_time(seconds, proven: bool, limit: float) -> float:
    """A run that did not end with a proof counts 4x its time limit (bench/fit_router.py:55-57)."""
    return float(seconds) if proven and seconds is not None else 4.0 * limit


def fit_router(records: list[dict], lam: float = 1e-2, folds: int = 5, seed: int = 1) -> dict:
    """Per-engine ridge fit of log(t + 1e-3) on phi (bench/fit_router.py:75-126).

    records: {"instance", "family", "features" (the router's feature record),
    "times": {engine: censored seconds}}.  Validation: `folds`-fold cross
    validation with each family's instances shuffled and dealt round-robin;
    reports the geometric-mean and worst regret time(chosen) / time(fastest)
    of the held-out predictions and of each fixed engine, and how often the
    fastest engine is picked.  The final model is fitted to all records."""
    import random

    names = sorted(r["instance"] for r in records)
    data = {r["instance"]: r for r in records}
    engines = [e for e in ENGINES if all(e in data[n]["times"] for n in names)]
    rng = random.Random(seed)
    fold = {}
    for fam in sorted({r["family"] for r in records}):
        members = [n for n in names if data[n]["family"] == fam]
        rng.shuffle(members)
        for i, n in enumerate(members):
            fold[n] = i % folds
    X = {n: phi(data[n]["features"]) for n in names}
    y = {n: {e: math.log(data[n]["times"][e] + 1e-3) for e in engines} for n in names}
    chosen = {}
    for f in range(folds):
        train = [n for n in names if fold[n] != f]
        test = [n for n in names if fold[n] == f]
        if not train or not test:
            continue
        W = {e: ridge([X[n] for n in train], [y[n][e] for n in train], lam) for e in engines}
        for n in test:
            pred = {e: float(np.dot(W[e], X[n])) for e in engines}
            chosen[n] = min(pred, key=pred.get)

    def regret(sel):
        r = [(data[n]["times"][sel[n]] + 1e-3) / (min(data[n]["times"][e] for e in engines) + 1e-3) for n in sel]
        return {"geo_mean": math.exp(sum(math.log(v) for v in r) / len(r)), "worst": max(r),
                "picks_fastest": sum(1 for v in r if v <= 1.0001) / len(r)}

    W = {e: ridge([X[n] for n in names], [y[n][e] for n in names], lam).tolist() for e in engines}
    for e in ENGINES:
        W.setdefault(e, [50.0] + [0.0] * 7)  # an engine without data is never chosen
    return {"weights": W, "features": FEATURES, "instances": len(names), "engines": engines,
            "validation": {"router_held_out": regret(chosen) if chosen else None,
                           **{f"always_{e}": regret({n: e for n in names}) for e in engines}}}


def crossover_from_sweep(sweep: dict) -> dict:
    """The CPU/GPU PDHG crossover re-fit from the output of the reference's
    own `pramana calibrate` verb run over this engine (calibrateBandwidth,
    src/parametric/parametric.cpp:537-590: rows with "nnz", "cpu" and "gpu"
    seconds per iteration)."""
    rows = [r for r in sweep["pdhg_iteration_timing"] if "gpu" in r and "cpu" in r]
    out = refit(rows, "gpu")
    out["gpu"] = sweep.get("gpu")
    out["cpu_threads"] = sweep.get("cpu_threads")
    out["sweep"] = [{"nnz": r["nnz"], "cpu_seconds_per_iteration": r["cpu"]["seconds_per_iteration"],
                     "gpu_seconds_per_iteration": r["gpu"]["seconds_per_iteration"],
                     "gpu_effective_GBs": r["gpu"]["effective_GBs"]} for r in rows]
    return out


if __name__ == "__main__":
    import json
    import sys

    if len(sys.argv) != 3 or sys.argv[1] != "crossover":
        raise SystemExit("usage: python router_fit.py crossover <calibrate output JSON>")
    print(json.dumps(crossover_from_sweep(json.load(open(sys.argv[2]))), indent=1))
