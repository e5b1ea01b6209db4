"""Router crossover re-fit (host numerics only): the least-squares ratio model,
the measured crossover and the weight update on synthetic sweeps whose answer
is known in closed form."""
import math

import numpy as np
import pytest

import router_fit as R


def synthetic_rows(a, c, sizes=(1e3, 4e3, 1.6e4, 6.4e4, 2.56e5, 1e6)):
    rows = []
    for s in sizes:
        nnz = 8 * s
        t_cpu = 2e-9 * nnz + 1e-6
        rows.append({"nnz": nnz, "cpu": {"seconds_per_iteration": t_cpu},
                     "gpu": {"seconds_per_iteration": t_cpu * math.exp(a + c * math.log(nnz))}})
    return rows


def test_ratio_fit_recovers_exact_model():
    a, c = 8.0, -0.7  # ratio 1 at nnz = exp(8 / 0.7) ~ 9.2e4, inside the sweep
    out = R.refit(synthetic_rows(a, c))
    assert out["ratio_fit"]["a"] == pytest.approx(a, abs=1e-9)
    assert out["ratio_fit"]["c"] == pytest.approx(c, abs=1e-9)
    # ratio 1 where a + c log(nnz) = 0
    assert out["crossover_nnz_fit"] == pytest.approx(math.exp(-a / c), rel=1e-9)
    # the measured crossover interpolates the same straight line exactly
    assert out["crossover_nnz_measured"] == pytest.approx(math.exp(-a / c), rel=1e-9)


def test_gpu_weights_are_cpu_weights_plus_ratio():
    out = R.refit(synthetic_rows(5.0, -0.5))
    w = out["weights"]
    for e in ("dual", "ipm", "pdhg-cpu"):
        assert w[e] == R.PRIOR[e]
    d = np.array(w["pdhg-gpu"]) - np.array(R.PRIOR["pdhg-cpu"])
    assert d[0] == pytest.approx(5.0) and d[R.NNZ_TERM] == pytest.approx(-0.5)
    assert np.all(np.delete(d, [0, R.NNZ_TERM]) == 0)
    # prediction identity: log t_gpu - log t_cpu = a + c log nnz for any phi
    phi = np.array([1.0, 9.1, 11.3, 2.0, 3.0, 1.5, 0.4, 0.0])
    assert phi @ np.array(w["pdhg-gpu"]) - phi @ np.array(w["pdhg-cpu"]) == pytest.approx(5.0 - 0.5 * 11.3)


def test_crossover_edge_cases():
    nnz = [1e3, 1e4, 1e5]
    assert R.measured_crossover(nnz, [1, 1, 1], [2, 2, 2]) is None      # GPU never faster
    assert R.measured_crossover(nnz, [2, 2, 2], [1, 1, 1]) == 1e3      # GPU always faster
    # sign change between the 2nd and 3rd point, log-linear interpolation
    x = R.measured_crossover(nnz, [1, 1, 1], [4, 2, 0.5])
    r1, r2 = math.log(2), math.log(0.5)
    assert x == pytest.approx(math.exp(math.log(1e4) + r1 / (r1 - r2) * (math.log(1e5) - math.log(1e4))))


def test_ridge_is_augmented_least_squares():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((30, 8))
    y = rng.standard_normal(30)
    lam = 0.37
    w = R.ridge(X, y, lam)
    Xa = np.vstack([X, math.sqrt(lam) * np.eye(8)])
    ya = np.concatenate([y, np.zeros(8)])
    assert np.allclose(w, np.linalg.lstsq(Xa, ya, rcond=None)[0], rtol=1e-10, atol=1e-12)


def test_fit_router_recovers_planted_weights_and_zero_regret():
    """Times generated exactly by planted per-engine weights: the fit recovers
    them (ridge term tiny) and the held-out choices are always the fastest."""
    rng = np.random.default_rng(5)
    planted = {e: rng.uniform(-1, 1, 8) * [3, 0.5, 0.5, 0.3, 0.3, 0.2, 1, 1] for e in R.ENGINES}
    records = []
    for q in range(60):
        f = {"rows": int(rng.integers(10, 10 ** 6)), "cols": int(rng.integers(10, 10 ** 6)),
             "nnz": int(rng.integers(100, 10 ** 7)), "avg_row_nnz": float(rng.uniform(1, 50)),
             "max_row_nnz": float(rng.uniform(1, 500)), "coefficient_range": float(10 ** rng.uniform(0, 8)),
             "equality_row_fraction": float(rng.uniform()), "free_column_fraction": float(rng.uniform())}
        t = {e: math.exp(float(np.dot(planted[e], R.phi(f)))) - 1e-3 for e in R.ENGINES}
        if min(t.values()) <= 0:
            continue
        records.append({"instance": f"i{q}", "family": "gen" if q % 2 else "netlib", "features": f, "times": t})
    out = R.fit_router(records, lam=1e-9)
    for e in R.ENGINES:
        assert np.allclose(out["weights"][e], planted[e], atol=1e-5)
    v = out["validation"]["router_held_out"]
    assert v["geo_mean"] == pytest.approx(1.0, abs=1e-6) and v["picks_fastest"] == 1.0


def test_censoring_rule():
    assert R.censored_time(2.5, True, 30.0) == 2.5
    assert R.censored_time(2.5, False, 30.0) == 120.0
    assert R.censored_time(None, True, 30.0) == 120.0


def test_phi_and_decision_match_the_reference_router(tmp_path):
    """The reference's own Router (src/router/router.cpp:24-28,157-203), loaded
    with random weights through --router-model in the reference CLI built over
    this engine: its predicted seconds equal exp(w . phi) of this module's phi
    on the features it reports, and its choice equals decide() (no GPU here,
    so pdhg-gpu is not available to either)."""
    import json

    import integration_lib as I

    if not I.available():
        pytest.skip("integration build absent (needs the reference tree)")
    import pdhg_amd as P

    if P.gpu_available():
        pytest.skip("decisions without a GPU are compared here")
    rng = np.random.default_rng(11)
    for name in ("afiro", "sc105", "kb2", "share1b"):
        w = {e: (rng.uniform(-1, 1, 8) * [4, 0.5, 0.5, 0.3, 0.3, 0.2, 1, 1]).tolist() for e in R.ENGINES}
        model = tmp_path / "router_model.json"
        model.write_text(json.dumps({"weights": w}))
        out = tmp_path / f"{name}.json"
        rc, stdout, stderr = I.cli(I.netlib_path(name), "--router-model", str(model), "--json", str(out), "--log", "0")
        assert rc == 0, stderr
        tel = json.loads(out.read_text())["telemetry"]["router"]
        assert tel["calibrated_model"] is True
        for e in ("dual", "ipm", "pdhg-cpu"):
            assert tel["predicted_seconds"][e] == pytest.approx(R.predicted_seconds(w[e], tel["features"]), rel=1e-12)
        assert tel["engine"] == R.decide(w, tel["features"], gpu=False)
