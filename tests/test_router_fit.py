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
