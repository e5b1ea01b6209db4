"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

* bit-for-bit equality with the reference's own PDHG compiled from its sources
  (oracle/_ref), on generated LPs, Netlib fixtures and a batched family;
* the closed-form optimum of the generated LPs (c'x* = b'y*);
* the safe-bound known answers of the reference tests (LB = 2.8 at
  y = (0.4, 0.2); any dual gives a valid bound), tests/cpp/test_core.cpp:175-193;
* the Netlib fixtures against scipy's HiGHS (a library routine) and the
  optima the reference tests assert (tests/cpp/test_solvers.cpp:27-28).
"""
import numpy as np
import pytest

import oracle_lib as O
import pdhg_amd as P
from conftest import load_netlib, load_plan
from lp_cases import degenerate_lp, infeasible_lp

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference (oracle/_ref) not built")


@pytest.fixture(scope="module", autouse=True)
def _threads():
    if O.ref_available():
        O.ref_set_threads(4)


def two_var_lp():
    # min x + y  s.t. x + 2y >= 4, 3x + y >= 6, x, y in [0, 10]: optimum 2.8 at (1.6, 1.2)
    start = np.array([0, 2, 4], np.int32)
    index = np.array([0, 1, 0, 1], np.int32)
    value = np.array([1.0, 3.0, 2.0, 1.0])
    inf = float("inf")
    return P.LP(2, 2, start, index, value, np.array([1.0, 1.0]), np.zeros(2), np.full(2, 10.0),
                np.array([4.0, 6.0]), np.array([inf, inf]))


@needs_ref
@pytest.mark.parametrize("seed,m,n", [(1, 1000, 2000), (11, 300, 500)])
def test_oracle_bitwise_equals_reference_random_lp(seed, m, n):
    lp = P.gen_random_lp(seed, m, n, 4.0, 0.0)
    o, xo, yo, _ = O.oracle_solve(lp, tol=1e-6)
    r, xr, yr = O.ref_solve(lp, tol=1e-6)
    assert o[0]["iterations"] == r["iterations"]
    assert o[0]["restarts"] == r["restarts"]
    assert o[0]["objective"] == r["objective"]
    assert o[0]["spectral_norm"] == r["spectral_norm"]
    assert np.array_equal(xo[0], xr) and np.array_equal(yo[0], yr)


@needs_ref
@pytest.mark.parametrize("name", ["afiro", "sc50a", "adlittle"])
def test_oracle_bitwise_equals_reference_netlib(name):
    lp, _ = load_netlib(name)
    o, xo, yo, _ = O.oracle_solve(lp, tol=1e-6, max_iter=200_000)
    r, xr, yr = O.ref_solve(lp, tol=1e-6, max_iter=200_000)
    assert (o[0]["iterations"], o[0]["status"], o[0]["objective"]) == (r["iterations"], r["status"], r["objective"])
    assert np.array_equal(xo[0], xr) and np.array_equal(yo[0], yr)


@needs_ref
def test_oracle_bitwise_equals_reference_degenerate_structure():
    """Empty, free and ranged rows, empty, free and fixed columns."""
    lp, _ = degenerate_lp()
    o, xo, yo, _ = O.oracle_solve(lp, tol=1e-6)
    r, xr, yr = O.ref_solve(lp, tol=1e-6)
    assert (o[0]["iterations"], o[0]["status"], o[0]["objective"]) == (r["iterations"], r["status"], r["objective"])
    assert o[0]["spectral_norm"] == r["spectral_norm"]
    assert np.array_equal(xo[0], xr) and np.array_equal(yo[0], yr)


@needs_ref
def test_oracle_iteration_limit_equals_reference_on_infeasible_lp():
    lp = infeasible_lp()
    o, xo, yo, _ = O.oracle_solve(lp, tol=1e-6, max_iter=3000)
    r, xr, yr = O.ref_solve(lp, tol=1e-6, max_iter=3000)
    assert (o[0]["iterations"], o[0]["status"]) == (r["iterations"], r["status"]) == (3000, P.ITERATION_LIMIT)
    assert np.array_equal(xo[0], xr) and np.array_equal(yo[0], yr)


def test_oracle_degenerate_structure_reaches_closed_form():
    lp, obj = degenerate_lp()
    o, xs, ys, _ = O.oracle_solve(lp, tol=1e-6)
    assert o[0]["status"] == P.OPTIMAL and o[0]["relative_kkt"] <= 1e-6
    assert abs(o[0]["objective"] - obj) <= 1e-5 * (1 + abs(obj))
    # the added columns at their optimal values: lower bound, initial point, fixed value
    assert xs[0][-3:].tolist() == [0.0, 0.0, 2.0]


@needs_ref
def test_oracle_batch_bitwise_equals_reference():
    lp, _ = load_netlib("afiro")
    costs = np.stack([lp.c.copy() for _ in range(4)])
    for k in range(4):
        costs[k, 1] = -0.4 + 0.1 * k
    o, xo, yo, _ = O.oracle_solve(lp, costs, tol=1e-7)
    r, xr, yr = O.ref_solve_batch(lp, costs, tol=1e-7)
    for k in range(4):
        assert o[k]["iterations"] == r[k]["iterations"] and o[k]["objective"] == r[k]["objective"]
    assert np.array_equal(xo, xr) and np.array_equal(yo, yr)


def test_oracle_reaches_closed_form_optimum():
    lp = P.gen_random_lp(1, 1000, 2000, 4.0, 0.0)
    obj = float(lp.c @ lp.xstar)
    assert abs(obj - float(lp.L @ lp.ystar)) <= 1e-9 * (1 + abs(obj))  # strong duality of the generated pair
    o, xs, ys, _ = O.oracle_solve(lp, tol=1e-6)
    assert o[0]["status"] == 1 and o[0]["relative_kkt"] <= 1e-6
    assert abs(o[0]["objective"] - obj) <= 1e-5 * (1 + abs(obj))


def test_oracle_netlib_known_optima():
    for name in ("afiro", "sc50a"):
        lp, opt = load_netlib(name)
        o, _, ys, _ = O.oracle_solve(lp, tol=1e-7, max_iter=500_000)
        assert o[0]["status"] == 1
        assert abs(o[0]["objective"] - opt) <= 1e-5 * (1 + abs(opt))
        lb, _, _ = O.oracle_safe_bound(lp, ys[0])
        assert lb <= opt + 1e-9


def test_oracle_trace_is_consistent():
    lp = P.gen_random_lp(3, 200, 400, 4.0, 0.0)
    o, _, _, tr = O.oracle_solve(lp, tol=1e-6, trace_cap=10000)
    assert [t["it"] for t in tr] == [64 * (i + 1) - 1 for i in range(len(tr))]
    assert tr[-1]["it"] == o[0]["iterations"]
    assert sum(t["restarted"] for t in tr) == o[0]["restarts"]


def test_safe_bound_known_answer():
    lp = two_var_lp()
    lb, inf, _ = O.oracle_safe_bound(lp, np.array([0.4, 0.2]))
    assert inf == 0 and lb <= 2.8 and abs(lb - 2.8) <= 1e-12
    rng = np.random.default_rng(1)
    for _ in range(200):
        y = rng.uniform(-0.5, 1.5, 2)
        assert O.oracle_safe_bound(lp, y)[0] <= 2.8 + 1e-12


def test_safe_bound_brute_force_vertices():
    """LB(y) equals the textbook Lagrangian bound min_x c'x + y'(b - Ax) over the
    box, which for this 2-variable LP is attained at box vertices."""
    lp = two_var_lp()
    A = lp.scipy_csc().toarray()
    rng = np.random.default_rng(7)
    for _ in range(50):
        y = np.maximum(rng.uniform(-0.2, 1.0, 2), 0)  # rows are >= rows: y >= 0 after projection
        z = lp.c - A.T @ y
        verts = [(a, b) for a in (0.0, 10.0) for b in (0.0, 10.0)]
        exact = min(z @ np.array(v) for v in verts) + y @ lp.L
        lb = O.oracle_safe_bound(lp, y)[0]
        assert lb <= exact + 1e-15 and exact - lb <= 1e-12 * (1 + abs(exact))


@needs_ref
def test_safe_bound_bitwise_equals_reference():
    lp = P.gen_random_lp(5, 400, 800, 4.0, 0.0)
    ub = lp.u.copy()
    ub[:] = 5.0  # finite boxes so the bound is finite for arbitrary duals
    lp2 = P.LP(lp.m, lp.n, lp.start, lp.index, lp.value, lp.c, lp.l, ub, lp.L, lp.U)
    rng = np.random.default_rng(2)
    for _ in range(10):
        y = lp.ystar + 1e-3 * rng.standard_normal(lp.m)
        assert O.oracle_safe_bound(lp2, y) == O.ref_safe_bound(lp2, y)
    assert O.oracle_safe_bound(lp, lp.ystar) == O.ref_safe_bound(lp, lp.ystar)


def highs_objective(lp):
    from scipy.optimize import linprog
    import scipy.sparse as sps

    A = lp.scipy_csc()
    fin = lambda v: np.isfinite(v) & (np.abs(v) < 1e20)  # noqa: E731
    eq = fin(lp.L) & fin(lp.U) & (lp.L == lp.U)
    ub_rows = fin(lp.U) & ~eq
    lb_rows = fin(lp.L) & ~eq
    A_ub = sps.vstack([A[ub_rows], -A[lb_rows]])
    b_ub = np.concatenate([lp.U[ub_rows], -lp.L[lb_rows]])
    bounds = [(lo if fin(np.array(lo)) else None, hi if fin(np.array(hi)) else None) for lo, hi in zip(lp.l, lp.u)]
    r = linprog(lp.c, A_ub=A_ub, b_ub=b_ub, A_eq=A[eq], b_eq=lp.L[eq], bounds=bounds, method="highs")
    assert r.status == 0
    return r.fun + lp.offset


@pytest.mark.parametrize("name", ["plan_6x4", "plan_10x12", "plan_20x12"])
def test_plan_fixtures_match_published_optima(name):
    """The parsed reference planning models reproduce the optimum the reference
    published for them (results/gpu/planning_engines.csv, dual simplex)."""
    lp, opt = load_plan(name)
    assert abs(highs_objective(lp) - opt) <= 1e-7 * (1 + abs(opt))


def test_oracle_reaches_published_plan_optimum():
    """The oracle PDHG on the reference's plan_6x4 model at rel-KKT 1e-6."""
    lp, opt = load_plan("plan_6x4")
    r, _, ys, _ = O.oracle_solve(lp, tol=1e-6)
    assert r[0]["status"] == 1
    assert abs(r[0]["objective"] - opt) <= 1e-5 * (1 + abs(opt))
    lb, _, _ = P.safe_dual_bound(lp, ys[0])
    assert lb <= opt + 1e-9 * (1 + abs(opt))


def test_netlib_fixtures_match_highs():
    from scipy.optimize import linprog

    for name in ("afiro", "sc50a", "adlittle", "blend", "share2b"):
        lp, opt = load_netlib(name)
        A = lp.scipy_csc()
        fin = lambda v: np.isfinite(v) & (np.abs(v) < 1e20)
        eq = fin(lp.L) & fin(lp.U) & (lp.L == lp.U)
        ub_rows = fin(lp.U) & ~eq
        lb_rows = fin(lp.L) & ~eq
        A_ub = __import__("scipy.sparse", fromlist=["vstack"]).vstack([A[ub_rows], -A[lb_rows]])
        b_ub = np.concatenate([lp.U[ub_rows], -lp.L[lb_rows]])
        bounds = [(lo if fin(np.array(lo)) else None, hi if fin(np.array(hi)) else None) for lo, hi in zip(lp.l, lp.u)]
        r = linprog(lp.c, A_ub=A_ub, b_ub=b_ub, A_eq=A[eq], b_eq=lp.L[eq], bounds=bounds, method="highs")
        assert r.status == 0
        assert abs(r.fun + lp.offset - opt) <= 1e-7 * (1 + abs(opt)), (name, r.fun, opt)


@pytest.mark.parametrize("which", ["random", "degenerate"])
def test_host_loop_over_oracle_ops_equals_oracle_solve(which):
    """The reference host loop restated in tests/pdhg_driver.py, run over the
    oracle's Backend ops one at a time, gives the oracle's solve bit for bit
    (iterations, restarts, objective, x, y) on two cost cases: this pins the
    op-level oracle (compared op by op with the device backend) and the
    driver (run over the device backend) to the solve that is itself pinned
    bitwise to the compiled reference."""
    import pdhg_driver as D
    from lp_cases import degenerate_costs, degenerate_lp

    if which == "random":
        lp = P.gen_random_lp(12, 100, 200)
        rng = np.random.default_rng(4)
        costs = np.stack([lp.c, lp.c + rng.uniform(0, 0.05, lp.n)])
    else:
        lp, _ = degenerate_lp(seed=5, m=60, n=120)
        costs = degenerate_costs(lp, 2)
    pr = O.oracle_prepare(lp)
    A, AT = D.scaled_csr_pair(lp, pr)
    ops = O.OracleOps(A, AT, pr["beta"] / pr["R"], pr["gamma"] / pr["C"], k=2)
    res = D.run_pdhg(ops, lp, pr, costs, tol=1e-6, max_iter=100000)
    ref, xs, ys, _ = O.oracle_solve(lp, costs, tol=1e-6, max_iter=100000)
    for b in range(2):
        assert res[b]["status"] == ref[b]["status"] == 1
        assert res[b]["iterations"] == ref[b]["iterations"]
        assert res[b]["restarts"] == ref[b]["restarts"]
        assert res[b]["objective"] == ref[b]["objective"]
        assert np.array_equal(res[b]["x"], xs[b]) and np.array_equal(res[b]["y"], ys[b])
