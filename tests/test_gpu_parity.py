"""GPU parity of the HIP engine against the FP64 oracle (runs on an MI355X).

Bar: SpMV within FP32 rounding of the FP64 product; full solves reach the
requested relative KKT (re-evaluated independently in FP64 on the host) and
their objectives agree with the oracle's and with the closed-form optimum.
"""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
import pdhg_amd as P
from conftest import load_netlib
from lp_cases import degenerate_costs, degenerate_lp, infeasible_lp

pytestmark = pytest.mark.gpu


def scaled_csr(lp):
    pr = O.oracle_prepare(lp)
    A = sp.csr_matrix((pr["val"], pr["idx"], pr["ptr"]), shape=(lp.m, lp.n))
    A.sort_indices()
    AT = A.T.tocsr()
    AT.sort_indices()
    rfac = pr["beta"] / pr["R"]
    cfac = pr["gamma"] / pr["C"]
    return A, AT, rfac, cfac


def host_kkt(lp, x, y, cost=None):
    """Relative KKT of an unscaled (x, y) in FP64 on the host, with the
    reference's definitions applied to the original problem (scaling factors
    1, beta = gamma = 1): relP, relD, gap."""
    c = lp.c if cost is None else cost
    A = lp.scipy_csc()
    ax = A @ x
    fin = lambda v: np.abs(v) < 1e20
    res = np.maximum(np.where(fin(lp.L), lp.L - ax, 0), 0) + np.maximum(np.where(fin(lp.U), ax - lp.U, 0), 0)
    d = c - A.T @ y
    hl, hu = fin(lp.l), fin(lp.u)
    dres = np.where(~hl & ~hu, np.abs(d), np.where(~hl, np.maximum(d, 0), np.where(~hu, np.maximum(-d, 0), 0)))
    pobj = c @ x + lp.offset
    dobj = (np.sum(np.where((y > 0) & fin(lp.L), y * np.where(fin(lp.L), lp.L, 0), 0))
            + np.sum(np.where((y < 0) & fin(lp.U), y * np.where(fin(lp.U), lp.U, 0), 0))
            + np.sum(np.where((d > 0) & hl, d * np.where(hl, lp.l, 0), 0))
            + np.sum(np.where((d < 0) & hu, d * np.where(hu, lp.u, 0), 0)) + lp.offset)
    bn = np.sqrt(np.sum(np.where(fin(lp.L), lp.L, 0) ** 2) + np.sum(np.where(fin(lp.U) & (lp.U != lp.L), lp.U, 0) ** 2))
    relP = np.linalg.norm(res) / (1 + bn)
    relD = np.linalg.norm(dres) / (1 + np.linalg.norm(c))
    gap = abs(pobj - dobj) / (1 + abs(pobj) + abs(dobj))
    return relP, relD, gap, pobj


@pytest.mark.parametrize("prec", [P.FP32, P.FP64])
@pytest.mark.parametrize("k", [1, 8, 64])
def test_spmv_matches_fp64_product(gpu, prec, k):
    lp = P.gen_random_lp(2, 3000, 6000, 5.12, 1.0)  # log-normal row lengths
    A, AT, rfac, cfac = scaled_csr(lp)
    be = P.Backend(A, AT, rfac, cfac, k=k, precision=prec)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((lp.n, k))
    y = rng.standard_normal((lp.m, k))
    be.set(P.V_X, x)
    be.set(P.V_Y, y)
    be.spmv(P.V_X, P.V_AX)
    be.spmvT(P.V_Y, P.V_ATY)
    ax = be.get(P.V_AX).reshape(lp.m, k)
    aty = be.get(P.V_ATY).reshape(lp.n, k)
    if prec == P.FP32:  # the product of the FP32-rounded operands, accumulated in FP32
        xr, yr = x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)
        Ar = sp.csr_matrix((A.data.astype(np.float32).astype(np.float64), A.indices, A.indptr), shape=A.shape)
        ATr = Ar.T.tocsr()
        tol_ax = 4e-7 * 64 * (abs(Ar) @ abs(xr)) + 1e-30
        tol_aty = 4e-7 * 64 * (abs(ATr) @ abs(yr)) + 1e-30
        ex, ey = Ar @ xr, ATr @ yr
    else:
        tol_ax = 1e-14 * (abs(A) @ abs(x)) + 1e-300
        tol_aty = 1e-14 * (abs(AT) @ abs(y)) + 1e-300
        ex, ey = A @ x, AT @ y
    assert np.all(np.abs(ax - ex) <= tol_ax)
    assert np.all(np.abs(aty - ey) <= tol_aty)
    be.close()


@pytest.mark.parametrize("prec", [P.FP64, P.FP32])
def test_sliced_ellpack_spmv_matches_fp64_product(gpu, prec):
    """Width 1: the backend keeps a sliced ELLPACK copy of A and A^T (slices
    of 64 rows padded with weight 0; rows reordered by length within windows
    of 512 when their own order would pad by more than a quarter) and uses it
    when it times faster than CSR.  Forced either way, both storages give A x
    and A^T y within rounding of the FP64 product; the log-normal LP's rows
    (up to 155 entries) take the reordered copy, and the matrix of 14 rows
    checks a slice that is not full.  Other widths refuse the sliced
    storage."""
    rng = np.random.default_rng(3)
    for lp in (P.gen_random_lp(2, 3000, 6000, 5.12, 1.0), P.gen_random_lp(4, 14, 30, 3.0, 0.0)):
        A, AT, rfac, cfac = scaled_csr(lp)
        be = P.Backend(A, AT, rfac, cfac, k=1, precision=prec)
        x = rng.standard_normal(lp.n)
        y = rng.standard_normal(lp.m)
        be.set(P.V_X, x)
        be.set(P.V_Y, y)
        eps = 4e-14 if prec == P.FP64 else 4e-7 * 64  # about 160 roundings at most per row
        if prec == P.FP32:
            x, y = x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)
            A = sp.csr_matrix((A.data.astype(np.float32).astype(np.float64), A.indices, A.indptr), shape=A.shape)
            AT = A.T.tocsr()
        for layout in ("csr", "sell"):
            for t in (False, True):
                be.set_spmv_layout(t, layout)
                assert be.spmv_layout(t) == layout
            be.spmv(P.V_X, P.V_AX)
            be.spmvT(P.V_Y, P.V_ATY)
            ax, aty = be.get(P.V_AX), be.get(P.V_ATY)
            assert np.all(np.abs(ax - A @ x) <= eps * (abs(A) @ abs(x)) + 1e-300), layout
            assert np.all(np.abs(aty - AT @ y) <= eps * (abs(AT) @ abs(y)) + 1e-300), layout
        be.close()
    lp = P.gen_random_lp(4, 14, 30, 3.0, 0.0)
    A, AT, rfac, cfac = scaled_csr(lp)
    be = P.Backend(A, AT, rfac, cfac, k=8, precision=prec)
    with pytest.raises(RuntimeError, match="width 1"):
        be.set_spmv_layout(False, "sell")
    with pytest.raises(RuntimeError, match="width 33..64"):
        be.set_spmv_layout(False, "tiles")
    be.close()


@pytest.mark.parametrize("prec", [P.FP64, P.FP32])
def test_tiled_spmv_equals_one_row_per_wavefront(gpu, prec):
    """Width 64: the op-level SpMV on CSR rows in 8-row LDS-staged tiles sums
    each row in the order of one row per wavefront, so the products are
    identical, and within rounding of the FP64 product (refinery LP, scaled)."""
    lp = P.gen_refinery(3)
    pr = P.prepare(lp)
    A = sp.csr_matrix((pr["val"], pr["idx"], pr["ptr"]), shape=(lp.m, lp.n))
    AT = A.T.tocsr()
    AT.sort_indices()
    be = P.Backend(A, AT, pr["beta"] / pr["R"], pr["gamma"] / pr["C"], k=64, precision=prec)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((lp.n, 64))
    y = rng.standard_normal((lp.m, 64))
    be.set(P.V_X, x)
    be.set(P.V_Y, y)
    out = {}
    for layout in ("csr", "tiles"):
        for t in (False, True):
            be.set_spmv_layout(t, layout)
            assert be.spmv_layout(t) == layout
        be.spmv(P.V_X, P.V_AX)
        be.spmvT(P.V_Y, P.V_ATY)
        out[layout] = (be.get(P.V_AX), be.get(P.V_ATY))
    be.close()
    assert np.array_equal(out["csr"][0], out["tiles"][0]) and np.array_equal(out["csr"][1], out["tiles"][1])
    if prec == P.FP64:
        ax, aty = out["tiles"]
        assert np.all(np.abs(ax.reshape(lp.m, 64) - A @ x) <= 1e-14 * (abs(A) @ abs(x)) + 1e-300)
        assert np.all(np.abs(aty.reshape(lp.n, 64) - AT @ y) <= 1e-14 * (abs(AT) @ abs(y)) + 1e-300)


@pytest.mark.parametrize("prec", [P.FP64, P.FP32, P.FP32_AUTO])
def test_solve_random_lp_matches_oracle(gpu, prec):
    """FP32 iterates follow their own trajectory, which on this LP depends on
    the SpMV summation order (the row mapping): the FP32 case fixes the
    mapping; FP32_AUTO (stall detection, FP64 continuation) must converge
    under the autotuned one."""
    lp = P.gen_random_lp(1, 1000, 2000, 4.0, 0.0)
    obj_star = float(lp.c @ lp.xstar)
    ref, _, _, trace = O.oracle_solve(lp, tol=1e-6, trace_cap=4096)
    fixed = dict(lanes_per_row=4, row_map=1, xcd_remap=0) if prec == P.FP32 else {}
    r, st, tr = P.solve(lp, P.default_options(tolerance=1e-6, precision=prec, **fixed), trace_capacity=4096)
    print(f"prec={prec} gpu iters {r.iterations} oracle {ref[0]['iterations']} obj {r.objective:.12g} "
          f"oracle {ref[0]['objective']:.12g} closed form {obj_star:.12g} restarts {r.restarts}/{ref[0]['restarts']}")
    assert r.status == P.OPTIMAL
    assert r.relative_kkt <= 1e-6
    relP, relD, gap, pobj = host_kkt(lp, r.x, r.y)
    assert max(relP, relD, gap) <= 1.5e-6, (relP, relD, gap)
    scale = 1 + abs(obj_star)
    assert abs(r.objective - ref[0]["objective"]) <= 1e-5 * scale
    assert abs(r.objective - obj_star) <= 1e-5 * scale
    if prec == P.FP64:  # same arithmetic up to FMA / summation order: same restart decisions
        assert abs(r.iterations - ref[0]["iterations"]) <= 4 * 64
        n = min(5, len(tr), len(trace))
        for t in range(n):
            assert tr[t][0] == trace[t]["it"]
            for col, key in ((1, "relP"), (2, "relD"), (3, "gap")):
                assert abs(tr[t][col] - trace[t][key]) <= 1e-6 * abs(trace[t][key]) + 1e-14


def test_sc50a_known_optimum(gpu):
    """Reference test pdhg_cpu_matches_optimum_and_gpu_agrees (tests/cpp/test_solvers.cpp:112-130)."""
    lp, opt = load_netlib("sc50a")
    r, _, _ = P.solve(lp, P.default_options(tolerance=1e-7, precision=P.FP64))
    assert r.status == P.OPTIMAL
    assert abs(r.objective - opt) <= 1e-5 * (1 + abs(opt))
    lb, _, _ = P.safe_dual_bound(lp, r.y)
    assert lb <= opt + 1e-9


def test_afiro_batch_equals_individual(gpu):
    """Reference test pdhg_batch_equals_individual_solves (tests/cpp/test_solvers.cpp:132-149)."""
    lp, _ = load_netlib("afiro")
    costs = np.stack([lp.c.copy() for _ in range(4)])
    for k in range(4):
        costs[k, 1] = -0.4 + 0.1 * k
    ref, _, _, _ = O.oracle_solve(lp, costs, tol=1e-7)
    # FP64 at the reference test's 1e-7; FP32 iterates at the engine's 1e-6 target.
    for prec, tol in ((P.FP64, 1e-7), (P.FP32, 1e-6)):
        res, _, _ = P.solve_batch(lp, costs, P.default_options(tolerance=tol, precision=prec))
        for k in range(4):
            print(prec, k, res[k].status, res[k].iterations, res[k].objective, ref[k]["objective"])
            assert res[k].status == P.OPTIMAL
            assert abs(res[k].objective - ref[k]["objective"]) <= 1e-5 * (1 + abs(ref[k]["objective"]))


_BATCH64 = {}


def batch64_case():
    if not _BATCH64:
        lp = P.gen_random_lp(11, 120, 240, 4.0, 0.0)
        rng = np.random.default_rng(3)
        # Non-negative cost increments keep every case bounded (x >= 0).
        costs = lp.c[None, :] + rng.uniform(0.0, 0.2, (64, lp.n))
        _BATCH64["v"] = (lp, costs, O.oracle_solve(lp, costs, tol=1e-7))
    return _BATCH64["v"]


@pytest.mark.parametrize("compact", [0, 1])
def test_batch64_fp64_matches_oracle(gpu, compact):
    """A full-width FP64 run (64 lanes, one wavefront per row, split-word
    vector layout) follows the oracle's batched solve case by case; with
    compaction the live cases also move between runs of narrower widths."""
    lp, costs, (ref, xs_o, ys_o, _) = batch64_case()
    res, st, _ = P.solve_batch(lp, costs, P.default_options(tolerance=1e-7, precision=P.FP64, compact=compact))
    assert st["batch"] == 64
    for k in range(64):
        assert ref[k]["status"] == 1 and res[k].status == P.OPTIMAL, k
        assert abs(res[k].iterations - ref[k]["iterations"]) <= 4 * 64 + 0.02 * ref[k]["iterations"], k
        obj = ref[k]["objective"]
        assert abs(res[k].objective - obj) <= 1e-6 * (1 + abs(obj)), k
        assert np.linalg.norm(res[k].y - ys_o[k]) <= 1e-4 * (1 + np.linalg.norm(ys_o[k])), k
        assert np.linalg.norm(res[k].x - xs_o[k]) <= 1e-4 * (1 + np.linalg.norm(xs_o[k])), k


def small_refinery():
    return P.gen_refinery(3, periods=2, streams=12, crudes=4, products=4, qualities=5, decades=3.0)


@pytest.mark.parametrize("prec", [P.FP64, P.FP32_AUTO])
def test_refinery_family_matches_oracle(gpu, prec):
    """Badly scaled blending family (10^U(-3,3) rows/columns), 8 price cases."""
    lp = small_refinery()
    _, costs = P.gen_family(lp, 4, 8, 0.3)
    ref, _, ys_o, _ = O.oracle_solve(lp, costs, tol=1e-6, max_iter=400_000)
    res, st, _ = P.solve_batch(lp, costs, P.default_options(tolerance=1e-6, precision=prec, max_iterations=400_000))
    # Certified dual bounds of the engine's duals agree with those of the oracle's duals.
    cert_o = P.certify_batch(lp, costs, None, ys_o, np.array([r["objective"] for r in ref]))
    cert_g = P.certify_batch(lp, costs, None, np.stack([r.y for r in res]), np.array([r.objective for r in res]))
    lb_rel = np.abs(cert_g[:, 0] - cert_o[:, 0]) / (1 + np.abs(cert_o[:, 0]))
    print(f"prec={prec} max relative difference of certified bounds vs oracle {lb_rel.max():.2e}")
    assert np.all(lb_rel <= 1e-4)
    cert = P.certify_batch(lp, costs, np.stack([r.x for r in res]), np.stack([r.y for r in res]),
                           np.array([r.objective for r in res]))
    for k in range(8):
        assert ref[k]["status"] == 1 and res[k].status == P.OPTIMAL
        obj = ref[k]["objective"]
        # Two rel-KKT <= 1e-6 answers of rows scaled by 10^U(-3,3): FP64
        # iterates track the oracle's path closely; FP32 iterates reach a
        # different point of the same tolerance, held to the certifier's
        # acceptance level.
        assert abs(res[k].objective - obj) <= (1e-5 if prec == P.FP64 else 1e-4) * (1 + abs(obj))
        relP, relD, gap, _ = host_kkt(lp, res[k].x, res[k].y, costs[k])
        assert max(relP, relD, gap) <= 2e-6, (k, relP, relD, gap)
        assert cert[k, 3] == 1.0
    if prec == P.FP64:
        for k in range(8):
            assert abs(res[k].iterations - ref[k]["iterations"]) <= 4 * 64 + 0.05 * ref[k]["iterations"]


def test_compaction_keeps_results(gpu):
    lp = small_refinery()
    _, costs = P.gen_family(lp, 9, 16, 0.3)
    on, st_on, _ = P.solve_batch(lp, costs, P.default_options(precision=P.FP64, compact=1))
    off, st_off, _ = P.solve_batch(lp, costs, P.default_options(precision=P.FP64, compact=0))
    for a, b in zip(on, off):
        assert a.status == b.status == P.OPTIMAL
        assert abs(a.iterations - b.iterations) <= 4 * 64
        assert abs(a.objective - b.objective) <= 1e-6 * (1 + abs(b.objective))
    assert st_on["kernel_seconds"] <= st_off["kernel_seconds"] * 1.5


@pytest.mark.parametrize("prec", [P.FP64, P.FP32_AUTO])
def test_packed_uneven_batch_matches_oracle(gpu, prec):
    """Unrelated LPs of different sizes in one packed run (persistent kernels,
    device work queue) agree with the oracle solving each one on its own."""
    lps = P.gen_uneven_batch(7, 12, 400, 12000)
    res, st = P.solve_packed(lps, P.default_options(tolerance=1e-6, precision=prec))
    assert st["batch"] == 12
    for lp, r in zip(lps, res):
        o, _, _, _ = O.oracle_solve(lp, tol=1e-6)
        obj_star = float(lp.c @ lp.xstar)
        assert r.status == P.OPTIMAL, (lp.m, lp.n, r)
        relP, relD, gap, _ = host_kkt(lp, r.x, r.y)
        assert max(relP, relD, gap) <= 1.5e-6
        assert abs(r.objective - obj_star) <= 1e-5 * (1 + abs(obj_star))
        assert abs(r.objective - o[0]["objective"]) <= 1e-5 * (1 + abs(obj_star))
        if prec == P.FP64:
            assert abs(r.iterations - o[0]["iterations"]) <= 4 * 64 + 0.02 * o[0]["iterations"]


def test_autotuner_choice_is_cached_and_keeps_results(gpu):
    """With the row mapping on auto the first run times the candidate mappings
    on the device and keeps the fastest per kernel; a second solve of the same
    matrix reuses the choice; the answers equal those of a fixed mapping."""
    lp = P.gen_random_lp(5, 3000, 6000, 5.12, 1.0)  # log-normal row lengths, ~30k iterations
    costs = np.tile(lp.c, (8, 1))
    auto = P.default_options(tolerance=1e-6, precision=P.FP64, kernel_timing_reps=50)
    r1, s1, _ = P.solve_batch(lp, costs, auto)
    r2, s2, _ = P.solve_batch(lp, costs, auto)
    fixed = P.default_options(tolerance=1e-6, precision=P.FP64, kernel_timing_reps=50, lanes_per_row=4, row_map=1,
                              xcd_remap=0)
    r3, s3, _ = P.solve_batch(lp, costs, fixed)
    assert s1["autotuned"] == 1 and s1["tune_seconds"] > 0
    assert s2["autotuned"] == 1 and s2["tune_seconds"] == 0  # cached
    assert s3["autotuned"] == 0 and s3["col_nnz_per_lane"] == 4 and s3["col_xcd"] == 0
    for s in (s1, s2):
        for d in ("col", "row"):
            assert s[f"{d}_nnz_per_lane"] in (0, 2, 4, 8, 16) and s[f"{d}_row_map"] in (1, 2)
    assert (s1["col_nnz_per_lane"], s1["row_nnz_per_lane"]) == (s2["col_nnz_per_lane"], s2["row_nnz_per_lane"])
    print(f"tuned col {s1['col_nnz_per_lane']}/{s1['col_row_map']}/{s1['col_xcd']} "
          f"row {s1['row_nnz_per_lane']}/{s1['row_row_map']}/{s1['row_xcd']} in {s1['tune_seconds'] * 1e3:.1f} ms; "
          f"kernels tuned {s1['col_kernel_ms'] * 1e3:.1f}+{s1['row_kernel_ms'] * 1e3:.1f} us, fixed "
          f"{s3['col_kernel_ms'] * 1e3:.1f}+{s3['row_kernel_ms'] * 1e3:.1f} us")
    # the fixed mapping is one of the candidates: the tuned one is not slower (timing noise allowed)
    assert s1["col_kernel_ms"] + s1["row_kernel_ms"] <= 1.15 * (s3["col_kernel_ms"] + s3["row_kernel_ms"])
    for a, b in zip(r1, r3):
        assert a.status == b.status == P.OPTIMAL
        assert abs(a.iterations - b.iterations) <= 4 * 64 + 0.02 * b.iterations
        assert abs(a.objective - b.objective) <= 1e-6 * (1 + abs(b.objective))


@pytest.mark.parametrize("prec", [P.FP64, P.FP32_AUTO])
@pytest.mark.parametrize("name", ["plan_6x4", "plan_10x12", "plan_20x12"])
def test_plan_models_reach_published_optima(gpu, name, prec):
    """The reference's own refinery planning models (data/gen of the reference)
    reach the optimum it published for them, with a certificate."""
    from conftest import load_plan

    lp, opt = load_plan(name)
    r, st, _ = P.solve(lp, P.default_options(tolerance=1e-6, precision=prec))
    print(f"{name} prec={prec}: {r.iterations} iterations, objective {r.objective:.10g} vs published {opt:.10g}, "
          f"switch {st['fp64_switch_iteration']}")
    assert r.status == P.OPTIMAL and r.relative_kkt <= 1e-6
    assert abs(r.objective - opt) <= 1e-5 * (1 + abs(opt))
    lb, _, _ = P.safe_dual_bound(lp, r.y)
    assert lb <= opt + 1e-9 * (1 + abs(opt))
    cert = P.certify(lp, r.x, r.y, r.objective)
    assert cert["accepted"], cert
    if prec == P.FP64 and name == "plan_6x4":
        o, _, _, _ = O.oracle_solve(lp, tol=1e-6)
        assert abs(r.iterations - o[0]["iterations"]) <= 4 * 64 + 0.02 * o[0]["iterations"]


@pytest.mark.parametrize("prec,tol", [(P.FP64, 1e-4), (P.FP32, 1e-4), (P.FP64, 1e-6), (P.FP32_AUTO, 1e-6)])
def test_lognormal_lp_full_size(gpu, prec, tol):
    """Single LP at full size (m=50,000, n=100,000, nnz ~5.12e5, log-normal row
    lengths, sigma=1): size-independent checks instead of the oracle: the
    generator's closed-form optimum and an independent FP64 host re-evaluation
    of the KKT residuals (its columns have no upper bounds, so the safe dual
    bound without implied bounds is not finite and pins nothing here)."""
    lp = P.gen_random_lp(2, 50_000, 100_000, 5.12, 1.0)  # config C1: seed 2
    obj_star = float(lp.c @ lp.xstar)
    r, st, _ = P.solve(lp, P.default_options(tolerance=tol, precision=prec))
    print(f"prec={prec} tol={tol}: {r.iterations} iterations, {st['solve_seconds']:.2f} s, objective error "
          f"{abs(r.objective - obj_star) / (1 + abs(obj_star)):.2e}")
    assert r.status == P.OPTIMAL and r.relative_kkt <= tol
    relP, relD, gap, _ = host_kkt(lp, r.x, r.y)
    assert max(relP, relD, gap) <= 1.5 * tol, (relP, relD, gap)
    assert abs(r.objective - obj_star) <= 10 * tol * (1 + abs(obj_star))


def test_refinery_lp_full_size(gpu):
    """The full refinery LP (config C2: 555,852 nonzeros, rows/columns scaled
    by 10^U(-3,3)), single instance, against the reference's own CPU PDHG on
    the same LP and tolerance (tests/golden/c3_reference_cases.json, "base",
    written by tests/golden/make_c3_reference.py).
    * FP64 iterates at rel-KKT 1e-4 and 1e-6: the reference's trajectory, so
      its iteration count within 4 checks and its objective within 1e-6.
    * FP32 iterates at 1e-4 (host re-evaluation of the residuals) and
      FP32_AUTO at 1e-6 (certified): another trajectory, stopped at another
      point that meets the same tolerance.  Relative KKT bounds the gap and
      the residuals, not the distance of the objective from the reference's
      answer: a primal residual within 1e-4 moves the objective by up to the
      duals times that residual.  Measured: 1.3e-3 at 1e-4 (the reference's
      own 1e-4 answer is 6e-5 from the 1e-6 optimum), so the bounds are 5e-3
      at 1e-4 and 1e-4 at 1e-6."""
    import json

    with open(Path(__file__).parent / "golden" / "c3_reference_cases.json") as f:
        ref = json.load(f)["base"]
    lp = P.gen_refinery(3)
    rr = ref["0.0001"]
    r, st, _ = P.solve(lp, P.default_options(tolerance=1e-4, precision=P.FP64))
    err = abs(r.objective - rr["objective"]) / (1 + abs(rr["objective"]))
    print(f"refinery 1e-4 FP64: {r.iterations} iterations (reference {rr['iterations']}), objective against the "
          f"reference's {err:.1e}")
    assert r.status == P.OPTIMAL and err <= 1e-6 and abs(r.iterations - rr["iterations"]) <= 4 * 64
    r, st, _ = P.solve(lp, P.default_options(tolerance=1e-4, precision=P.FP32))
    err = abs(r.objective - rr["objective"]) / (1 + abs(rr["objective"]))
    print(f"refinery 1e-4 FP32: {r.iterations} iterations, {st['solve_seconds']:.2f} s, objective against the "
          f"reference's {err:.1e}")
    assert r.status == P.OPTIMAL and r.relative_kkt <= 1e-4
    relP, relD, gap, _ = host_kkt(lp, r.x, r.y)
    assert max(relP, relD, gap) <= 1.5e-4, (relP, relD, gap)
    assert err <= 5e-3
    out = {}
    for prec in (P.FP64, P.FP32_AUTO):
        r, st, _ = P.solve(lp, P.default_options(tolerance=1e-6, precision=prec))
        rr = ref["1e-06"]
        err = abs(r.objective - rr["objective"]) / (1 + abs(rr["objective"]))
        print(f"refinery 1e-6 prec={prec}: {r.iterations} iterations (reference {rr['iterations']}), "
              f"{st['solve_seconds']:.2f} s, objective {r.objective:.12g}, against the reference's {err:.1e}, "
              f"FP64 from {st['fp64_switch_iteration']}")
        assert r.status == P.OPTIMAL and r.relative_kkt <= 1e-6
        relP, relD, gap, _ = host_kkt(lp, r.x, r.y)
        assert max(relP, relD, gap) <= 1.5e-6, (prec, relP, relD, gap)
        cert = P.certify(lp, r.x, r.y, r.objective)
        assert cert["accepted"], cert
        if prec == P.FP64:
            assert err <= 1e-6 and abs(r.iterations - rr["iterations"]) <= 4 * 64
        else:
            assert err <= 1e-4
        out[prec] = r.objective
    assert abs(out[P.FP32_AUTO] - out[P.FP64]) <= 1e-4 * (1 + abs(out[P.FP64]))


def test_streamed_family_equals_fixed_batches(gpu):
    """More cases than one run holds: lanes freed by finished cases are
    refilled from the queue (continuous batching).  Each case's trajectory
    depends only on its own data, so in FP64 a case streamed through the
    100-case run ends exactly as it does in a plain 64-case run."""
    lp = P.gen_random_lp(11, 120, 240, 4.0, 0.0)
    rng = np.random.default_rng(8)
    costs = lp.c[None, :] + rng.uniform(0.0, 0.2, (100, lp.n))
    # No compaction: every case then runs at width 64 with the same row
    # mapping (summation order) in all three solves.
    opts = P.default_options(tolerance=1e-7, precision=P.FP64, compact=0)
    streamed, st, _ = P.solve_batch(lp, costs, opts)
    head, _, _ = P.solve_batch(lp, costs[:64], opts)
    tail, _, _ = P.solve_batch(lp, costs[36:], opts)
    assert st["batch"] == 64
    assert all(r.status == P.OPTIMAL for r in streamed)
    for q in range(100):
        ref = head[q] if q < 64 else tail[q - 36]
        assert streamed[q].iterations == ref.iterations, q
        assert streamed[q].objective == ref.objective, q
        assert np.array_equal(streamed[q].x, ref.x) and np.array_equal(streamed[q].y, ref.y), q
    # the streamed run is shorter than the two fixed batches back to back
    print(f"streamed run {st['iterations']} iterations for 100 cases; "
          f"max case {max(r.iterations for r in streamed)}")


@pytest.mark.parametrize("prec", [P.FP64, P.FP32])
def test_tiled_rows_match_one_row_per_wavefront(gpu, prec):
    """Width 64 with the rows in LDS-staged tiles of each shape (row_map 3, 4,
    5: 16, 8 or 4 rows per tile, 64, 32 or 16 KB of staged columns): each
    row's entries are summed in the order of the one-row-per-wavefront
    mapping (row_map 1), so every case ends with identical iterates.  A side
    whose longest row exceeds the shape's staging budget keeps the plain plan
    (on the log-normal LP the longest rows of A have 155 entries)."""
    import scipy.sparse as sp

    rng = np.random.default_rng(12)
    small = P.gen_random_lp(11, 120, 240, 4.0, 0.0)
    wide = P.gen_random_lp(2, 3000, 6000, 5.12, 1.0)
    vbytes = 8 if prec == P.FP64 else 4
    for lp in (small, wide):
        A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n))
        longest_col = int(np.diff(A.indptr).max())  # rows of A^T: the column kernel's side
        longest_row = int(np.diff(A.tocsr().indptr).max())
        costs = lp.c[None, :] + rng.uniform(0.0, 0.2, (64, lp.n))
        runs = {}
        for row_map in (1, 3, 4, 5):
            o = P.default_options(tolerance=1e-7, precision=prec, row_map=row_map, xcd_remap=1,
                                  max_iterations=20000, compact=0)
            runs[row_map] = P.solve_batch(lp, costs, o)
        plain, st1, _ = runs[1]
        assert (st1["col_row_map"], st1["row_row_map"]) == (1, 1)
        for row_map, budget in ((3, 64 * 1024), (4, 32 * 1024), (5, 16 * 1024)):
            tiled, st, _ = runs[row_map]
            cols = budget // (64 * vbytes)
            expect = (row_map if longest_col <= cols else 1, row_map if longest_row <= cols else 1)
            assert (st["col_row_map"], st["row_row_map"]) == expect, (row_map, expect)
            for q in range(64):
                assert tiled[q].status == plain[q].status and tiled[q].iterations == plain[q].iterations, q
                assert tiled[q].objective == plain[q].objective, q
                assert np.array_equal(tiled[q].x, plain[q].x) and np.array_equal(tiled[q].y, plain[q].y), q
            print(f"m={lp.m} n={lp.n} row_map {row_map}: tiled sides {expect}, "
                  f"{sum(r.status == P.OPTIMAL for r in tiled)}/64 optimal, iterations {st['iterations']}")


def test_graph_replays_equal_direct_launches(gpu):
    """Blocks of check_every iterations replay hipGraphs, thousands of times a
    solve, with the engine's own work between the replays: record fetches,
    lanes loaded with new cases, compaction into narrower runs, the packed
    runs' task-list rebuilds and hand-offs.  Every such run equals the same
    run launched kernel by kernel, bit for bit: a width-64 batch on the
    LDS-staged tiles, 96 cases streamed through a run of width 64, and a
    packed batch of uneven LPs."""
    lp = small_refinery()
    _, costs = P.gen_family(lp, 9, 96, 0.3)
    lps = P.gen_uneven_batch(7, 12, 400, 12000)

    def same(ra, rb):
        for a, b in zip(ra, rb):
            assert a.status == b.status and a.iterations == b.iterations and a.objective == b.objective
            assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)

    out = {}
    for graph in (1, 0):
        fixed = dict(tolerance=1e-6, precision=P.FP64, use_graph=graph, xcd_remap=1)
        out[graph] = (P.solve_batch(lp, costs[:64], P.default_options(row_map=3, **fixed)),
                      P.solve_batch(lp, costs, P.default_options(row_map=1, lanes_per_row=4, **fixed)),
                      P.solve_packed(lps, P.default_options(tolerance=1e-6, precision=P.FP64, use_graph=graph)))
    (tiled_g, streamed_g, packed_g), (tiled_d, streamed_d, packed_d) = out[1], out[0]
    assert tiled_g[1]["graph_launches"] > 0 and tiled_d[1]["graph_launches"] == 0
    assert streamed_g[1]["graph_launches"] > 0 and tiled_g[1]["col_row_map"] == 3
    same(tiled_g[0], tiled_d[0])
    same(streamed_g[0], streamed_d[0])
    same(packed_g[0], packed_d[0])
    print(f"tiled batch {tiled_g[1]['graph_launches']} replays, streamed {streamed_g[1]['graph_launches']} replays, "
          f"packed {packed_g[1]['graph_launches']} replays: equal to direct launches")


def test_streamed_family_time_limit(gpu):
    """A run's time limit (a host-side clock, the reference's Deadline,
    src/util/timer.h:22-33) stops a continuous batch: converged cases keep
    their answers, running ones report TIME_LIMIT with the iterations they
    ran, and cases never started report TIME_LIMIT with none."""
    lp = P.gen_refinery(3)
    _, costs = P.gen_family(lp, 4, 96, 0.3)
    o = P.default_options(tolerance=1e-6, precision=P.FP64, time_limit=0.3)
    res, st, _ = P.solve_batch(lp, costs, o)
    status = [r.status for r in res]
    assert all(s in (P.OPTIMAL, P.TIME_LIMIT) for s in status)
    assert status.count(P.TIME_LIMIT) >= 32  # the 32 cases past the first 64 never ran
    never = [r for r in res if r.status == P.TIME_LIMIT and r.iterations == 0]
    running = [r for r in res if r.status == P.TIME_LIMIT and r.iterations > 0]
    assert len(never) >= 32 and len(running) >= 1
    assert all(r.iterations < 1_000_000 for r in res)
    print(f"{len(running)} stopped while running, {len(never)} never started, {status.count(P.OPTIMAL)} converged")


def test_packed_handoff_with_iteration_limit(gpu):
    """The last few live LPs of a packed FP64 run move to single-LP runs with
    their state.  With an iteration limit between the 9th and the 12th LP's
    iteration count (not a multiple of the check interval), the handed-off
    LPs either converge, as their one-by-one solves do, or stop at exactly
    the limit."""
    lps = P.gen_uneven_batch(7, 12, 400, 12000)
    o = P.default_options(tolerance=1e-6, precision=P.FP64)
    free, _ = P.solve_packed(lps, o)
    its = sorted(r.iterations for r in free)
    cap = (its[8] + its[11]) // 2 + 37
    cap += 1 if cap % 64 == 0 else 0
    o = P.default_options(tolerance=1e-6, precision=P.FP64, max_iterations=cap)
    res, st = P.solve_packed(lps, o)
    capped = 0
    for lp, r in zip(lps, res):
        one, _, _ = P.solve(lp, o)
        assert r.iterations <= cap
        if r.status == P.ITERATION_LIMIT:
            capped += 1
            assert r.iterations == cap
        else:
            assert r.status == P.OPTIMAL and one.status == P.OPTIMAL
            assert abs(r.objective - one.objective) <= 1e-5 * (1 + abs(one.objective))
            assert abs(r.iterations - one.iterations) <= 4 * 64
    assert capped >= 1
    print(f"cap {cap}: {capped} of 12 at the limit")


_REPRO = """
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import pdhg_amd as P
lp = P.gen_random_lp(11, 120, 240, 4.0, 0.0)
costs = lp.c[None, :] + np.random.default_rng(5).uniform(0.0, 0.2, (48, lp.n))
o = P.default_options(tolerance=1e-7, precision=P.FP64, lanes_per_row=4, row_map=1, xcd_remap=1)
res, st, _ = P.solve_batch(lp, costs, o)
np.save(sys.argv[2], np.concatenate([np.stack([r.x for r in res]).ravel(), [r.iterations for r in res]]))
"""


def test_fixed_mapping_is_reproducible_across_processes(gpu, tmp_path):
    """With a fixed row mapping the run's widths follow from the live counts
    alone (compaction to the narrowest power of two), so two processes solving
    the same family produce identical bits; autotuned mappings and
    cost-based compaction depend on timings."""
    import subprocess
    import sys as _sys
    pkg = str(Path(P.__file__).resolve().parent)
    outs = []
    for k in range(2):
        f = tmp_path / f"run{k}.npy"
        subprocess.run([_sys.executable, "-c", _REPRO, pkg, str(f)], check=True, timeout=300)
        outs.append(np.load(f))
    assert np.array_equal(outs[0], outs[1])


def test_streamed_family_iteration_limit(gpu):
    """Continuous batching with an iteration limit that is not a multiple of
    the check interval.  Each case is checked on its own iteration count,
    every 64 iterations and at its limit (src/pdhg/pdhg.cpp:593), so a lane
    reaching its limit mid-block does not add checks to the other lanes:
    converged and capped cases of the streamed run equal those of plain
    64-case runs, and a capped case reports exactly the limit."""
    lp = P.gen_random_lp(11, 120, 240, 4.0, 0.0)
    rng = np.random.default_rng(8)
    costs = lp.c[None, :] + rng.uniform(0.0, 0.2, (100, lp.n))
    opts = P.default_options(tolerance=1e-7, precision=P.FP64, compact=0)
    free, _, _ = P.solve_batch(lp, costs[:64], opts)
    cap = int(np.median([r.iterations for r in free])) + 37
    cap += 1 if cap % 64 == 0 else 0
    opts = P.default_options(tolerance=1e-7, precision=P.FP64, compact=0, max_iterations=cap)
    streamed, st, _ = P.solve_batch(lp, costs, opts)
    head, _, _ = P.solve_batch(lp, costs[:64], opts)
    tail, _, _ = P.solve_batch(lp, costs[36:], opts)
    assert st["batch"] == 64
    status = [r.status for r in streamed]
    assert status.count(P.ITERATION_LIMIT) >= 10 and status.count(P.OPTIMAL) >= 10, status
    for q in range(100):
        ref = head[q] if q < 64 else tail[q - 36]
        r = streamed[q]
        assert r.status == ref.status, q
        assert r.iterations == ref.iterations, q
        assert r.iterations <= cap, q
        if r.status == P.ITERATION_LIMIT:
            assert r.iterations == cap, q
        assert r.objective == ref.objective and r.relative_kkt == ref.relative_kkt, q
        assert np.array_equal(r.x, ref.x) and np.array_equal(r.y, ref.y), q


def test_degenerate_structure_matches_oracle(gpu):
    """Empty, free and ranged rows, empty, free and fixed columns (lp_cases):
    an FP64 solve against the oracle and the closed-form optimum, and a
    width-64 FP64 batch whose LDS-staged tiles hold the empty rows and
    columns, case by case against the oracle's batch and bit-equal to one
    row per wavefront."""
    lp, obj = degenerate_lp()
    ref, _, _, _ = O.oracle_solve(lp, tol=1e-6)
    r, _, _ = P.solve(lp, P.default_options(tolerance=1e-6, precision=P.FP64))
    print(f"single: {r.iterations} iterations (oracle {ref[0]['iterations']}), objective {r.objective:.10g} "
          f"(oracle {ref[0]['objective']:.10g}, closed form {obj:.10g})")
    assert r.status == P.OPTIMAL and ref[0]["status"] == P.OPTIMAL
    assert abs(r.iterations - ref[0]["iterations"]) <= 4 * 64
    assert abs(r.objective - ref[0]["objective"]) <= 1e-6 * (1 + abs(obj))
    assert abs(r.objective - obj) <= 1e-5 * (1 + abs(obj))
    relP, relD, gap, _ = host_kkt(lp, r.x, r.y)
    assert max(relP, relD, gap) <= 1.5e-6, (relP, relD, gap)
    assert np.allclose(r.x[-3:], [0.0, 0.0, 2.0], rtol=0, atol=1e-12)
    assert np.all(np.abs(r.y[-2:]) <= 1e-12)  # the empty free row and the free row

    costs = degenerate_costs(lp, 64)
    ref_b, _, _, _ = O.oracle_solve(lp, costs, tol=1e-6)
    runs = {}
    for row_map in (1, 3):
        o = P.default_options(tolerance=1e-6, precision=P.FP64, row_map=row_map, xcd_remap=1, compact=0)
        runs[row_map] = P.solve_batch(lp, costs, o)
    (plain, st1, _), (tiled, st3, _) = runs[1], runs[3]
    assert (st1["col_row_map"], st1["row_row_map"]) == (1, 1)
    assert (st3["col_row_map"], st3["row_row_map"]) == (3, 3)
    for q in range(64):
        a, b, o = tiled[q], plain[q], ref_b[q]
        assert a.iterations == b.iterations and a.objective == b.objective, q
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y), q
        assert a.status == P.OPTIMAL and o["status"] == P.OPTIMAL, q
        assert abs(a.iterations - o["iterations"]) <= 4 * 64 + 0.02 * o["iterations"], q
        assert abs(a.objective - o["objective"]) <= 1e-6 * (1 + abs(o["objective"])), q
    print(f"batch: iterations {min(r.iterations for r in tiled)}..{max(r.iterations for r in tiled)}")


@pytest.mark.parametrize("k", [1, 64])
def test_infeasible_lp_stops_at_iteration_limit(gpu, k):
    """PDHG has no infeasibility detection (src/pdhg/pdhg.cpp:593-661): on
    x1 + x2 <= -1, x >= 0 the solve runs to its limit and reports
    ITERATION_LIMIT with exactly that many iterations, as the oracle does."""
    lp = infeasible_lp()
    ref, _, _, _ = O.oracle_solve(lp, tol=1e-6, max_iter=3000)
    assert (ref[0]["status"], ref[0]["iterations"]) == (P.ITERATION_LIMIT, 3000)
    o = P.default_options(tolerance=1e-6, precision=P.FP64, max_iterations=3000)
    if k == 1:
        res = [P.solve(lp, o)[0]]
    else:
        res = P.solve_batch(lp, np.repeat(lp.c[None, :], k, axis=0), o)[0]
    for r in res:
        assert (r.status, r.iterations) == (P.ITERATION_LIMIT, 3000)
        assert r.relative_kkt > 1e-3


def test_packed_edge_cases_match_oracle(gpu):
    """The packed kernels on the edge-case LPs: two degenerate-structure LPs,
    an infeasible LP and two ordinary ones in one FP64 run with an iteration
    limit past the degenerate LPs' own counts.  The solvable LPs agree with
    the oracle; the infeasible one stops at exactly the limit."""
    cap = 30001
    deg = [degenerate_lp(seed=5), degenerate_lp(seed=6, m=150, n=320)]
    plain = P.gen_uneven_batch(7, 2, 400, 3000)
    lps = [deg[0][0], infeasible_lp(), plain[0], deg[1][0], plain[1]]
    res, st = P.solve_packed(lps, P.default_options(tolerance=1e-6, precision=P.FP64, max_iterations=cap))
    assert st["batch"] == len(lps)
    assert (res[1].status, res[1].iterations) == (P.ITERATION_LIMIT, cap)
    for q in (0, 2, 3, 4):
        o, _, _, _ = O.oracle_solve(lps[q], tol=1e-6, max_iter=cap)
        r = res[q]
        print(f"LP {q}: {r.iterations} iterations (oracle {o[0]['iterations']}), objective {r.objective:.10g} "
              f"(oracle {o[0]['objective']:.10g})")
        assert r.status == P.OPTIMAL and o[0]["status"] == P.OPTIMAL, q
        assert abs(r.iterations - o[0]["iterations"]) <= 4 * 64 + 0.02 * o[0]["iterations"], q
        assert abs(r.objective - o[0]["objective"]) <= 1e-6 * (1 + abs(o[0]["objective"])), q
    for q, (lp, obj) in zip((0, 3), deg):
        assert abs(res[q].objective - obj) <= 1e-5 * (1 + abs(obj))
