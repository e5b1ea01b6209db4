"""CPU-side tests of the product library: the C-ABI surface, the host set-up
(bit-identical to the reference), the certifier and the generators.  No GPU
compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import pdhg_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference (oracle/_ref) not built")


def header_symbols():
    text = open(os.path.join(ROOT, "include", "pdhg_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdhg_[a-zA-Z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = P.lib()
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(P.EXPORTS) == syms


def test_no_gpu_means_loud_failure():
    if P.gpu_available():
        pytest.skip("a GPU is present")
    info = P.gpu_info()
    assert not info["available"] and info["reason"]
    lp = P.gen_random_lp(1, 50, 100)
    with pytest.raises(RuntimeError, match="GPU unavailable"):
        P.solve(lp)


def test_bad_input_is_rejected():
    lp = P.gen_random_lp(1, 50, 100)
    bad = P.LP(lp.m, lp.n, lp.start, lp.index.copy(), lp.value, lp.c, lp.l, lp.u, lp.L, lp.U)
    bad.index[3] = lp.m + 5
    with pytest.raises(RuntimeError, match="row index out of range"):
        P.safe_dual_bound(bad, np.zeros(lp.m))


def test_bad_solver_input_is_rejected_before_device_work():
    """NaN column data is rejected as the reference's Model::validate does
    (src/core/model.cpp:111-113), NULL arrays and out-of-range options by the
    ABI's own checks; all before any device is touched, so the messages are
    the same with or without a GPU."""
    lp = P.gen_random_lp(1, 50, 100)
    for field in ("c", "l", "u"):
        v = getattr(lp, field).copy()
        v[7] = np.nan
        bad = P.LP(lp.m, lp.n, lp.start, lp.index, lp.value, lp.c, lp.l, lp.u, lp.L, lp.U)
        setattr(bad, field, v)
        with pytest.raises(RuntimeError, match="NaN in column data"):
            P.solve(bad)
        with pytest.raises(RuntimeError, match="NaN in column data"):
            P.solve_packed([lp, bad])
    s = lp.struct()
    s.col_start = None
    res = P.ResultStruct()
    assert P.lib().pdhg_solve(C.byref(s), None, C.byref(res), None) == -1
    assert b"null column pointers" in P.lib().pdhg_last_error()
    for kw, msg in ((dict(precision=7), "unknown precision"), (dict(row_map=6), "row_map"),
                    (dict(lanes_per_row=65), "lanes_per_row"), (dict(xcd_remap=2), "xcd_remap"),
                    (dict(tolerance=float("nan")), "tolerance is NaN"), (dict(check_every=1), "check_every"),
                    (dict(max_iterations=0), "max_iterations"), (dict(device=-1), "device ordinal"),
                    (dict(device=1 << 30), "device ordinal")):
        with pytest.raises(RuntimeError, match=msg):
            P.solve(lp, P.default_options(**kw))
        with pytest.raises(RuntimeError, match=msg):
            P.solve_packed([lp], P.default_options(**kw))


@pytest.mark.parametrize("gen", ["c0", "c1", "netlib", "refinery"])
def test_prepare_bitwise_equals_oracle(gen):
    from conftest import load_netlib

    if gen == "c0":
        lp = P.gen_random_lp(1, 1000, 2000, 4.0, 0.0)
    elif gen == "c1":
        lp = P.gen_random_lp(2, 5000, 10000, 5.12, 1.0)
    elif gen == "netlib":
        lp, _ = load_netlib("share2b")
    else:
        lp = P.gen_refinery(3, periods=2, products=6, qualities=5)
    a, b = P.prepare(lp), O.oracle_prepare(lp)
    for key in ("R", "C", "ptr", "idx", "val"):
        assert np.array_equal(a[key], b[key]), key
    for key in ("beta", "gamma", "normA"):
        assert a[key] == b[key], key


@needs_ref
def test_prepare_norm_equals_reference():
    lp = P.gen_random_lp(1, 1000, 2000, 4.0, 0.0)
    r, _, _ = O.ref_solve(lp, tol=1e-6, max_iter=64)
    assert P.prepare(lp)["normA"] == r["spectral_norm"]


def test_generator_random_lp_closed_form():
    lp = P.gen_random_lp(1, 1000, 2000, 4.0, 0.0)
    A = lp.scipy_csc()
    assert 7800 <= lp.nnz <= 8000
    assert np.allclose(A @ lp.xstar, lp.L, rtol=0, atol=1e-12) and np.array_equal(lp.L, lp.U)
    s = lp.c - A.T @ lp.ystar
    assert np.all(s >= -1e-12)
    assert np.all(np.minimum(s, lp.xstar) <= 1e-12)  # complementarity
    assert np.sum(lp.xstar == 0) == 1000
    assert np.all(lp.l == 0) and np.all(np.isinf(lp.u))


def test_certifier_on_unbounded_columns():
    """Config-0 LPs have x >= 0 without upper bounds, and their mixed-sign
    equality rows imply none.  The safe bound at the generator's exact dual
    is finite and certifies the optimum.  Moving one reduced cost just below
    zero (as in any inexact PDHG dual) makes the rigorous bound -inf
    (src/cert/certifier.cpp:28-118), so such answers are never accepted."""
    lp = P.gen_random_lp(1, 200, 400, 4.0, 0.0)
    A = lp.scipy_csc()
    obj = float(lp.c @ lp.xstar)
    lb, _, _ = P.safe_dual_bound(lp, lp.ystar)
    assert np.isfinite(lb) and abs(lb - obj) <= 1e-9 * (1 + abs(obj))
    assert P.certify(lp, lp.xstar, lp.ystar, obj)["accepted"]
    s = lp.c - A.T @ lp.ystar
    j = int(np.argmin(np.where(A.getnnz(axis=0) > 0, s, np.inf)))
    i = int(A[:, j].nonzero()[0][0])
    y = lp.ystar.copy()
    y[i] += (s[j] + 1e-9) / A[i, j]  # (c - A^T y)_j = -1e-9
    assert (lp.c - A.T @ y)[j] < 0
    assert P.safe_dual_bound(lp, y)[0] == -np.inf
    assert not P.certify(lp, lp.xstar, y, obj)["accepted"]


def test_generator_lognormal_rows():
    lp = P.gen_random_lp(2, 50000, 100000, 5.12, 1.0)
    assert abs(lp.nnz - 5.12e5) / 5.12e5 < 0.03
    rows = np.diff(lp.scipy_csc().tocsr().indptr)
    assert rows.max() > 8 * rows.mean()  # heavy tail of a log-normal (sigma = 1)


def test_generator_refinery_shape():
    lp = P.gen_refinery(3)
    assert abs(lp.nnz - 5.6e5) / 5.6e5 < 0.05
    assert np.all(np.isfinite(lp.u)) and np.all(lp.l == 0)  # every column bounded
    A = lp.scipy_csc()
    assert np.all(lp.L <= 1e-12) and np.all(lp.U >= -1e-12)  # x = 0 is feasible
    vals = np.abs(A.data)
    assert vals.max() / vals.min() > 1e8  # deliberately bad scaling
    th, cs = P.gen_family(lp, 4, 64, 0.3)
    assert th.shape == (64, 4) and np.all(np.abs(th) <= 0.3)
    changed = np.nonzero(np.any(cs != lp.c, axis=0))[0]
    assert 0 < changed.size <= 4 * 12
    assert np.allclose(P.family_costs(lp, th), cs)


def test_safe_bound_matches_oracle_bitwise():
    lp = P.gen_random_lp(5, 400, 800, 4.0, 0.0)
    rng = np.random.default_rng(3)
    for _ in range(10):
        y = lp.ystar + 1e-3 * rng.standard_normal(lp.m)
        assert P.safe_dual_bound(lp, y) == O.oracle_safe_bound(lp, y)


def test_safe_bound_known_answer():
    from test_oracle import two_var_lp

    lp = two_var_lp()
    lb, inf, _ = P.safe_dual_bound(lp, np.array([0.4, 0.2]))
    assert inf == 0 and lb <= 2.8 and 2.8 - lb <= 1e-12
    rng = np.random.default_rng(1)
    for _ in range(200):
        assert P.safe_dual_bound(lp, rng.uniform(-0.5, 1.5, 2))[0] <= 2.8 + 1e-12


@needs_ref
def test_family_certificate_matches_reference():
    lp = P.gen_refinery(3, periods=2, products=6, qualities=5)
    rng = np.random.default_rng(4)
    y = rng.standard_normal(lp.m) * 1e-2
    cert = P.certify(lp, None, y, 0.0)
    assert cert["bound"] == O.ref_family_bound(lp, y)


def test_certificate_rejects_wrong_objective():
    from test_oracle import two_var_lp

    lp = two_var_lp()
    good = P.certify(lp, np.array([1.6, 1.2]), np.array([0.4, 0.2]), 2.8)
    assert good["accepted"] and good["primal_violation"] == 0
    bad = P.certify(lp, np.array([2.0, 2.0]), np.array([0.4, 0.2]), 4.0)
    assert not bad["accepted"]
    out = P.certify_batch(lp, None, np.array([[1.6, 1.2], [2.0, 2.0]]), np.array([[0.4, 0.2], [0.4, 0.2]]),
                          np.array([2.8, 4.0]), threads=2)
    assert out[0, 3] == 1.0 and out[1, 3] == 0.0


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N with N > 1 exits with status 2 and
    says why, before touching a device."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 2
    assert "one GPU only" in out.stderr and out.stdout == ""


def test_backend_create_rejects_bad_csr_before_device_work():
    """pdhg_backend_create checks its CSR arrays as model_from checks a C-ABI
    LP (null arrays, ptr[0] == 0, monotone pointers, index range, finite
    values and scale factors) before it touches a device, so a malformed call
    returns NULL with a message instead of reaching the kernels."""
    import scipy.sparse as sp

    lp = P.gen_random_lp(2, 30, 60)
    A = lp.scipy_csc().tocsr()
    AT = A.T.tocsr()
    rf, cf = np.ones(lp.m), np.ones(lp.n)
    L = P.lib()

    def create(aptr=None, aidx=None, aval=None, tptr=None, rfac=rf, cfac=cf, precision=P.FP64):
        arrs = [np.ascontiguousarray(a) for a in (
            A.indptr if aptr is None else aptr, A.indices if aidx is None else aidx, A.data if aval is None else aval,
            AT.indptr if tptr is None else tptr, AT.indices, AT.data)]
        ints = [a.astype(np.int32) for a in (arrs[0], arrs[1], arrs[3], arrs[4])]
        h = L.pdhg_backend_create(lp.m, lp.n, P._i(ints[0]), P._i(ints[1]), P._d(arrs[2].astype(np.float64)),
                                  P._i(ints[2]), P._i(ints[3]), P._d(arrs[5].astype(np.float64)),
                                  None if rfac is None else P._d(rfac), None if cfac is None else P._d(cfac),
                                  1, precision, 0)
        assert not h
        return L.pdhg_last_error().decode()

    p = A.indptr.copy()
    p[0] = 1
    assert "start at 0" in create(aptr=p)
    p = A.indptr.copy()
    p[5], p[6] = p[6], p[5] - 1  # not monotone
    assert "not monotone" in create(aptr=p)
    i = A.indices.copy()
    i[4] = lp.n
    assert "column index out of range" in create(aidx=i)
    v = A.data.copy()
    v[2] = np.inf
    assert "non-finite value" in create(aval=v)
    assert "null scale factors" in create(rfac=None)
    r = rf.copy()
    r[3] = np.nan
    assert "non-finite row factor" in create(rfac=r)
    assert "backend precision" in create(precision=P.FP32_AUTO)
    tp = AT.indptr.copy()
    last = int(np.nonzero(np.diff(tp))[0][-1])  # the last row of A^T with entries
    tp[last + 1:] -= 1
    assert "nonzero counts differ" in create(tptr=tp)
    h = L.pdhg_backend_create(lp.m, lp.n, None, None, None, None, None, None, None, None, 1, P.FP64, 0)
    assert not h and b"null row pointers" in L.pdhg_last_error()
    if not P.gpu_available():  # valid input gets as far as the device check
        assert "GPU unavailable" in create()
