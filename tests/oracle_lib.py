"""ctypes access to the two CPU checkers (test infrastructure only).

* oracle/liboracle.so        - FP64 restatement of the reference PDHG + safe bound.
* oracle/_ref/libpramana_ref.so - the reference solver's own sources compiled
  from the reference tree (absent when the reference was not available at
  build time; tests that need it skip).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libpramana_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OracleResult(C.Structure):
    _fields_ = [("status", C.c_int), ("restarts", C.c_int), ("iterations", C.c_longlong),
                ("objective", C.c_double), ("dual_objective", C.c_double), ("relative_kkt", C.c_double),
                ("rel_primal", C.c_double), ("rel_dual", C.c_double), ("rel_gap", C.c_double),
                ("primal_weight", C.c_double), ("spectral_norm", C.c_double)]


class OracleTrace(C.Structure):
    _fields_ = [("it", C.c_longlong), ("relP", C.c_double), ("relD", C.c_double), ("gap", C.c_double),
                ("pobj", C.c_double), ("dobj", C.c_double), ("omega", C.c_double), ("restarted", C.c_int),
                ("pad", C.c_int)]


class RefResult(C.Structure):
    _fields_ = [("status", C.c_int), ("restarts", C.c_int), ("iterations", C.c_longlong),
                ("objective", C.c_double), ("dual_objective", C.c_double), ("relative_kkt", C.c_double),
                ("primal_residual", C.c_double), ("dual_residual", C.c_double), ("gap", C.c_double),
                ("primal_weight", C.c_double), ("spectral_norm", C.c_double), ("setup_seconds", C.c_double),
                ("solve_seconds", C.c_double)]


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _lp_args(lp):
    return (lp.m, lp.n, _i(lp.start), _i(lp.index), _d(lp.value), _d(lp.c), _d(lp.l), _d(lp.u), _d(lp.L),
            _d(lp.U))


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE):
            raise RuntimeError(f"{ORACLE} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE)
        L.oracle_last_error.restype = C.c_char_p
        assert L.oracle_result_size() == C.sizeof(OracleResult)
        assert L.oracle_trace_size() == C.sizeof(OracleTrace)
        _oracle = L
    return _oracle


def ref_available():
    return os.path.exists(REF)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF)
        L.ref_last_error.restype = C.c_char_p
        assert L.ref_result_size() == C.sizeof(RefResult)
        _ref = L
    return _ref


def oracle_prepare(lp, ruiz=10, pc=1, rescale=1):
    R, Cs, sc = np.zeros(lp.m), np.zeros(lp.n), np.zeros(3)
    ptr, idx, val = np.zeros(lp.m + 1, np.int32), np.zeros(lp.nnz, np.int32), np.zeros(lp.nnz)
    rc = oracle().oracle_prepare(lp.m, lp.n, _i(lp.start), _i(lp.index), _d(lp.value), _d(lp.c), _d(lp.L),
                                 _d(lp.U), ruiz, pc, rescale, _d(R), _d(Cs), _d(sc), _i(ptr), _i(idx), _d(val))
    assert rc == 0, oracle().oracle_last_error()
    return dict(R=R, C=Cs, beta=sc[0], gamma=sc[1], normA=sc[2], ptr=ptr, idx=idx, val=val)


def oracle_solve(lp, costs=None, tol=1e-6, max_iter=1_000_000, check_every=64, trace_cap=0):
    """Returns (list of result dicts, xs (k x n), ys (k x m), trace list)."""
    k = 1 if costs is None else costs.shape[0]
    costs = None if costs is None else np.ascontiguousarray(costs, dtype=np.float64)
    res = (OracleResult * k)()
    xs, ys = np.zeros((k, lp.n)), np.zeros((k, lp.m))
    tr = (OracleTrace * max(1, trace_cap))()
    tl = C.c_int(0)
    rc = oracle().oracle_solve(*_lp_args(lp), C.c_double(lp.offset), k, _d(costs), C.c_double(tol),
                               C.c_longlong(max_iter), check_every, res, _d(xs), _d(ys), tr if trace_cap else None,
                               trace_cap, C.byref(tl))
    assert rc == 0, oracle().oracle_last_error()
    out = [{f: getattr(res[b], f) for f, _ in OracleResult._fields_} for b in range(k)]
    trace = [{f: getattr(tr[t], f) for f, _ in OracleTrace._fields_} for t in range(tl.value)]
    return out, xs, ys, trace


def oracle_safe_bound(lp, y, c=None, l=None, u=None):
    c = lp.c if c is None else np.ascontiguousarray(c, dtype=np.float64)
    l = lp.l if l is None else np.ascontiguousarray(l, dtype=np.float64)
    u = lp.u if u is None else np.ascontiguousarray(u, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    b, inf, pert = C.c_double(), C.c_int(), C.c_double()
    rc = oracle().oracle_safe_dual_bound(lp.m, lp.n, _i(lp.start), _i(lp.index), _d(lp.value), _d(c), _d(l), _d(u),
                                         _d(lp.L), _d(lp.U), _d(y), C.byref(b), C.byref(inf), C.byref(pert))
    assert rc == 0
    return b.value, inf.value, pert.value


def ref_set_threads(n):
    return ref().ref_set_threads(n)


def ref_solve(lp, tol=1e-6, max_iter=1_000_000, check_every=64):
    r = RefResult()
    x, y = np.zeros(lp.n), np.zeros(lp.m)
    rc = ref().ref_solve_pdhg(*_lp_args(lp), C.c_double(lp.offset), C.c_double(tol), C.c_longlong(max_iter),
                              check_every, C.byref(r), _d(x), _d(y))
    assert rc == 0, ref().ref_last_error()
    return {f: getattr(r, f) for f, _ in RefResult._fields_}, x, y


def ref_solve_batch(lp, costs, tol=1e-6, max_iter=1_000_000):
    costs = np.ascontiguousarray(costs, dtype=np.float64)
    k = costs.shape[0]
    res = (RefResult * k)()
    xs, ys = np.zeros((k, lp.n)), np.zeros((k, lp.m))
    rc = ref().ref_solve_batch_cost(*_lp_args(lp), C.c_double(lp.offset), k, _d(costs), C.c_double(tol),
                                    C.c_longlong(max_iter), res, _d(xs), _d(ys))
    assert rc == 0, ref().ref_last_error()
    return [{f: getattr(res[b], f) for f, _ in RefResult._fields_} for b in range(k)], xs, ys


def ref_safe_bound(lp, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    b, inf, pert = C.c_double(), C.c_int(), C.c_double()
    rc = ref().ref_safe_dual_bound(*_lp_args(lp), _d(y), C.byref(b), C.byref(inf), C.byref(pert))
    assert rc == 0
    return b.value, inf.value, pert.value


def ref_family_bound(lp, y, cost=None):
    c = lp.c if cost is None else np.ascontiguousarray(cost, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    b = C.c_double()
    rc = ref().ref_family_bound(lp.m, lp.n, _i(lp.start), _i(lp.index), _d(lp.value), _d(c), _d(lp.l), _d(lp.u),
                                _d(lp.L), _d(lp.U), C.c_double(lp.offset), _d(y), C.byref(b))
    assert rc == 0, ref().ref_last_error()
    return b.value


class OracleOps:
    """The oracle's serial CpuBackend restatement (src/pdhg/pdhg.cpp:75-214)
    driven one op at a time, with the same vector ids and k-interleaved layout
    as pdhg_amd.Backend."""

    def __init__(self, A_csr, AT_csr, rfac, cfac, k=1):
        L = oracle()
        L.oracle_ops_create.restype = C.c_void_p
        L.oracle_ops_create.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _ip, _ip, _dp, _dp, _dp, C.c_int]
        for name in ("oracle_ops_destroy", "oracle_ops_primal", "oracle_ops_dual"):
            getattr(L, name).argtypes = [C.c_void_p]
        for name in ("oracle_ops_set", "oracle_ops_get"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, _dp]
        for name in ("oracle_ops_spmv", "oracle_ops_spmvT"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.oracle_ops_set_steps.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_ops_halpern.argtypes = [C.c_void_p, C.c_double]
        L.oracle_ops_restart.argtypes = [C.c_void_p, _ip]
        L.oracle_ops_kkt.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_ops_diff2.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        self.L = L
        self.m, self.n, self.k = A_csr.shape[0], A_csr.shape[1], k
        self._keep = [np.ascontiguousarray(a, dtype=t) for a, t in (
            (A_csr.indptr, np.int32), (A_csr.indices, np.int32), (A_csr.data, np.float64),
            (AT_csr.indptr, np.int32), (AT_csr.indices, np.int32), (AT_csr.data, np.float64),
            (rfac, np.float64), (cfac, np.float64))]
        a = self._keep
        self.h = L.oracle_ops_create(self.m, self.n, _i(a[0]), _i(a[1]), _d(a[2]), _i(a[3]), _i(a[4]), _d(a[5]),
                                     _d(a[6]), _d(a[7]), k)
        assert self.h, L.oracle_last_error()

    def close(self):
        if self.h:
            self.L.oracle_ops_destroy(self.h)
            self.h = None

    __del__ = close

    def _len(self, v):
        return (self.m if v >= 8 else self.n) * self.k

    def set(self, v, data):
        a = np.ascontiguousarray(data, dtype=np.float64).ravel()
        assert a.size == self._len(v)
        assert self.L.oracle_ops_set(self.h, v, _d(a)) == 0

    def get(self, v):
        a = np.zeros(self._len(v))
        assert self.L.oracle_ops_get(self.h, v, _d(a)) == 0
        return a

    def spmv(self, i, o):
        self.L.oracle_ops_spmv(self.h, i, o)

    def spmvT(self, i, o):
        self.L.oracle_ops_spmvT(self.h, i, o)

    def set_steps(self, tau, sigma):
        t = np.ascontiguousarray(tau, dtype=np.float64)
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        self.L.oracle_ops_set_steps(self.h, _d(t), _d(s))

    def primal_step(self):
        self.L.oracle_ops_primal(self.h)

    def dual_step(self):
        self.L.oracle_ops_dual(self.h)

    def halpern(self, rho=1.0):
        self.L.oracle_ops_halpern(self.h, rho)

    def restart(self, mask):
        mk = np.ascontiguousarray(mask, dtype=np.int32)
        self.L.oracle_ops_restart(self.h, _i(mk))

    def kkt(self):
        r3, c3 = np.zeros(3 * self.k), np.zeros(3 * self.k)
        self.L.oracle_ops_kkt(self.h, _d(r3), _d(c3))
        return r3, c3

    def diff2(self, a, b):
        out = np.zeros(self.k)
        self.L.oracle_ops_diff2(self.h, a, b, _d(out))
        return out
