"""ctypes binding of libpdhg_hip.so (include/pdhg_hip.h).

Thin Python mirror of the engine API for tests and benchmarks: the reference's
solvePdhg / solvePdhgBatch / gpuAvailable / gpuInfoJson (src/pdhg/pdhg.h:73-80),
the Backend op set (src/pdhg/pdhg.cpp:51-70), the safe-bound certifier and the
synthetic LP generators.  All arithmetic happens in the library; this module
only moves numpy arrays across the C-ABI.  There is no CPU fallback: solve
entries raise when the HIP device is missing.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdhg_hip.so")

OPTIMAL, TIME_LIMIT, ITERATION_LIMIT = 1, 5, 6
FP32, FP64, FP32_AUTO = 0, 1, 2
(V_X, V_XP, V_X0, V_ATY, V_ATYP, V_C, V_LO, V_UP,
 V_Y, V_YP, V_Y0, V_AX, V_AXP, V_AX0, V_RL, V_RU) = range(16)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class LpStruct(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("col_start", _ip), ("row_index", _ip), ("value", _dp),
                ("cost", _dp), ("col_lower", _dp), ("col_upper", _dp), ("row_lower", _dp), ("row_upper", _dp),
                ("obj_offset", C.c_double)]


class CaseStruct(C.Structure):
    _fields_ = [("cost", _dp), ("col_lower", _dp), ("col_upper", _dp), ("row_lower", _dp), ("row_upper", _dp),
                ("x0", _dp), ("y0", _dp), ("primal_weight", C.c_double)]


class Options(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_longlong), ("check_every", C.c_int),
                ("use_gpu", C.c_int), ("ruiz_iterations", C.c_int), ("pock_chambolle", C.c_int),
                ("rescale", C.c_int), ("reflection", C.c_double), ("log_interval", C.c_int),
                ("precision", C.c_int), ("use_graph", C.c_int), ("lanes_per_row", C.c_int), ("row_map", C.c_int),
                ("xcd_remap", C.c_int), ("time_limit", C.c_double), ("device", C.c_int),
                ("kernel_timing_reps", C.c_int), ("compact", C.c_int), ("progress_seconds", C.c_double),
                ("warm_fallback", C.c_longlong), ("should_stop", C.c_void_p), ("stop_user", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_longlong), ("restarts", C.c_int), ("primal_residual", C.c_double),
                ("dual_residual", C.c_double), ("gap", C.c_double), ("primal_weight", C.c_double),
                ("setup_seconds", C.c_double), ("solve_seconds", C.c_double), ("kernel_seconds", C.c_double),
                ("transfer_seconds", C.c_double), ("bytes_uploaded", C.c_longlong),
                ("bytes_downloaded", C.c_longlong), ("kernel_launches", C.c_longlong),
                ("graph_launches", C.c_longlong), ("spectral_norm", C.c_double), ("batch", C.c_int),
                ("precision", C.c_int), ("fp64_switch_iteration", C.c_longlong),
                ("spmv_bytes_per_iter", C.c_double), ("iter_bytes", C.c_double),
                ("col_kernel_ms", C.c_double), ("row_kernel_ms", C.c_double), ("col_kernel_bytes", C.c_double),
                ("row_kernel_bytes", C.c_double),
                ("trace", _dp), ("trace_capacity", C.c_int), ("trace_length", C.c_int),
                ("backend", C.c_char * 32), ("device_name", C.c_char * 64),
                ("autotuned", C.c_int), ("col_nnz_per_lane", C.c_int), ("col_row_map", C.c_int), ("col_xcd", C.c_int),
                ("row_nnz_per_lane", C.c_int), ("row_row_map", C.c_int), ("row_xcd", C.c_int),
                ("tune_seconds", C.c_double), ("col_rows_per_wave", C.c_int), ("row_rows_per_wave", C.c_int),
                ("warm_fallbacks", C.c_int), ("iter_kernel_ms", C.c_double)]


class ResultStruct(C.Structure):
    _fields_ = [("status", C.c_int), ("restarts", C.c_int), ("iterations", C.c_longlong),
                ("objective", C.c_double), ("dual_objective", C.c_double), ("relative_kkt", C.c_double),
                ("rel_primal", C.c_double), ("rel_dual", C.c_double), ("rel_gap", C.c_double),
                ("primal_weight", C.c_double), ("x", _dp), ("row_dual", _dp)]


# Every symbol include/pdhg_hip.h declares (checked by the CPU test suite).
EXPORTS = [
    "pdhg_last_error", "pdhg_default_options", "pdhg_gpu_available", "pdhg_gpu_info_json",
    "pdhg_solve", "pdhg_solve_batch", "pdhg_solve_packed", "pdhg_prepare",
    "pdhg_backend_create", "pdhg_backend_destroy", "pdhg_backend_set", "pdhg_backend_get",
    "pdhg_backend_spmv", "pdhg_backend_spmvT", "pdhg_backend_set_steps", "pdhg_backend_primal_step",
    "pdhg_backend_dual_step", "pdhg_backend_halpern", "pdhg_backend_restart", "pdhg_backend_kkt",
    "pdhg_backend_diff2", "pdhg_backend_time_spmv", "pdhg_backend_spmv_layout",
    "pdhg_backend_set_spmv_layout", "pdhg_kkt_noalloc",
    "pdhg_safe_dual_bound", "pdhg_certify", "pdhg_certify_batch",
    "pdhg_gen_random_lp", "pdhg_gen_refinery", "pdhg_gen_family", "pdhg_family_costs",
    "pdhg_gen_calibration", "pdhg_read_mps", "pdhg_model_view", "pdhg_model_free",
    "pdhg_pod_basis", "pdhg_decode_bf16", "pdhg_mlp_create", "pdhg_mlp_destroy", "pdhg_mlp_param_count",
    "pdhg_mlp_get_params", "pdhg_mlp_set_params", "pdhg_mlp_forward", "pdhg_mlp_loss_grad", "pdhg_mlp_train",
]

_lib = None


def lib() -> C.CDLL:
    """Loads libpdhg_hip.so (built in-tree by `make -C optimized-gpu-accelerator_amd`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {HERE}`")
        L = C.CDLL(LIB_PATH)
        L.pdhg_last_error.restype = C.c_char_p
        L.pdhg_gpu_available.restype = C.c_int
        L.pdhg_gen_random_lp.restype = C.c_void_p
        L.pdhg_gen_random_lp.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_double, C.c_double]
        L.pdhg_gen_refinery.restype = C.c_void_p
        L.pdhg_gen_refinery.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.pdhg_read_mps.restype = C.c_void_p
        L.pdhg_read_mps.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        L.pdhg_gen_calibration.restype = C.c_void_p
        L.pdhg_gen_calibration.argtypes = [C.c_uint64, C.c_int]
        L.pdhg_gen_family.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_double, _dp, _dp]
        L.pdhg_family_costs.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.pdhg_model_view.argtypes = [C.c_void_p, C.POINTER(LpStruct), C.POINTER(_dp), C.POINTER(_dp)]
        L.pdhg_model_free.argtypes = [C.c_void_p]
        L.pdhg_solve.argtypes = [C.POINTER(LpStruct), C.POINTER(Options), C.POINTER(ResultStruct), C.POINTER(Stats)]
        L.pdhg_solve_batch.argtypes = [C.POINTER(LpStruct), C.POINTER(CaseStruct), C.c_int, C.POINTER(Options),
                                       C.POINTER(ResultStruct), C.POINTER(Stats)]
        L.pdhg_solve_packed.argtypes = [C.POINTER(LpStruct), C.c_int, C.POINTER(Options), C.POINTER(ResultStruct),
                                        C.POINTER(Stats)]
        L.pdhg_prepare.argtypes = [C.POINTER(LpStruct), C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip, _dp]
        L.pdhg_safe_dual_bound.argtypes = [C.POINTER(LpStruct), _dp, _dp, _ip, _dp]
        L.pdhg_certify.argtypes = [C.POINTER(LpStruct), _dp, _dp, C.c_double, _dp]
        L.pdhg_certify_batch.argtypes = [C.POINTER(LpStruct), C.POINTER(CaseStruct), C.c_int, _dp, _dp, _dp,
                                         C.c_int, _dp]
        L.pdhg_backend_create.restype = C.c_void_p
        L.pdhg_backend_create.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _ip, _ip, _dp, _dp, _dp, C.c_int,
                                          C.c_int, C.c_int]
        L.pdhg_backend_destroy.argtypes = [C.c_void_p]
        for name in ("pdhg_backend_set", "pdhg_backend_get"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, _dp]
        for name in ("pdhg_backend_spmv", "pdhg_backend_spmvT"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.pdhg_backend_set_steps.argtypes = [C.c_void_p, _dp, _dp]
        L.pdhg_backend_primal_step.argtypes = [C.c_void_p]
        L.pdhg_backend_dual_step.argtypes = [C.c_void_p]
        L.pdhg_backend_halpern.argtypes = [C.c_void_p, C.c_double]
        L.pdhg_backend_restart.argtypes = [C.c_void_p, _ip]
        L.pdhg_backend_kkt.argtypes = [C.c_void_p, _dp, _dp]
        L.pdhg_backend_diff2.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.pdhg_backend_time_spmv.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.pdhg_backend_spmv_layout.argtypes = [C.c_void_p, C.c_int, _ip]
        L.pdhg_backend_set_spmv_layout.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.pdhg_kkt_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4 + [C.c_void_p] + \
            [C.c_void_p] * 5 + [C.c_void_p] * 3 + [C.c_void_p]
        L.pdhg_pod_basis.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.pdhg_decode_bf16.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp, C.c_int, _dp]
        L.pdhg_mlp_create.restype = C.c_void_p
        L.pdhg_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int]
        L.pdhg_mlp_destroy.argtypes = [C.c_void_p]
        L.pdhg_mlp_param_count.argtypes = [C.c_void_p]
        L.pdhg_mlp_get_params.argtypes = [C.c_void_p, _dp]
        L.pdhg_mlp_set_params.argtypes = [C.c_void_p, _dp]
        L.pdhg_mlp_forward.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
        L.pdhg_mlp_loss_grad.argtypes = [C.c_void_p, _dp, _dp, C.c_int, _dp, _dp]
        L.pdhg_mlp_train.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int, C.c_double, _dp]
        _lib = L
    return _lib


def _err() -> str:
    return lib().pdhg_last_error().decode()


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {_err()}")


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def gpu_available() -> bool:
    return bool(lib().pdhg_gpu_available())


def gpu_info() -> dict:
    buf = C.create_string_buffer(2048)
    _check(lib().pdhg_gpu_info_json(buf, 2048), "gpu_info")
    return json.loads(buf.value.decode())


def default_options(**kw) -> Options:
    o = Options()
    lib().pdhg_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


@dataclass
class LP:
    """LP in minimization form with a CSC matrix (numpy arrays, owned)."""
    m: int
    n: int
    start: np.ndarray
    index: np.ndarray
    value: np.ndarray
    c: np.ndarray
    l: np.ndarray
    u: np.ndarray
    L: np.ndarray
    U: np.ndarray
    offset: float = 0.0
    xstar: np.ndarray | None = None
    ystar: np.ndarray | None = None
    _model: object = None  # generator handle (family maps need it)

    def __post_init__(self):
        self.start = np.ascontiguousarray(self.start, dtype=np.int32)
        self.index = np.ascontiguousarray(self.index, dtype=np.int32)
        for f in ("value", "c", "l", "u", "L", "U"):
            setattr(self, f, np.ascontiguousarray(getattr(self, f), dtype=np.float64))

    @property
    def nnz(self) -> int:
        return int(self.start[-1])

    def struct(self) -> LpStruct:
        s = LpStruct(self.m, self.n, _i(self.start), _i(self.index), _d(self.value), _d(self.c), _d(self.l),
                     _d(self.u), _d(self.L), _d(self.U), float(self.offset))
        s._keep = self  # keep arrays alive
        return s

    def scipy_csc(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.value, self.index, self.start), shape=(self.m, self.n))

    def __del__(self):
        if self._model is not None and _lib is not None:
            _lib.pdhg_model_free(self._model)
            self._model = None


def _from_model(h) -> LP:
    if not h:
        raise RuntimeError(f"generator: {_err()}")
    s = LpStruct()
    xs, ys = _dp(), _dp()
    _check(lib().pdhg_model_view(h, C.byref(s), C.byref(xs), C.byref(ys)), "model_view")
    nnz = s.col_start[s.n]

    def arr(p, k, dt=np.float64):
        return np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True) if k > 0 else np.zeros(0, dt)

    lp = LP(s.m, s.n, arr(s.col_start, s.n + 1, np.int32), arr(s.row_index, nnz, np.int32), arr(s.value, nnz),
            arr(s.cost, s.n), arr(s.col_lower, s.n), arr(s.col_upper, s.n), arr(s.row_lower, s.m),
            arr(s.row_upper, s.m), s.obj_offset)
    lp.xstar = arr(xs, s.n) if xs else None
    lp.ystar = arr(ys, s.m) if ys else None
    lp._model = h
    return lp


def gen_random_lp(seed: int, m: int, n: int, nnz_per_col: float = 4.0, row_sigma: float = 0.0) -> LP:
    """Random LP with a known optimal pair (x*, y*) (configs of 1000 x 2000 and 5e4 x 1e5)."""
    return _from_model(lib().pdhg_gen_random_lp(seed, m, n, nnz_per_col, row_sigma))


def gen_refinery(seed: int, periods=12, streams=40, crudes=8, products=36, qualities=30, decades=3.0) -> LP:
    """Refinery-style multi-period blending LP (about 5.6e5 nonzeros at the defaults)."""
    return _from_model(lib().pdhg_gen_refinery(seed, periods, streams, crudes, products, qualities, decades))


def read_mps(path: str):
    """An MPS file (free or fixed format, plain or gzip) as an LP in minimisation form,
    and its sense (1, or -1 for OBJSENSE MAX: costs and offset negated)."""
    sense = C.c_int(1)
    h = lib().pdhg_read_mps(os.fsencode(path), C.byref(sense))
    if not h:
        raise RuntimeError(f"read_mps: {_err()}")
    return _from_model(h), sense.value


def gen_calibration(seed: int, s: int) -> LP:
    return _from_model(lib().pdhg_gen_calibration(seed, s))


def gen_family(base: LP, seed: int, k: int, rng: float = 0.3):
    """k price cases of a refinery model: returns (thetas k x 4, costs k x n)."""
    th = np.zeros((k, 4))
    cs = np.zeros((k, base.n))
    _check(lib().pdhg_gen_family(base._model, seed, k, rng, _d(th), _d(cs)), "gen_family")
    return th, cs


def family_costs(base: LP, thetas: np.ndarray) -> np.ndarray:
    th = np.ascontiguousarray(thetas, dtype=np.float64).reshape(-1, 4)
    cs = np.zeros((th.shape[0], base.n))
    _check(lib().pdhg_family_costs(base._model, th.shape[0], _d(th), _d(cs)), "family_costs")
    return cs


@dataclass
class Result:
    status: int
    iterations: int
    objective: float
    dual_objective: float
    relative_kkt: float
    rel_primal: float
    rel_dual: float
    rel_gap: float
    restarts: int
    primal_weight: float
    x: np.ndarray
    y: np.ndarray


def _result(r: ResultStruct, x, y) -> Result:
    return Result(r.status, r.iterations, r.objective, r.dual_objective, r.relative_kkt, r.rel_primal, r.rel_dual,
                  r.rel_gap, r.restarts, r.primal_weight, x, y)


def stats_dict(st: Stats) -> dict:
    d = {f: getattr(st, f) for f, _ in Stats._fields_ if f not in ("trace", "backend", "device_name")}
    d["backend"] = st.backend.decode()
    d["device"] = st.device_name.decode()
    return d


def solve(lp: LP, opts: Options | None = None, trace_capacity: int = 0):
    """solvePdhg: returns (Result, stats dict, trace array (k x 8) or None)."""
    opts = opts or default_options()
    x, y = np.zeros(lp.n), np.zeros(lp.m)
    r = ResultStruct()
    r.x, r.row_dual = _d(x), _d(y)
    st = Stats()
    tr = None
    if trace_capacity > 0:
        tr = np.zeros((trace_capacity, 8))
        st.trace, st.trace_capacity = _d(tr), trace_capacity
    s = lp.struct()
    _check(lib().pdhg_solve(C.byref(s), C.byref(opts), C.byref(r), C.byref(st)), "pdhg_solve")
    return _result(r, x, y), stats_dict(st), (tr[: st.trace_length] if tr is not None else None)


def solve_batch(lp: LP, costs: np.ndarray | None = None, opts: Options | None = None, k: int | None = None,
                x0: np.ndarray | None = None, y0: np.ndarray | None = None, trace_capacity: int = 0,
                primal_weight=None):
    """solvePdhgBatch over cost cases (k x n) with optional warm starts (k x n, k x m) and
    initial primal weights (scalar or k values; None keeps the reference's)."""
    opts = opts or default_options()
    if costs is not None:
        costs = np.ascontiguousarray(costs, dtype=np.float64)
        k = costs.shape[0]
    assert k is not None
    xs, ys = np.zeros((k, lp.n)), np.zeros((k, lp.m))
    cases = (CaseStruct * k)()
    if x0 is not None:
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
    if y0 is not None:
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
    for b in range(k):
        if costs is not None:
            cases[b].cost = _d(costs[b])
        if x0 is not None:
            cases[b].x0 = _d(x0[b])
        if y0 is not None:
            cases[b].y0 = _d(y0[b])
        if primal_weight is not None:
            cases[b].primal_weight = float(np.broadcast_to(np.asarray(primal_weight, dtype=np.float64), (k,))[b])
    res = (ResultStruct * k)()
    for b in range(k):
        res[b].x, res[b].row_dual = _d(xs[b]), _d(ys[b])
    st = Stats()
    tr = None
    if trace_capacity > 0:
        tr = np.zeros((trace_capacity, 8))
        st.trace, st.trace_capacity = _d(tr), trace_capacity
    s = lp.struct()
    _check(lib().pdhg_solve_batch(C.byref(s), cases, k, C.byref(opts), res, C.byref(st)), "pdhg_solve_batch")
    out = [_result(res[b], xs[b], ys[b]) for b in range(k)]
    return out, stats_dict(st), (tr[: st.trace_length] if tr is not None else None)


def prepare(lp: LP, ruiz: int = 10, pc: int = 1, rescale: int = 1) -> dict:
    """Host set-up of the engine: scaling factors, beta, gamma, ||A||, scaled CSR."""
    R, Cs, sc = np.zeros(lp.m), np.zeros(lp.n), np.zeros(3)
    ptr, idx, val = np.zeros(lp.m + 1, np.int32), np.zeros(lp.nnz, np.int32), np.zeros(lp.nnz)
    s = lp.struct()
    _check(lib().pdhg_prepare(C.byref(s), ruiz, pc, rescale, _d(R), _d(Cs), _d(sc), _i(ptr), _i(idx), _d(val)),
           "prepare")
    return dict(R=R, C=Cs, beta=sc[0], gamma=sc[1], normA=sc[2], ptr=ptr, idx=idx, val=val)


def safe_dual_bound(lp: LP, y: np.ndarray):
    y = np.ascontiguousarray(y, dtype=np.float64)
    b, inf, pert = C.c_double(), C.c_int(), C.c_double()
    s = lp.struct()
    _check(lib().pdhg_safe_dual_bound(C.byref(s), _d(y), C.byref(b), C.byref(inf), C.byref(pert)), "safe_bound")
    return b.value, inf.value, pert.value


def certify(lp: LP, x, y, objective: float) -> dict:
    out = np.zeros(4)
    x = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    s = lp.struct()
    _check(lib().pdhg_certify(C.byref(s), _d(x), _d(y), float(objective), _d(out)), "certify")
    return {"bound": out[0], "rel_gap": out[1], "primal_violation": out[2], "accepted": bool(out[3])}


def certify_batch(lp: LP, costs, xs, ys, objectives, threads: int = 16) -> np.ndarray:
    """Family acceptance for k cost cases; returns k x 4 (bound, rel gap, primal viol, accepted)."""
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    k = ys.shape[0]
    cases = (CaseStruct * k)()
    costs = None if costs is None else np.ascontiguousarray(costs, dtype=np.float64)
    if costs is not None:
        for b in range(k):
            cases[b].cost = _d(costs[b])
    xs = None if xs is None else np.ascontiguousarray(xs, dtype=np.float64)
    obj = np.ascontiguousarray(objectives, dtype=np.float64)
    out = np.zeros((k, 4))
    s = lp.struct()
    _check(lib().pdhg_certify_batch(C.byref(s), cases, k, _d(xs), _d(ys), _d(obj), threads, _d(out)),
           "certify_batch")
    return out


class Backend:
    """The reference Backend op set (src/pdhg/pdhg.cpp:51-70) on the GPU."""

    def __init__(self, A_csr, AT_csr, rfac, cfac, k=1, precision=FP32, device=0):
        self.m, self.n, self.k = A_csr.shape[0], A_csr.shape[1], k
        self._keep = []

        def ints(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            self._keep.append(a)
            return _i(a)

        def dbl(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return _d(a)

        self.h = lib().pdhg_backend_create(self.m, self.n, ints(A_csr.indptr), ints(A_csr.indices), dbl(A_csr.data),
                                           ints(AT_csr.indptr), ints(AT_csr.indices), dbl(AT_csr.data), dbl(rfac),
                                           dbl(cfac), k, precision, device)
        if not self.h:
            raise RuntimeError(f"backend_create: {_err()}")

    def close(self):
        if self.h:
            lib().pdhg_backend_destroy(self.h)
            self.h = None

    __del__ = close

    def _len(self, v):
        return (self.m if v >= V_Y else self.n) * self.k

    def set(self, v, data):
        a = np.ascontiguousarray(data, dtype=np.float64).ravel()
        assert a.size == self._len(v)
        _check(lib().pdhg_backend_set(self.h, v, _d(a)), "set")

    def get(self, v):
        a = np.zeros(self._len(v))
        _check(lib().pdhg_backend_get(self.h, v, _d(a)), "get")
        return a

    def spmv(self, i, o):
        _check(lib().pdhg_backend_spmv(self.h, i, o), "spmv")

    def spmvT(self, i, o):
        _check(lib().pdhg_backend_spmvT(self.h, i, o), "spmvT")

    def set_steps(self, tau, sigma):
        t = np.ascontiguousarray(tau, dtype=np.float64)
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        _check(lib().pdhg_backend_set_steps(self.h, _d(t), _d(s)), "set_steps")

    def primal_step(self):
        _check(lib().pdhg_backend_primal_step(self.h), "primal")

    def dual_step(self):
        _check(lib().pdhg_backend_dual_step(self.h), "dual")

    def halpern(self, rho=1.0):
        _check(lib().pdhg_backend_halpern(self.h, rho), "halpern")

    def restart(self, mask):
        mk = np.ascontiguousarray(mask, dtype=np.int32)
        _check(lib().pdhg_backend_restart(self.h, _i(mk)), "restart")

    def kkt(self):
        r3, c3 = np.zeros(3 * self.k), np.zeros(3 * self.k)
        _check(lib().pdhg_backend_kkt(self.h, _d(r3), _d(c3)), "kkt")
        return r3, c3

    def diff2(self, a, b):
        out = np.zeros(self.k)
        _check(lib().pdhg_backend_diff2(self.h, a, b, _d(out)), "diff2")
        return out

    def time_spmv(self, transpose: bool, reps: int = 100) -> float:
        ms = C.c_double()
        _check(lib().pdhg_backend_time_spmv(self.h, int(transpose), reps, C.byref(ms)), "time_spmv")
        return ms.value

    LAYOUTS = ("csr", "sell", "tiles")

    def spmv_layout(self, transpose: bool) -> str:
        """Storage of the SpMV: "csr" (tuned row mapping), "sell" (sliced ELLPACK, width 1) or
        "tiles" (CSR rows in LDS-staged tiles, width 64)."""
        v = C.c_int()
        _check(lib().pdhg_backend_spmv_layout(self.h, int(transpose), C.byref(v)), "spmv_layout")
        return self.LAYOUTS[v.value]

    def set_spmv_layout(self, transpose: bool, layout: str) -> None:
        """Forces a storage with its default plan: "csr" (one row per wavefront), "sell" (width 1
        only) or "tiles" (8-row tiles, width 64 only)."""
        _check(lib().pdhg_backend_set_spmv_layout(self.h, int(transpose), self.LAYOUTS.index(layout)),
               "set_spmv_layout")


# ---------------------------------------------------------------------------
# Learned warm start: POD basis, BF16-MFMA decode and the MLP
# ---------------------------------------------------------------------------
def pod_basis(Z: np.ndarray, rank: int):
    """POD of snapshots Z (k x d): returns (mean d, basis d x rank, coeffs k x rank, energy rank, used)."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    k, d = Z.shape
    mean, basis = np.zeros(d), np.zeros((d, rank))
    coeffs, energy = np.zeros((k, rank)), np.zeros(rank)
    used = lib().pdhg_pod_basis(_d(Z), k, d, rank, _d(mean), _d(basis), _d(coeffs), _d(energy))
    if used < 0:
        raise RuntimeError(f"pod_basis: {_err()}")
    return mean, basis, coeffs, energy, used


def decode_bf16(basis: np.ndarray, mean: np.ndarray, coeffs: np.ndarray) -> np.ndarray:
    """mean + coeffs @ basis.T on BF16 MFMA tiles (FP32 accumulation)."""
    basis = np.ascontiguousarray(basis, dtype=np.float64)
    mean = np.ascontiguousarray(mean, dtype=np.float64)
    coeffs = np.ascontiguousarray(np.atleast_2d(coeffs), dtype=np.float64)
    d, r = basis.shape
    out = np.zeros((coeffs.shape[0], d))
    _check(lib().pdhg_decode_bf16(_d(basis), _d(mean), d, r, _d(coeffs), coeffs.shape[0], _d(out)), "decode")
    return out


class Mlp:
    """in -> hidden (GELU) -> hidden (GELU) -> out, trained on the GPU (BF16 MFMA, FP32 accumulate)."""

    def __init__(self, n_in=4, hidden=128, n_out=32, seed=0, device=0):
        self.shape = (n_in, hidden, n_out)
        self.h = lib().pdhg_mlp_create(n_in, hidden, n_out, seed, device)
        if not self.h:
            raise RuntimeError(f"mlp_create: {_err()}")
        self.n_params = lib().pdhg_mlp_param_count(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().pdhg_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params)
        _check(lib().pdhg_mlp_get_params(self.h, _d(p)), "mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.size == self.n_params
        _check(lib().pdhg_mlp_set_params(self.h, _d(p)), "mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.zeros((X.shape[0], self.shape[2]))
        _check(lib().pdhg_mlp_forward(self.h, _d(X), X.shape[0], _d(Y)), "mlp_forward")
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        loss, g = C.c_double(), np.zeros(self.n_params)
        _check(lib().pdhg_mlp_loss_grad(self.h, _d(X), _d(T), X.shape[0], C.byref(loss), _d(g)), "mlp_loss_grad")
        return loss.value, g

    def train(self, X, T, steps=2000, lr=1e-3):
        X = np.ascontiguousarray(X, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        losses = np.zeros(steps)
        _check(lib().pdhg_mlp_train(self.h, _d(X), _d(T), X.shape[0], steps, lr, _d(losses)), "mlp_train")
        return losses


def solve_packed(lps: list, opts: Options | None = None):
    """Unrelated LPs solved together by the persistent packed kernels: (results, stats)."""
    opts = opts or default_options()
    k = len(lps)
    arr = (LpStruct * k)()
    keep = []
    for b, lp in enumerate(lps):
        s = lp.struct()
        keep.append(s)
        arr[b] = s
    xs = [np.zeros(lp.n) for lp in lps]
    ys = [np.zeros(lp.m) for lp in lps]
    res = (ResultStruct * k)()
    for b in range(k):
        res[b].x, res[b].row_dual = _d(xs[b]), _d(ys[b])
    st = Stats()
    _check(lib().pdhg_solve_packed(arr, k, C.byref(opts), res, C.byref(st)), "pdhg_solve_packed")
    return [_result(res[b], xs[b], ys[b]) for b in range(k)], stats_dict(st)


def gen_uneven_batch(seed: int, count: int, nnz_lo: float = 2e3, nnz_hi: float = 2e5, upper: float = 2.0) -> list:
    """`count` random LPs of the 1000 x 2000-style generator (m = n / 2, 4 entries per
    column) with nonzero counts log-uniform in [nnz_lo, nnz_hi].  Each gets the
    column box x <= `upper` (None keeps x >= 0 only).  The generator's optimum has
    x* <= 1, so a box of 2 leaves (x*, y*) optimal; it gives every column a
    finite bound, without which the safe dual bound of a PDHG dual is -inf as
    soon as one reduced cost is slightly negative."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        nnz = float(np.exp(rng.uniform(np.log(nnz_lo), np.log(nnz_hi))))
        n = max(8, int(round(nnz / 4)))
        lp = gen_random_lp(seed * 100003 + b, max(4, n // 2), n, 4.0, 0.0)
        if upper is not None:
            assert lp.xstar is None or float(lp.xstar.max()) < upper
            lp.u = np.full(lp.n, float(upper))
        out.append(lp)
    return out
