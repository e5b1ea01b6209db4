"""ctypes access to the reference solver built with this engine plugged in
(oracle/_ref/libpramana_hip.so: the reference's own C API, solve() pipeline,
crossover, certifier, router and family driver, compiled from the reference
tree by `make -C oracle hip`, with its PDHG entries provided by
optimized-gpu-accelerator_amd/integration/pdhg_hip_adapter.cpp over
libpdhg_hip.so), and the matching CLI oracle/_ref/pramana_hip.  Test
infrastructure only; absent when the reference tree was not there at build
time."""
import ctypes as C
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libpramana_hip.so")
CLI = os.path.join(ROOT, "oracle", "_ref", "pramana_hip")
GOLDEN = os.path.join(ROOT, "tests", "golden")
_lib = None


def available():
    return os.path.exists(LIB) and os.path.exists(CLI)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        L.pramana_last_error.restype = C.c_char_p
        L.pramana_model_read.restype = C.c_void_p
        L.pramana_model_read.argtypes = [C.c_char_p]
        L.pramana_model_free.argtypes = [C.c_void_p]
        L.pramana_solve.restype = C.c_void_p
        L.pramana_solve.argtypes = [C.c_void_p, C.c_char_p]
        L.pramana_result_status.argtypes = [C.c_void_p]
        L.pramana_result_objective.restype = C.c_double
        L.pramana_result_objective.argtypes = [C.c_void_p]
        L.pramana_result_certified.argtypes = [C.c_void_p]
        L.pramana_result_json.restype = C.c_char_p
        L.pramana_result_json.argtypes = [C.c_void_p, C.c_int]
        L.pramana_result_free.argtypes = [C.c_void_p]
        for name in ("pramana_family", "pramana_gpu_info"):
            getattr(L, name).restype = C.c_void_p
        L.pramana_family.argtypes = [C.c_void_p, C.c_char_p]
        L.pramana_free_string.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _take(p):
    if not p:
        raise RuntimeError(lib().pramana_last_error().decode())
    s = C.cast(p, C.c_char_p).value.decode()
    lib().pramana_free_string(p)
    return s


def gpu_info():
    return json.loads(_take(lib().pramana_gpu_info()))


def netlib_path(name):
    return os.path.join(GOLDEN, f"netlib_{name}.mps.gz")


def netlib_optima():
    """Certified optimal objective of each Netlib fixture (the reference's own
    benchmark record; tests/golden/make_netlib_mps.py)."""
    return {k: v["optimum"] for k, v in json.load(open(os.path.join(GOLDEN, "netlib_optima.json"))).items()}


def solve(path, **options):
    """pramana_solve on an MPS file: (status, objective, certified, result JSON)."""
    m = lib().pramana_model_read(os.fsencode(path))
    if not m:
        raise RuntimeError(lib().pramana_last_error().decode())
    try:
        r = lib().pramana_solve(m, json.dumps(options).encode())
        if not r:
            raise RuntimeError(lib().pramana_last_error().decode())
        try:
            out = (lib().pramana_result_status(r), lib().pramana_result_objective(r),
                   lib().pramana_result_certified(r), json.loads(lib().pramana_result_json(r, 0).decode()))
        finally:
            lib().pramana_result_free(r)
    finally:
        lib().pramana_model_free(m)
    return out


def family(path, spec):
    m = lib().pramana_model_read(os.fsencode(path))
    if not m:
        raise RuntimeError(lib().pramana_last_error().decode())
    try:
        return json.loads(_take(lib().pramana_family(m, json.dumps(spec).encode())))
    finally:
        lib().pramana_model_free(m)


def cli(*args, timeout=600):
    """Runs the reference CLI built over this engine; returns (rc, stdout, stderr)."""
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout, p.stderr
