"""pdhg_kkt_noalloc, the allocation-free variant of pdhg_backend_kkt (block
partials in the library's fixed device scratch array), called through the
C-ABI on device arrays this test allocates, against pdhg_backend_kkt on the
same data: bit-identical sums (same kernels, same fixed summation order), on
two calls in a row, the second with a smaller input, so stale partials of
the first call cannot leak into the second."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import pdhg_amd as P
import pdhg_driver as D
from lp_cases import degenerate_costs, degenerate_lp

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceArrays:
    """Device copies of host arrays through the HIP runtime (the test's own
    allocations; the entry point under test allocates nothing)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p

    def get(self, p, count):
        out = np.zeros(count)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), D2H) == 0
        return out

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def case(lp, prec, k, seed):
    pr = O.oracle_prepare(lp)
    A, AT = D.scaled_csr_pair(lp, pr)
    rfac, cfac = pr["beta"] / pr["R"], pr["gamma"] / pr["C"]
    be = P.Backend(A, AT, rfac, cfac, k=k, precision=prec)
    rng = np.random.default_rng(seed)
    d = D.case_data(lp, pr, lp.c[None, :] + rng.uniform(0, 0.1, (k, lp.n)))
    vecs = {D.C: d["c"], D.LO: d["lo"], D.UP: d["up"], D.RL: d["rl"], D.RU: d["ru"],
            D.XP: rng.standard_normal((lp.n, k)), D.ATYP: rng.standard_normal((lp.n, k)),
            D.YP: rng.standard_normal((lp.m, k)), D.AXP: rng.standard_normal((lp.m, k))}
    for v, a in vecs.items():
        be.set(v, a)
    r3, c3 = be.kkt()
    # the same data as the backend holds (its FP32 rounding included)
    dt = np.float64 if prec == P.FP64 else np.float32
    host = {v: be.get(v).astype(dt) for v in vecs}
    return host, rfac, cfac, r3, c3


def call_noalloc(dev, lp, prec, k, host, rfac, cfac):
    a = {v: dev.put(x) for v, x in host.items()}
    rows, cols = dev.empty(3 * k * 8), dev.empty(3 * k * 8)
    rc = P.lib().pdhg_kkt_noalloc(lp.m, lp.n, k, prec, a[D.AXP], a[D.RL], a[D.RU], a[D.YP], dev.put(rfac),
                                  a[D.XP], a[D.C], a[D.ATYP], a[D.LO], a[D.UP], dev.put(cfac), rows, cols, None)
    assert rc == 0, P.lib().pdhg_last_error()
    dev.sync()
    return dev.get(rows, 3 * k), dev.get(cols, 3 * k)


def test_kkt_noalloc_equals_backend_kkt_twice(gpu):
    dev = DeviceArrays()
    try:
        big, _ = degenerate_lp(seed=5, m=600, n=1200)  # free, fixed, empty and ranged rows and columns
        small = P.gen_random_lp(3, 40, 90)
        for lp, prec, k in ((big, P.FP64, 64), (small, P.FP32, 3), (small, P.FP64, 1)):
            host, rfac, cfac, r3, c3 = case(lp, prec, k, seed=k)
            g3, gc3 = call_noalloc(dev, lp, prec, k, host, rfac, cfac)
            assert np.array_equal(g3, r3), (prec, k, g3, r3)
            assert np.array_equal(gc3, c3), (prec, k, gc3, c3)
            assert np.all(np.isfinite(g3)) and np.any(g3 != 0)
    finally:
        dev.free()


def test_kkt_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    nul = C.c_void_p(0)
    rc = P.lib().pdhg_kkt_noalloc(10, 10, 65, P.FP64, *([nul] * 13), None)
    assert rc == -1 and b"1..64" in P.lib().pdhg_last_error()
