"""The op-level device backend (the reference's Backend vtable,
src/pdhg/pdhg.cpp:51-70, exported as pdhg_backend_*) against the oracle's
serial CpuBackend restatement (src/pdhg/pdhg.cpp:75-214), op by op and over
whole solves driven by the reference's own host loop (tests/pdhg_driver.py).

Per op, both backends start from the same state (the device's FP32 rounding
of it is copied back into the oracle), so each comparison sees one op's
arithmetic only.  Bars: copies (restart) exact; elementwise ops within a few
units of rounding of the operand magnitudes (FP64: 2^-52, FP32: 2^-23);
SpMV within the row's sum of |a_ij x_j| times 64 units of rounding (FP32
accumulates in FP32); KKT and diff2 partial sums accumulate in FP64 on both
sides (from the same FP32 or FP64 operands), so they agree to 1e-12
relative."""
import numpy as np
import pytest

import oracle_lib as O
import pdhg_amd as P
import pdhg_driver as D
from lp_cases import degenerate_costs, degenerate_lp

pytestmark = pytest.mark.gpu

EPS = {P.FP64: 2.0 ** -52, P.FP32: 2.0 ** -23}


def make_pair(prec, k, seed=1):
    lp, _ = degenerate_lp(seed=5, m=200, n=400)  # free, fixed, empty and ranged rows / columns
    pr = O.oracle_prepare(lp)
    A, AT = D.scaled_csr_pair(lp, pr)
    rfac, cfac = pr["beta"] / pr["R"], pr["gamma"] / pr["C"]
    be = P.Backend(A, AT, rfac, cfac, k=k, precision=prec)
    ora = O.OracleOps(A, AT, rfac, cfac, k=k)
    rng = np.random.default_rng(seed)
    d = D.case_data(lp, pr, degenerate_costs(lp, k))
    state = {D.C: d["c"], D.LO: d["lo"], D.UP: d["up"], D.RL: d["rl"], D.RU: d["ru"]}
    for v in (D.X, D.XP, D.X0, D.ATY, D.ATYP):
        state[v] = rng.standard_normal((lp.n, k))
    for v in (D.Y, D.YP, D.Y0, D.AX, D.AXP, D.AX0):
        state[v] = rng.standard_normal((lp.m, k))
    for v, a in state.items():
        be.set(v, a)
        ora.set(v, be.get(v))  # the oracle takes the device's rounding of the inputs
    return lp, A, AT, be, ora


def sync(be, ora):
    for v in range(16):
        be.set(v, ora.get(v))
        ora.set(v, be.get(v))


def close(got, want, scale, units, eps):
    bad = np.abs(got - want) > units * eps * scale + 1e-300
    assert not bad.any(), (np.flatnonzero(bad)[:5], got[bad][:5], want[bad][:5])


@pytest.mark.parametrize("prec,k", [(P.FP64, 1), (P.FP64, 8), (P.FP64, 64), (P.FP32, 8), (P.FP32, 64)])
def test_backend_ops_match_oracle_op_by_op(gpu, prec, k):
    lp, A, AT, be, ora = make_pair(prec, k)
    eps = EPS[prec]
    absA, absAT = abs(A), abs(AT)

    def op_spmv(transpose, vin, vout):
        sync(be, ora)
        (be.spmvT if transpose else be.spmv)(vin, vout)
        (ora.spmvT if transpose else ora.spmv)(vin, vout)
        M = absAT if transpose else absA
        scale = (M @ np.abs(ora.get(vin)).reshape(M.shape[1], k)).ravel()
        close(be.get(vout), ora.get(vout), scale, 64, eps)

    rng = np.random.default_rng(7)
    tau, sigma = rng.uniform(0.1, 2.0, k), rng.uniform(0.1, 2.0, k)
    be.set_steps(tau, sigma)
    ora.set_steps(tau, sigma)
    op_spmv(True, D.Y, D.ATY)
    # primal step: x+ = clamp(x - tau (c - aty), l, u)
    sync(be, ora)
    be.primal_step()
    ora.primal_step()
    x, c, aty = (ora.get(v).reshape(lp.n, k) for v in (D.X, D.C, D.ATY))
    close(be.get(D.XP), ora.get(D.XP), (np.abs(x) + tau * (np.abs(c) + np.abs(aty))).ravel(), 4, eps)
    op_spmv(False, D.XP, D.AXP)
    # reflected dual step (src/pdhg/pdhg.cpp:126-137)
    sync(be, ora)
    be.dual_step()
    ora.dual_step()
    y, axp, ax = (ora.get(v).reshape(lp.m, k) for v in (D.Y, D.AXP, D.AX))
    rl, ru = ora.get(D.RL).reshape(lp.m, k), ora.get(D.RU).reshape(lp.m, k)
    fb = lambda v: np.where(np.isfinite(v), np.abs(v), 0.0)  # noqa: E731
    close(be.get(D.YP), ora.get(D.YP),
          (np.abs(y) + sigma * (2 * np.abs(axp) + np.abs(ax) + fb(rl) + fb(ru))).ravel(), 8, eps)
    op_spmv(True, D.YP, D.ATYP)
    # KKT partial sums and fixed-point norms: FP64 sums on both sides
    sync(be, ora)
    r3g, c3g = be.kkt()
    r3o, c3o = ora.kkt()
    # The dual-objective and objective parts are signed sums: their bar is
    # relative to the sum of the terms' magnitudes.
    yp = ora.get(D.YP).reshape(lp.m, k)
    xp, atyp = ora.get(D.XP).reshape(lp.n, k), ora.get(D.ATYP).reshape(lp.n, k)
    lo, up = ora.get(D.LO).reshape(lp.n, k), ora.get(D.UP).reshape(lp.n, k)
    mag_r = np.zeros(3 * k)
    mag_c = np.zeros(3 * k)
    mag_r[1::3] = (np.abs(yp) * (fb(rl) + fb(ru))).sum(0)
    mag_c[1::3] = (np.abs(c) * np.abs(xp)).sum(0)
    mag_c[2::3] = (np.abs(c - atyp) * (fb(lo) + fb(up))).sum(0)
    assert np.all(np.abs(r3g - r3o) <= 1e-12 * (np.abs(r3o) + mag_r) + 1e-300), (r3g, r3o)
    assert np.all(np.abs(c3g - c3o) <= 1e-12 * (np.abs(c3o) + mag_c) + 1e-300), (c3g, c3o)
    for a, b in ((D.X, D.XP), (D.Y, D.YP), (D.XP, D.X0), (D.YP, D.Y0)):
        np.testing.assert_allclose(be.diff2(a, b), ora.diff2(a, b), rtol=1e-12)
    # Halpern (rho = 1, then after a masked restart rho = 0.5: the restart
    # resets the masked instances' counters, which sets their lambda)
    for rho, mask in ((1.0, None), (0.5, (np.arange(k) % 3 == 0).astype(np.int32))):
        sync(be, ora)
        if mask is not None:
            be.restart(mask)
            ora.restart(mask)
            for v in (D.X, D.X0, D.Y, D.Y0, D.AX, D.AX0):
                assert np.array_equal(be.get(v), ora.get(v)), v  # copies are exact
        scale = {}
        for z, zp, z0 in ((D.X, D.XP, D.X0), (D.Y, D.YP, D.Y0), (D.AX, D.AXP, D.AX0)):
            scale[z] = 2 * np.abs(ora.get(zp)) + np.abs(ora.get(z)) + np.abs(ora.get(z0))
        be.halpern(rho)
        ora.halpern(rho)
        for z in (D.X, D.Y, D.AX):
            close(be.get(z), ora.get(z), scale[z], 8, eps)


@pytest.mark.parametrize("prec,k,tol", [(P.FP64, 1, 1e-6), (P.FP64, 3, 1e-6), (P.FP32, 2, 1e-4)])
def test_reference_host_loop_over_device_backend(gpu, prec, k, tol):
    """The reference's runPdhg loop driving the device backend op by op reaches
    the oracle's answer: FP64 within 4 checks of the oracle's iteration count
    and its objective within 1e-6 relative; FP32 iterates at 1e-4 within
    1e-3 of the closed-form optimum."""
    lp = P.gen_random_lp(12, 100, 200)
    rng = np.random.default_rng(4)
    costs = np.stack([lp.c] + [lp.c + rng.uniform(0, 0.05, lp.n) for _ in range(k - 1)])
    pr = O.oracle_prepare(lp)
    A, AT = D.scaled_csr_pair(lp, pr)
    be = P.Backend(A, AT, pr["beta"] / pr["R"], pr["gamma"] / pr["C"], k=k, precision=prec)
    res = D.run_pdhg(be, lp, pr, costs, tol=tol, max_iter=200000)
    ref, _, _, _ = O.oracle_solve(lp, costs, tol=tol, max_iter=200000)
    for b in range(k):
        assert res[b]["status"] == 1, res[b]["relative_kkt"]
        if prec == P.FP64:
            assert abs(res[b]["iterations"] - ref[b]["iterations"]) <= 4 * 64
            assert abs(res[b]["objective"] - ref[b]["objective"]) <= 1e-6 * (1 + abs(ref[b]["objective"]))
        else:
            assert abs(res[b]["objective"] - ref[b]["objective"]) <= 1e-3 * (1 + abs(ref[b]["objective"]))
    star = float(lp.c @ lp.xstar) + lp.offset
    assert abs(res[0]["objective"] - star) <= (1e-5 if prec == P.FP64 else 1e-3) * (1 + abs(star))
