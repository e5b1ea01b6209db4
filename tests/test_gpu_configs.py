"""The benchmark configurations at their stated sizes on the GPU.

C3: the 64-case price family of the full refinery LP (seed 3; theta seed 4,
range 0.3; 555,852 nonzeros, rows and columns scaled by 10^U(-3,3)) in one
batched run of width 64, autotuned row mapping, hipGraphs of 64 iterations,
FP64 iterates, rel-KKT 1e-6: from the cold start, and from the learned warm
start with its per-case cold-start fallback.  The reference analogue is the
K=64 family block of compareFamilyStrategies
(src/parametric/parametric.cpp:483-530).  These cases run 10^5 - 10^6
iterations, beyond what the FP64 oracle finishes in a test, so every answer
is checked by size-independent properties: an FP64 host re-evaluation of the
KKT residuals and the reference's certifier acceptance (implied bounds, safe
dual bound, |obj - LB| <= 1e-4 (1 + |obj|)).  Six of the cases are also
checked against the reference's own CPU PDHG (its sources, FP64), solved
ahead of time into tests/golden/c3_reference_cases.json by
tests/golden/make_c3_reference.py: iterations within 4 checks, objective and
certified dual bound within 1e-6 relative.

C4's uneven batch: 256 random LPs of the config-0 generator (seed 7, nnz
log-uniform in [2e3, 2e5], a non-binding box x <= 2) in one packed run:
each converged answer against its generator's closed-form optimum (relative
to the size of the objective's terms), the host KKT re-evaluation and the
certifier."""
import json
import os

import numpy as np
import pytest

import pdhg_amd as P
from test_gpu_parity import host_kkt

pytestmark = pytest.mark.gpu

REFERENCE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_reference_cases.json")


def check_against_reference(res, bounds, label):
    """Iterations within 4 checks of the reference's (its test_solvers.cpp:128
    bound), objective and certified dual bound within 1e-6 relative."""
    with open(REFERENCE) as f:
        cases = json.load(f)["cases"]
    worst = [0.0, 0.0, 0]
    for q, rec in cases.items():
        r = res[int(q)]
        assert r.status == P.OPTIMAL and rec["status"] == P.OPTIMAL, (label, q)
        dobj = abs(r.objective - rec["objective"]) / (1 + abs(rec["objective"]))
        dlb = abs(bounds[int(q)] - rec["safe_bound"]) / (1 + abs(rec["safe_bound"]))
        dit = abs(r.iterations - rec["iterations"])
        assert dit <= 4 * 64, (label, q, r.iterations, rec["iterations"])
        assert dobj <= 1e-6 and dlb <= 1e-6, (label, q, dobj, dlb)
        worst = [max(worst[0], dobj), max(worst[1], dlb), max(worst[2], dit)]
    print(f"{label} against the reference's CPU PDHG on {len(cases)} cases: objective {worst[0]:.1e}, "
          f"bound {worst[1]:.1e}, iterations {worst[2]} apart at most")


@pytest.fixture(scope="module")
def family():
    lp = P.gen_refinery(3)
    thetas, costs = P.gen_family(lp, 4, 64, 0.3)
    return lp, thetas, costs


def check_family(lp, costs, res, label):
    xs = np.stack([r.x for r in res])
    ys = np.stack([r.y for r in res])
    obj = np.array([r.objective for r in res])
    cert = P.certify_batch(lp, costs, xs, ys, obj)
    conv = [q for q, r in enumerate(res) if r.status == P.OPTIMAL]
    for q in conv:
        assert res[q].relative_kkt <= 1e-6
        relP, relD, gap, pobj = host_kkt(lp, res[q].x, res[q].y, costs[q])
        assert max(relP, relD, gap) <= 2e-6, (label, q, relP, relD, gap)
        assert abs(pobj - res[q].objective) <= 1e-9 * (1 + abs(pobj))
        assert cert[q, 3] == 1.0, (label, q, cert[q])
    for q, r in enumerate(res):
        if r.status != P.OPTIMAL:  # the reference's cap: stopped exactly at its limit
            assert r.status == P.ITERATION_LIMIT, (label, q, r.status)
    return conv, cert


def test_c3_family_cold_batch(gpu, family):
    lp, _, costs = family
    res, st, _ = P.solve_batch(lp, costs, P.default_options(tolerance=1e-6, precision=P.FP64))
    conv, cert = check_family(lp, costs, res, "cold")
    its = np.array([r.iterations for r in res])
    print(f"C3 cold: {len(conv)}/64 converged, {int(cert[:, 3].sum())} certified, run {st['iterations']} "
          f"iterations, median case {np.median(its):.0f}, {st['solve_seconds']:.1f} s, tuned col_row_map "
          f"{st['col_row_map']} row_row_map {st['row_row_map']}")
    assert st["batch"] == 64 and st["autotuned"] == 1 and st["graph_launches"] > 0
    check_against_reference(res, cert[:, 0], "C3 cold")
    # The hardest cases of this ill-conditioned family need about the
    # reference's 1M-iteration cap: at most two stop there.
    assert len(conv) >= 62
    for q, r in enumerate(res):
        if r.status != P.OPTIMAL:
            assert r.iterations == 1_000_000 and r.relative_kkt <= 1e-4, (q, r.relative_kkt)


def test_c3_family_warm_start_with_cold_fallback(gpu, family):
    """The learned warm start (POD basis + MLP trained on the GPU on 64 other
    parameters of the family) with the per-case cold-start fallback: a warm
    case not converged after 256,000 iterations restarts in its lane from the
    reference's cold start.  Every warm answer meets the same tolerance and
    certifier as the cold ones, and the fallback recovers the cases on which
    the warm start stalls."""
    import warmstart as W

    lp, thetas, costs = family
    o = P.default_options(tolerance=1e-6, precision=P.FP64)
    tht, ct = P.gen_family(lp, 5, 64, 0.3)
    tr, _, _ = P.solve_batch(lp, ct, o)
    ok = [q for q in range(64) if tr[q].status == P.OPTIMAL]
    ws = W.WarmStart(lp, rank=32, hidden=128).fit(tht[ok], np.stack([tr[q].x for q in ok]),
                                                 np.stack([tr[q].y for q in ok]), steps=3000, lr=1e-3,
                                                 omegas=np.array([tr[q].primal_weight for q in ok]))
    x0, y0 = ws.predict(thetas)
    w0 = ws.primal_weights(thetas)
    o.warm_fallback = 256_000  # a multiple of the check interval (64)
    res, st, _ = P.solve_batch(lp, costs, o, x0=x0, y0=y0, primal_weight=w0)
    conv, cert = check_family(lp, costs, res, "warm")
    its = np.array([r.iterations for r in res])
    print(f"C3 warm: {len(conv)}/64 converged, {int(cert[:, 3].sum())} certified, {st['warm_fallbacks']} cold "
          f"fallbacks, run {st['iterations']} iterations, median case {np.median(its):.0f}, "
          f"{st['solve_seconds']:.1f} s")
    assert len(conv) >= 62
    # Cases that converged warm did so within the budget; the others ran it
    # out and continued from the cold start with a cap of their own.
    for q, r in enumerate(res):
        if r.status != P.OPTIMAL:  # the warm budget plus a full cold cap of its own
            assert r.iterations == 256_000 + 1_000_000, (q, r.iterations)
    assert st["warm_fallbacks"] == sum(1 for r in res if r.iterations > 256_000)
    assert sum(1 for r in res if r.iterations < 256_000) >= 16  # the warm start solves these early


def test_c4_uneven_batch_full_size(gpu):
    lps = P.gen_uneven_batch(7, 256)
    nnz = np.array([q.nnz for q in lps])
    assert nnz.min() >= 1500 and nnz.max() <= 2.1e5 and len(lps) == 256
    # A cap of 400k iterations keeps the test short: the slowest of these
    # boxed LPs need up to and beyond the reference's 1M (16 of 256 stop
    # there), and a capped LP stops at exactly its cap.
    cap = 400_000
    res, st = P.solve_packed(lps, P.default_options(tolerance=1e-6, precision=P.FP64, max_iterations=cap))
    conv = 0
    certified = 0
    for lp, r in zip(lps, res):
        if r.status != P.OPTIMAL:
            assert r.status == P.ITERATION_LIMIT and r.iterations == cap
            continue
        conv += 1
        obj_star = float(lp.c @ lp.xstar) + lp.offset
        # These objectives are sums of terms of both signs that nearly cancel
        # (|c'x*| ~ 0.3 against sum |c_j x*_j| ~ 10^2): the error is measured
        # against the size of the terms.
        scale = 1 + float(np.abs(lp.c * lp.xstar).sum())
        assert abs(r.objective - obj_star) <= 1e-6 * scale, (lp.m, lp.n, r.objective, obj_star)
        relP, relD, gap, _ = host_kkt(lp, r.x, r.y)
        assert max(relP, relD, gap) <= 1.5e-6, (lp.m, lp.n, relP, relD, gap)
        c = P.certify(lp, r.x, r.y, r.objective)
        certified += int(c["accepted"])
    print(f"C4 packed: {conv}/256 converged, {certified} certified, {st['iterations']} iterations, "
          f"{st['solve_seconds']:.1f} s, {nnz.sum()} nonzeros")
    assert conv >= 200
    assert certified == conv
