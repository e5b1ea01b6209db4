"""GPU tests of the learned warm start: the HIP MLP kernels against the FP64
oracle, the BF16-MFMA basis decode, and the gain of warm-started PDHG over
cold-started PDHG on a held-out parametric family (same tolerance, same
certifier)."""
import os
import sys

import numpy as np
import pytest

import pdhg_amd as P
import warmstart as W

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def test_mlp_forward_and_gradient_match_oracle(gpu):
    net = P.Mlp(4, 128, 32, seed=1)
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (96, 4))
    T = rng.standard_normal((96, 32))
    p = net.params
    Y = net.forward(X)
    Yo, _ = M.forward(p, X, 4, 128, 32)
    # BF16 operands (8-bit mantissa) with FP32 accumulation.
    assert np.linalg.norm(Y - Yo) <= 2e-2 * np.linalg.norm(Yo)
    loss, g = net.loss_grad(X, T)
    lo, go = M.loss_grad(p, X, T, 4, 128, 32)
    assert abs(loss - lo) <= 2e-2 * lo
    assert np.linalg.norm(g - go) <= 3e-2 * np.linalg.norm(go)
    assert g @ go / (np.linalg.norm(g) * np.linalg.norm(go)) > 0.999
    # Per block (W1, b1, W2, b2, W3, b3): each gradient piece agrees.
    sizes = [128 * 4, 128, 128 * 128, 128, 32 * 128, 32]
    o = 0
    for s in sizes:
        assert np.linalg.norm(g[o:o + s] - go[o:o + s]) <= 5e-2 * np.linalg.norm(go[o:o + s]) + 1e-9
        o += s


def test_mlp_training_reduces_loss(gpu):
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (256, 4))
    T = np.stack([np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2], np.cos(X[:, 3]) - X[:, 0] ** 2], 1)
    net = P.Mlp(4, 64, 2, seed=2)
    losses = net.train(X, T, steps=1500, lr=3e-3)
    assert losses[-1] < 0.02 * losses[0]
    lo, _ = M.loss_grad(net.params, X, T, 4, 64, 2)
    assert abs(lo - losses[-1]) <= 0.2 * lo + 1e-4  # the oracle agrees with the trained fit


def test_decode_matches_host_product(gpu):
    rng = np.random.default_rng(2)
    d, r, N = 1000, 32, 20
    U = rng.standard_normal((d, r))
    mean = rng.standard_normal(d)
    A = rng.standard_normal((N, r))
    out = P.decode_bf16(U, mean, A)
    ref = mean + A @ U.T
    assert np.linalg.norm(out - ref) <= 1e-2 * np.linalg.norm(ref)


def small_family():
    lp = P.gen_refinery(3, periods=2, streams=12, crudes=4, products=4, qualities=5, decades=3.0)
    return lp


def test_warm_start_beats_cold_start_on_held_out_family(gpu):
    lp = small_family()
    th_train, c_train = P.gen_family(lp, 5, 64, 0.3)
    th_eval, c_eval = P.gen_family(lp, 6, 16, 0.3)
    opts = P.default_options(tolerance=1e-6, precision=P.FP64, max_iterations=400_000)
    train, _, _ = P.solve_batch(lp, c_train, opts)
    assert all(r.status == P.OPTIMAL for r in train)
    ws = W.WarmStart(lp, rank=32, hidden=128).fit(th_train, np.stack([r.x for r in train]),
                                                 np.stack([r.y for r in train]), steps=2000, lr=1e-3,
                                                 omegas=np.array([r.primal_weight for r in train]))
    assert ws.losses[-1] < 0.2 * ws.losses[0]
    assert ws.weight_losses[-1] < ws.weight_losses[0]
    cold, _, _ = P.solve_batch(lp, c_eval, opts)
    x0, y0 = ws.predict(th_eval)
    warm, _, _ = P.solve_batch(lp, c_eval, opts, x0=x0, y0=y0)
    xm, ym = ws.predict(th_eval, coeffs=np.zeros((16, ws.rank)))
    mean_only, _, _ = P.solve_batch(lp, c_eval, opts, x0=xm, y0=ym)
    w0 = ws.primal_weights(th_eval)
    assert w0.shape == (16,) and np.all(np.isfinite(w0)) and np.all(w0 > 0)
    warm_w, _, _ = P.solve_batch(lp, c_eval, opts, x0=x0, y0=y0, primal_weight=w0)
    ic = np.array([r.iterations for r in cold], float)
    iw = np.array([r.iterations for r in warm], float)
    im = np.array([r.iterations for r in mean_only], float)
    iww = np.array([r.iterations for r in warm_w], float)
    print(f"iterations cold {ic.mean():.0f}  mean-of-snapshots start {im.mean():.0f}  learned start {iw.mean():.0f}"
          f"  learned start + learned primal weight {iww.mean():.0f}")
    for group in (cold, warm, warm_w):
        assert all(r.status == P.OPTIMAL and r.relative_kkt <= 1e-6 for r in group)
        cert = P.certify_batch(lp, c_eval, np.stack([r.x for r in group]), np.stack([r.y for r in group]),
                               np.array([r.objective for r in group]))
        assert np.all(cert[:, 3] == 1.0)
    # Two answers at rel-KKT <= 1e-6 of an LP with rows scaled by 10^U(-3,3):
    # the objective tolerance is the certifier's acceptance level.
    rel = [abs(a.objective - b.objective) / (1 + abs(a.objective)) for a, b in zip(cold, warm)]
    print(f"max relative objective difference warm vs cold {max(rel):.2e}")
    assert max(rel) <= 1e-4
    assert iw.mean() < ic.mean()
