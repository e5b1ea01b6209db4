"""CPU tests of the warm-start building blocks.

* The FP64 MLP oracle (oracle/mlp_oracle.py) is pinned by central finite
  differences of its own loss, so a dropped term or a transposed weight in its
  backward pass fails here before it is used to check the HIP kernels.
* The POD basis (host FP64 Jacobi eigensolver in the library) is checked
  against numpy's SVD of the centred snapshots: same energies, orthonormal
  modes spanning the same subspace, exact reconstruction at full rank.
"""
import os
import sys

import numpy as np

import pdhg_amd as P

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, hid, n_out, N = 4, 9, 5, 13
    npar = hid * n_in + hid + hid * hid + hid + n_out * hid + n_out
    p = rng.standard_normal(npar) * 0.5
    X = rng.uniform(-1, 1, (N, n_in))
    T = rng.standard_normal((N, n_out))
    _, g = M.loss_grad(p, X, T, n_in, hid, n_out)
    h = 1e-6
    for i in rng.choice(npar, 60, replace=False):
        e = np.zeros(npar)
        e[i] = h
        fd = (M.loss_grad(p + e, X, T, n_in, hid, n_out)[0] - M.loss_grad(p - e, X, T, n_in, hid, n_out)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * (1 + abs(fd)), (i, fd, g[i])


def test_gelu_is_exact_erf_form():
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    # GELU(z) = z Phi(z); Phi(0) = 1/2, Phi(-z) = 1 - Phi(z).
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert np.allclose(M.gelu(z) - M.gelu(-z), z)
    h = 1e-6
    assert np.allclose((M.gelu(z + h) - M.gelu(z - h)) / (2 * h), M.gelu_grad(z), atol=1e-8)


def test_pod_basis_matches_svd():
    rng = np.random.default_rng(3)
    k, d, r = 40, 300, 8
    # Snapshots near a 6-dimensional affine subspace plus noise.
    Z = rng.standard_normal(d) + rng.standard_normal((k, 6)) @ rng.standard_normal((6, d)) \
        + 1e-3 * rng.standard_normal((k, d))
    mean, U, A, energy, used = P.pod_basis(Z, r)
    assert used == r
    Zc = Z - Z.mean(0)
    s = np.linalg.svd(Zc, compute_uv=False)
    assert np.allclose(mean, Z.mean(0))
    assert np.allclose(energy, (s[:r] ** 2) / np.sum(s ** 2), rtol=1e-8, atol=1e-14)
    # Gram-matrix (snapshot) POD: mode accuracy ~ eps / energy, 1e-8 for the noise modes here.
    assert np.allclose(U.T @ U, np.eye(r), atol=1e-6)
    assert np.allclose(A, Zc @ U, atol=1e-6 * np.abs(Zc).max())
    # The leading modes span the principal subspace of numpy's SVD.
    Vt = np.linalg.svd(Zc, full_matrices=False)[2][:6]
    assert np.allclose(np.abs(np.linalg.svd(U[:, :6].T @ Vt.T, compute_uv=False)), 1.0, atol=1e-8)


def test_pod_full_rank_reconstructs_snapshots():
    rng = np.random.default_rng(4)
    k, d = 12, 50
    Z = rng.standard_normal((k, d))
    mean, U, A, energy, used = P.pod_basis(Z, k)
    assert used == k - 1  # centring removes one dimension
    assert np.allclose(mean + A @ U.T, Z, atol=1e-10)
