// Proper-orthogonal-decomposition (POD) basis of PDHG solution snapshots for
// the learned warm start: centred snapshots Zc (k x d), their Gram matrix
// G = Zc Zc^T diagonalised by the cyclic Jacobi method (FP64, host), and the
// leading `rank` modes u_i = Zc^T v_i / sqrt(lambda_i), orthonormal in R^d.
// The warm start has no counterpart in the reference (x0, y0 are hard-wired
// there, src/pdhg/pdhg.cpp:509,519).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>
#include <vector>

#include "core.h"

namespace pdhgamd {

namespace {

// Cyclic Jacobi eigenvalue iteration for a symmetric k x k matrix (row-major,
// overwritten).  Returns eigenvalues; V holds eigenvectors as columns.
std::vector<double> jacobi_eigen(std::vector<double>& G, int k, std::vector<double>& V) {
  V.assign(static_cast<size_t>(k) * k, 0.0);
  for (int i = 0; i < k; ++i) V[static_cast<size_t>(i) * k + i] = 1.0;
  auto g = [&](int r, int c) -> double& { return G[static_cast<size_t>(r) * k + c]; };
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0, diag = 0;
    for (int p = 0; p < k; ++p) {
      diag += g(p, p) * g(p, p);
      for (int q = p + 1; q < k; ++q) off += g(p, q) * g(p, q);
    }
    if (off <= 1e-30 * std::max(diag, 1e-300)) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = g(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        // Rotation angle zeroing G(p, q) (Golub & Van Loan, Algorithm 8.4.1).
        const double tau = (g(q, q) - g(p, p)) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int r = 0; r < k; ++r) {  // G <- G J
          const double grp = g(r, p), grq = g(r, q);
          g(r, p) = c * grp - s * grq;
          g(r, q) = s * grp + c * grq;
        }
        for (int r = 0; r < k; ++r) {  // G <- J^T G
          const double gpr = g(p, r), gqr = g(q, r);
          g(p, r) = c * gpr - s * gqr;
          g(q, r) = s * gpr + c * gqr;
        }
        for (int r = 0; r < k; ++r) {
          double& vp = V[static_cast<size_t>(r) * k + p];
          double& vq = V[static_cast<size_t>(r) * k + q];
          const double a = vp, b = vq;
          vp = c * a - s * b;
          vq = s * a + c * b;
        }
      }
  }
  std::vector<double> lam(k);
  for (int i = 0; i < k; ++i) lam[i] = g(i, i);
  return lam;
}

}  // namespace

// Z: k x d snapshots (row-major).  Outputs mean[d], basis[d x rank] (row-major,
// basis[j * rank + i] = u_i[j]), coeffs[k x rank] and energy[rank] (fraction of
// the snapshot variance each mode carries).  Returns the rank actually used
// (modes with negligible variance are left zero).
int pod_basis(const double* Z, int k, int d, int rank, double* mean, double* basis, double* coeffs, double* energy) {
  PDHG_REQUIRE(k >= 1 && d >= 1 && rank >= 1, "pod: bad sizes");
  std::fill(mean, mean + d, 0.0);
  for (int s = 0; s < k; ++s)
    for (int j = 0; j < d; ++j) mean[j] += Z[static_cast<size_t>(s) * d + j];
  for (int j = 0; j < d; ++j) mean[j] /= k;
  // Gram matrix of the centred snapshots, rows split over threads.
  std::vector<double> G(static_cast<size_t>(k) * k, 0.0);
  const int nt = std::max(1, std::min(16, static_cast<int>(std::thread::hardware_concurrency())));
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
      pool.emplace_back([&, t] {
        for (int a = t; a < k; a += nt)
          for (int b = a; b < k; ++b) {
            double acc = 0;
            const double *za = Z + static_cast<size_t>(a) * d, *zb = Z + static_cast<size_t>(b) * d;
            for (int j = 0; j < d; ++j) acc += (za[j] - mean[j]) * (zb[j] - mean[j]);
            G[static_cast<size_t>(a) * k + b] = G[static_cast<size_t>(b) * k + a] = acc;
          }
      });
    for (auto& th : pool) th.join();
  }
  std::vector<double> V;
  std::vector<double> lam = jacobi_eigen(G, k, V);
  std::vector<int> order(k);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return lam[a] > lam[b]; });
  double total = 0;
  for (double l : lam) total += std::max(l, 0.0);
  std::fill(basis, basis + static_cast<size_t>(d) * rank, 0.0);
  std::fill(coeffs, coeffs + static_cast<size_t>(k) * rank, 0.0);
  std::fill(energy, energy + rank, 0.0);
  int used = 0;
  for (int i = 0; i < rank && i < k; ++i) {
    const double l = lam[order[i]];
    if (!(l > 1e-14 * std::max(total, 1e-300))) break;
    const double sl = std::sqrt(l);
    for (int j = 0; j < d; ++j) {
      double acc = 0;
      for (int s = 0; s < k; ++s)
        acc += (Z[static_cast<size_t>(s) * d + j] - mean[j]) * V[static_cast<size_t>(s) * k + order[i]];
      basis[static_cast<size_t>(j) * rank + i] = acc / sl;
    }
    for (int s = 0; s < k; ++s) coeffs[static_cast<size_t>(s) * rank + i] = sl * V[static_cast<size_t>(s) * k + order[i]];
    energy[i] = l / total;
    ++used;
  }
  return used;
}

}  // namespace pdhgamd
