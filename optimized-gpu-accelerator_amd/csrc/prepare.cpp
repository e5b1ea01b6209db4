#include "prepare.h"

#include <algorithm>
#include <cmath>

namespace pdhgamd {

namespace {

// Nearest power of two, exponent clamped to +-60; non-positive or non-finite
// values map to 1 (src/lp/scaling.cpp:9-14).  Power-of-two factors make the
// scaling of A, bounds and costs exact.
double nearest_pow2(double v) {
  if (!(v > 0) || !std::isfinite(v)) return 1.0;
  int e = static_cast<int>(std::lround(std::log2(v)));
  return std::ldexp(1.0, std::clamp(e, -60, 60));
}

// |a_ij| R_i C_j of the current scaled matrix.
struct ScaledView {
  const Csc& a;
  const std::vector<double>& R;
  const std::vector<double>& C;
  double at(int j, int e) const { return std::fabs(a.value[e]) * C[j] * R[a.index[e]]; }
};

// Ruiz infinity-norm equilibration followed by one Pock-Chambolle (alpha = 1)
// step (src/lp/scaling.cpp:97-139).  Row and column factors of a Ruiz pass
// both come from the same sweep over A.
void equilibrate(const Csc& a, int iters, bool pc, std::vector<double>& R, std::vector<double>& C) {
  R.assign(a.rows, 1.0);
  C.assign(a.cols, 1.0);
  ScaledView s{a, R, C};
  std::vector<double> rowMax(a.rows), colMax(a.cols);
  for (int pass = 0; pass < iters; ++pass) {
    std::fill(rowMax.begin(), rowMax.end(), 0.0);
    for (int j = 0; j < a.cols; ++j) {
      double cm = 0;
      for (int e = a.start[j]; e < a.start[j + 1]; ++e) {
        const double v = s.at(j, e);
        cm = std::max(cm, v);
        double& rm = rowMax[a.index[e]];
        rm = std::max(rm, v);
      }
      colMax[j] = cm;
    }
    for (int i = 0; i < a.rows; ++i)
      if (rowMax[i] > 0) R[i] /= std::sqrt(rowMax[i]);
    for (int j = 0; j < a.cols; ++j)
      if (colMax[j] > 0) C[j] /= std::sqrt(colMax[j]);
  }
  if (pc) {
    std::vector<double> rowSum(a.rows, 0.0), colSum(a.cols, 0.0);
    for (int j = 0; j < a.cols; ++j)
      for (int e = a.start[j]; e < a.start[j + 1]; ++e) {
        const double v = s.at(j, e);
        colSum[j] += v;
        rowSum[a.index[e]] += v;
      }
    for (int i = 0; i < a.rows; ++i)
      if (rowSum[i] > 0) R[i] /= std::sqrt(rowSum[i]);
    for (int j = 0; j < a.cols; ++j)
      if (colSum[j] > 0) C[j] /= std::sqrt(colSum[j]);
  }
  std::transform(R.begin(), R.end(), R.begin(), nearest_pow2);
  std::transform(C.begin(), C.end(), C.begin(), nearest_pow2);
}

// sigma_max(A) from 40 power steps on A^T A starting at 1/sqrt(n)
// (src/pdhg/pdhg.cpp:404-422).  A x is a column scatter that skips zero x_j
// and A^T y a column gather, as SparseMatrix::multiply / multiplyTranspose
// (src/core/sparse.cpp:77-92).
double spectral_norm(const Csc& a, int steps) {
  const int n = a.cols, m = a.rows;
  if (a.nnz() == 0) return 1.0;
  std::vector<double> x(n, 1.0 / std::sqrt(std::max(1, n))), ax(m), atax(n);
  double est = 1;
  for (int it = 0; it < steps; ++it) {
    std::fill(ax.begin(), ax.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      double xj = x[j];
      if (xj == 0.0) continue;
      xj *= 1.0;
      for (int e = a.start[j]; e < a.start[j + 1]; ++e) ax[a.index[e]] += a.value[e] * xj;
    }
    std::fill(atax.begin(), atax.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      double g = 0;
      for (int e = a.start[j]; e < a.start[j + 1]; ++e) g += a.value[e] * ax[a.index[e]];
      atax[j] += 1.0 * g;
    }
    double nrm = 0;
    for (double v : atax) nrm += v * v;
    nrm = std::sqrt(nrm);
    if (nrm == 0) return 1.0;
    est = std::sqrt(nrm);
    for (int j = 0; j < n; ++j) x[j] = atax[j] / nrm;
  }
  return est;
}

}  // namespace

Prepared prepare(const Model& lp, int ruizIterations, bool pockChambolle, bool rescale) {
  Prepared P;
  P.m = lp.m;
  P.n = lp.n;
  P.offset = lp.offset;
  Csc a = lp.A;
  a.rows = lp.m;
  if (ruizIterations > 0 && a.nnz() > 0) {
    equilibrate(a, ruizIterations, pockChambolle, P.R, P.C);
  } else {
    P.R.assign(P.m, 1.0);
    P.C.assign(P.n, 1.0);
  }
  for (int j = 0; j < a.cols; ++j)
    for (int e = a.start[j]; e < a.start[j + 1]; ++e) a.value[e] *= P.R[a.index[e]] * P.C[j];
  if (rescale) {
    double cn = 0, bn = 0;
    for (int j = 0; j < P.n; ++j) cn += sq(lp.c[j] * P.C[j]);
    for (int i = 0; i < P.m; ++i) {
      if (finite_bound(lp.L[i])) bn += sq(lp.L[i] * P.R[i]);
      if (finite_bound(lp.U[i]) && lp.U[i] != lp.L[i]) bn += sq(lp.U[i] * P.R[i]);
    }
    P.gamma = 1.0 + std::sqrt(cn);
    P.beta = 1.0 + std::sqrt(bn);
  }
  P.A = csr_of(a);
  P.AT = csr_of_transpose(a);
  P.rfac.resize(P.m);
  P.cfac.resize(P.n);
  for (int i = 0; i < P.m; ++i) P.rfac[i] = P.beta / P.R[i];
  for (int j = 0; j < P.n; ++j) P.cfac[j] = P.gamma / P.C[j];
  P.normA = spectral_norm(a, 40) * 1.01;
  return P;
}

}  // namespace pdhgamd
