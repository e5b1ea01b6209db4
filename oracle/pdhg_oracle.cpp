// ORACLE - test infrastructure only.
//
// A serial FP64 CPU restatement of the reference's restarted reflected-Halpern
// PDHG and of its Neumaier-Shcherbina safe dual bound.  Only tests/, the smoke
// check and bench.py's CPU-baseline leg may load this library, and only as the
// checker.  The product (the HIP engine) never links or calls it.
//
// Every function cites the reference file:line it restates.  Floating-point
// operations are issued in the same order as the reference so that, compiled
// with -ffp-contract=off like the reference, the two produce identical bits.
// The reference CPU backend splits SpMV rows and elementwise loops over threads
// but every output element is computed by one serial expression, and its KKT /
// diff2 reductions are serial (src/pdhg/pdhg.cpp:93-205), so a serial
// restatement is bit-compatible with any thread count.
//
// Pinning (tests/test_oracle.py): bitwise equality with the compiled reference
// (oracle/_ref) on generated LPs and the afiro / sc50a fixtures; the closed-form
// optimum of the generated LPs; the 2-variable safe-bound known answer 2.8 of
// tests/cpp/test_core.cpp:175-193 of the reference.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kInfBound = 1e20;  // src/util/common.h:15
inline bool fin(double v) { return std::fabs(v) < kInfBound; }
inline double sq(double v) { return v * v; }

thread_local std::string gErr;

struct Csc {  // column-major matrix, src/core/sparse.h:10-16
  int rows = 0, cols = 0;
  std::vector<int> start, index;
  std::vector<double> value;
};

struct Csr {
  int rows = 0, cols = 0;
  std::vector<int> ptr, idx;
  std::vector<double> val;
};

// CSC of the transpose == CSR of the matrix (src/core/sparse.cpp:59-75).
Csr csrOf(const Csc& a) {
  Csr t;
  t.rows = a.rows;
  t.cols = a.cols;
  const int nnz = a.start[a.cols];
  std::vector<int> cnt(a.rows + 1, 0);
  for (int e = 0; e < nnz; ++e) cnt[a.index[e] + 1]++;
  for (int i = 0; i < a.rows; ++i) cnt[i + 1] += cnt[i];
  t.ptr = cnt;
  t.idx.resize(nnz);
  t.val.resize(nnz);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (int j = 0; j < a.cols; ++j)
    for (int e = a.start[j]; e < a.start[j + 1]; ++e) {
      int p = pos[a.index[e]]++;
      t.idx[p] = j;
      t.val[p] = a.value[e];
    }
  return t;
}

// CSR of A^T is the CSC of A itself (src/pdhg/pdhg.cpp:35-43).
Csr csrOfTranspose(const Csc& a) {
  Csr t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.ptr = a.start;
  t.idx = a.index;
  t.val = a.value;
  return t;
}

// src/lp/scaling.cpp:9-14
double pow2Round(double v) {
  if (!(v > 0) || !std::isfinite(v)) return 1.0;
  int e = static_cast<int>(std::lround(std::log2(v)));
  e = std::max(-60, std::min(60, e));
  return std::ldexp(1.0, e);
}

// Ruiz (infinity norm) + Pock-Chambolle alpha=1, factors rounded to powers of
// two (src/lp/scaling.cpp:97-139).
void ruiz(const Csc& a, int iters, bool pc, std::vector<double>& R, std::vector<double>& C) {
  const int m = a.rows, n = a.cols;
  R.assign(m, 1.0);
  C.assign(n, 1.0);
  std::vector<double> rmax(m), cmax(n);
  for (int it = 0; it < iters; ++it) {
    std::fill(rmax.begin(), rmax.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      double c = 0;
      for (int e = a.start[j]; e < a.start[j + 1]; ++e) {
        double v = std::fabs(a.value[e]) * C[j] * R[a.index[e]];
        c = std::max(c, v);
        rmax[a.index[e]] = std::max(rmax[a.index[e]], v);
      }
      cmax[j] = c;
    }
    for (int i = 0; i < m; ++i)
      if (rmax[i] > 0) R[i] /= std::sqrt(rmax[i]);
    for (int j = 0; j < n; ++j)
      if (cmax[j] > 0) C[j] /= std::sqrt(cmax[j]);
  }
  if (pc) {
    std::vector<double> rs(m, 0.0), cs(n, 0.0);
    for (int j = 0; j < n; ++j)
      for (int e = a.start[j]; e < a.start[j + 1]; ++e) {
        double v = std::fabs(a.value[e]) * C[j] * R[a.index[e]];
        cs[j] += v;
        rs[a.index[e]] += v;
      }
    for (int i = 0; i < m; ++i)
      if (rs[i] > 0) R[i] /= std::sqrt(rs[i]);
    for (int j = 0; j < n; ++j)
      if (cs[j] > 0) C[j] /= std::sqrt(cs[j]);
  }
  for (auto& v : R) v = pow2Round(v);
  for (auto& v : C) v = pow2Round(v);
}

// ||A||_2 estimate: 40 power steps on A^T A, column-major products as in
// SparseMatrix::multiply / multiplyTranspose (src/pdhg/pdhg.cpp:404-422,
// src/core/sparse.cpp:77-92).
double powerIteration(const Csc& a, int iters) {
  const int n = a.cols, m = a.rows;
  if (a.start[n] == 0) return 1.0;
  std::vector<double> x(n, 1.0 / std::sqrt(std::max(1, n))), y(m), z(n);
  double s = 1;
  for (int it = 0; it < iters; ++it) {
    std::fill(y.begin(), y.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      double xj = x[j];
      if (xj == 0.0) continue;
      xj *= 1.0;
      for (int e = a.start[j]; e < a.start[j + 1]; ++e) y[a.index[e]] += a.value[e] * xj;
    }
    std::fill(z.begin(), z.end(), 0.0);
    for (int j = 0; j < n; ++j) {
      double acc = 0;
      for (int e = a.start[j]; e < a.start[j + 1]; ++e) acc += a.value[e] * y[a.index[e]];
      z[j] += 1.0 * acc;
    }
    double nz = 0;
    for (double v : z) nz += v * v;
    nz = std::sqrt(nz);
    if (nz == 0) return 1.0;
    s = std::sqrt(nz);
    for (int j = 0; j < n; ++j) x[j] = z[j] / nz;
  }
  return s;
}

struct Prepared {  // src/pdhg/pdhg.cpp:394-402
  int n = 0, m = 0;
  Csr A, AT;
  std::vector<double> R, C, rfac, cfac;
  double beta = 1, gamma = 1, normA = 1, offset = 0;
};

struct Lp {
  int m, n;
  Csc A;
  const double *c, *l, *u, *L, *U;
  double offset;
};

// src/pdhg/pdhg.cpp:424-459
Prepared prepare(const Lp& lp, int ruizIters, bool pc, bool rescale) {
  Prepared P;
  P.n = lp.n;
  P.m = lp.m;
  P.offset = lp.offset;
  Csc A = lp.A;
  if (ruizIters > 0 && A.start[A.cols] > 0) {
    ruiz(A, ruizIters, pc, P.R, P.C);
  } else {
    P.R.assign(P.m, 1.0);
    P.C.assign(P.n, 1.0);
  }
  for (int j = 0; j < A.cols; ++j)  // src/lp/scaling.cpp:28-32
    for (int e = A.start[j]; e < A.start[j + 1]; ++e) A.value[e] *= P.R[A.index[e]] * P.C[j];
  if (rescale) {
    double cn = 0, bn = 0;
    for (int j = 0; j < P.n; ++j) cn += sq(lp.c[j] * P.C[j]);
    for (int i = 0; i < P.m; ++i) {
      if (fin(lp.L[i])) bn += sq(lp.L[i] * P.R[i]);
      if (fin(lp.U[i]) && lp.U[i] != lp.L[i]) bn += sq(lp.U[i] * P.R[i]);
    }
    P.gamma = 1.0 + std::sqrt(cn);
    P.beta = 1.0 + std::sqrt(bn);
  }
  P.A = csrOf(A);
  P.AT = csrOfTranspose(A);
  P.rfac.resize(P.m);
  P.cfac.resize(P.n);
  for (int i = 0; i < P.m; ++i) P.rfac[i] = P.beta / P.R[i];
  for (int j = 0; j < P.n; ++j) P.cfac[j] = P.gamma / P.C[j];
  P.normA = powerIteration(A, 40) * 1.01;
  return P;
}

enum V { X, XP, X0, ATY, ATYP, C, LO, UP, Y, YP, Y0, AX, AXP, AX0, RL, RU, NV };

// Serial restatement of CpuBackend (src/pdhg/pdhg.cpp:75-214).
struct Cpu {
  const Csr *A = nullptr, *AT = nullptr;
  std::vector<double> rfac, cfac;
  int n = 0, m = 0, k = 1;
  std::vector<double> v[NV];
  std::vector<int> iter;
  std::vector<double> tau, sigma;

  void init(const Prepared& P, int kk) {
    A = &P.A;
    AT = &P.AT;
    rfac = P.rfac;
    cfac = P.cfac;
    n = P.n;
    m = P.m;
    k = kk;
    for (int q = 0; q < NV; ++q) v[q].assign(static_cast<size_t>(q >= Y ? m : n) * k, 0.0);
    iter.assign(k, 0);
    tau.assign(k, 0);
    sigma.assign(k, 0);
  }
  void mult(const Csr& M, const std::vector<double>& in, std::vector<double>& out) {  // :93-103
    for (int i = 0; i < M.rows; ++i)
      for (int q = 0; q < k; ++q) {
        double s = 0;
        for (int p = M.ptr[i]; p < M.ptr[i + 1]; ++p) s += M.val[p] * in[static_cast<size_t>(M.idx[p]) * k + q];
        out[static_cast<size_t>(i) * k + q] = s;
      }
  }
  void spmv(int in, int out) { mult(*A, v[in], v[out]); }
  void spmvT(int in, int out) { mult(*AT, v[in], v[out]); }
  void primal() {  // :116-125
    auto &x = v[X], &aty = v[ATY], &c = v[C], &lo = v[LO], &up = v[UP], &xp = v[XP];
    for (size_t t = 0; t < x.size(); ++t) {
      double r = x[t] - tau[t % k] * (c[t] - aty[t]);
      r = std::max(r, lo[t]);
      r = std::min(r, up[t]);
      xp[t] = r;
    }
  }
  void dual() {  // :126-137
    auto &y = v[Y], &axp = v[AXP], &ax = v[AX], &rl = v[RL], &ru = v[RU], &yp = v[YP];
    for (size_t t = 0; t < y.size(); ++t) {
      double s = sigma[t % k];
      double w0 = y[t] - s * (2.0 * axp[t] - ax[t]);
      double w = -w0 / s;
      w = std::max(w, rl[t]);
      w = std::min(w, ru[t]);
      yp[t] = w0 + s * w;
    }
  }
  void halpernOne(int z, int zp, int z0, double rho) {  // :138-146
    auto &a = v[z], &p = v[zp], &o = v[z0];
    for (size_t t = 0; t < a.size(); ++t) {
      double kk2 = iter[t % k];
      double lam = (kk2 + 1.0) / (kk2 + 2.0);
      a[t] = lam * ((1.0 + rho) * p[t] - rho * a[t]) + (1.0 - lam) * o[t];
    }
  }
  void halpern(double rho) {  // :147-152
    halpernOne(X, XP, X0, rho);
    halpernOne(Y, YP, Y0, rho);
    halpernOne(AX, AXP, AX0, rho);
    for (auto& it : iter) ++it;
  }
  void restart(const std::vector<int>& mask) {  // :153-165
    auto cp = [&](int src, int d1, int d2) {
      for (size_t t = 0; t < v[src].size(); ++t)
        if (mask[t % k]) v[d1][t] = v[d2][t] = v[src][t];
    };
    cp(XP, X, X0);
    cp(YP, Y, Y0);
    cp(AXP, AX, AX0);
    for (int b = 0; b < k; ++b)
      if (mask[b]) iter[b] = 0;
  }
  void kkt(std::vector<double>& r3, std::vector<double>& c3) {  // :166-197
    r3.assign(3 * k, 0.0);
    c3.assign(3 * k, 0.0);
    auto &axp = v[AXP], &rl = v[RL], &ru = v[RU], &yp = v[YP];
    for (int i = 0; i < m; ++i)
      for (int b = 0; b < k; ++b) {
        size_t t = static_cast<size_t>(i) * k + b;
        double a = axp[t];
        double res = (std::max(rl[t] - a, 0.0) + std::max(a - ru[t], 0.0)) * rfac[i];
        r3[3 * b] += res * res;
        double y = yp[t];
        if (y > 0 && rl[t] > -1e300) r3[3 * b + 1] += y * rl[t];
        else if (y < 0 && ru[t] < 1e300) r3[3 * b + 1] += y * ru[t];
        r3[3 * b + 2] += y * y;
      }
    auto &xp = v[XP], &c = v[C], &atyp = v[ATYP], &lo = v[LO], &up = v[UP];
    for (int j = 0; j < n; ++j)
      for (int b = 0; b < k; ++b) {
        size_t t = static_cast<size_t>(j) * k + b;
        double d = c[t] - atyp[t];
        bool hl = lo[t] > -1e300, hu = up[t] < 1e300;
        double res = 0;
        if (!hl && !hu) res = std::fabs(d);
        else if (!hl) res = std::max(d, 0.0);
        else if (!hu) res = std::max(-d, 0.0);
        res *= cfac[j];
        c3[3 * b] += res * res;
        c3[3 * b + 1] += c[t] * xp[t];
        if (d > 0 && hl) c3[3 * b + 2] += d * lo[t];
        else if (d < 0 && hu) c3[3 * b + 2] += d * up[t];
      }
  }
  void diff2(int a, int b, std::vector<double>& out) {  // :198-205
    out.assign(k, 0.0);
    for (size_t t = 0; t < v[a].size(); ++t) {
      double d = v[a][t] - v[b][t];
      out[t % k] += d * d;
    }
  }
};

}  // namespace

extern "C" {

// Flat result record (ctypes mirror in tests/oracle_lib.py).
struct OracleResult {
  int status;  // 1 Optimal, 5 TimeLimit, 6 IterationLimit (src/util/common.h:34-46)
  int restarts;
  long long iterations;
  double objective, dual_objective, relative_kkt;
  double rel_primal, rel_dual, rel_gap, primal_weight, spectral_norm;
};

// Per-check trace record of instance 0 (for trajectory comparisons).
struct OracleTrace {
  long long it;
  double relP, relD, gap, pobj, dobj, omega;
  int restarted, pad;
};

const char* oracle_last_error() { return gErr.c_str(); }
int oracle_result_size() { return static_cast<int>(sizeof(OracleResult)); }
int oracle_trace_size() { return static_cast<int>(sizeof(OracleTrace)); }

// Scaling / norm of the engine set-up (src/pdhg/pdhg.cpp:424-459).  Outputs
// R[m], C[n], {beta, gamma, normA}, and the scaled CSR of A (ptr[m+1],
// idx[nnz], val[nnz]); any output pointer may be NULL.
int oracle_prepare(int m, int n, const int* start, const int* index, const double* value, const double* c,
                   const double* L, const double* U, int ruizIters, int pc, int rescale, double* R, double* C,
                   double* scalars, int* ptr, int* idx, double* val) {
  try {
    Lp lp{m, n, {}, c, nullptr, nullptr, L, U, 0.0};
    lp.A.rows = m;
    lp.A.cols = n;
    lp.A.start.assign(start, start + n + 1);
    lp.A.index.assign(index, index + start[n]);
    lp.A.value.assign(value, value + start[n]);
    Prepared P = prepare(lp, ruizIters, pc != 0, rescale != 0);
    if (R) std::copy(P.R.begin(), P.R.end(), R);
    if (C) std::copy(P.C.begin(), P.C.end(), C);
    if (scalars) {
      scalars[0] = P.beta;
      scalars[1] = P.gamma;
      scalars[2] = P.normA;
    }
    if (ptr) std::copy(P.A.ptr.begin(), P.A.ptr.end(), ptr);
    if (idx) std::copy(P.A.idx.begin(), P.A.idx.end(), idx);
    if (val) std::copy(P.A.val.begin(), P.A.val.end(), val);
    return 0;
  } catch (const std::exception& e) {
    gErr = e.what();
    return -1;
  }
}

// Restated runPdhg (src/pdhg/pdhg.cpp:461-687) on the serial CPU backend.
// k cases share A and bounds; costs (k x n row-major) may be NULL (= base c).
// trace (capacity traceCap records, instance 0) may be NULL.
int oracle_solve(int m, int n, const int* start, const int* index, const double* value, const double* c0,
                 const double* l0, const double* u0, const double* L0, const double* U0, double offset, int k,
                 const double* costs, double tol, long long maxIter, int checkEvery, OracleResult* res,
                 double* xs, double* ys, OracleTrace* trace, int traceCap, int* traceLen) {
  try {
    Lp lp{m, n, {}, c0, l0, u0, L0, U0, offset};
    lp.A.rows = m;
    lp.A.cols = n;
    lp.A.start.assign(start, start + n + 1);
    lp.A.index.assign(index, index + start[n]);
    lp.A.value.assign(value, value + start[n]);
    Prepared P = prepare(lp, 10, true, true);
    Cpu be;
    be.init(P, k);
    if (traceLen) *traceLen = 0;

    // Case vectors (:461-481) and scaled interleaved data (:506-545).
    std::vector<double> cNorm(k, 0.0), rhsNorm(k, 0.0);
    std::vector<double> cv(static_cast<size_t>(n) * k), lo(cv.size()), up(cv.size()), rl(static_cast<size_t>(m) * k),
        ru(rl.size()), x(cv.size()), y(rl.size(), 0.0);
    std::vector<double> omega(k, 1.0);
    for (int b = 0; b < k; ++b) {
      const double* cb = costs ? costs + static_cast<size_t>(b) * n : c0;
      for (int j = 0; j < n; ++j) cNorm[b] += cb[j] * cb[j];
      cNorm[b] = std::sqrt(cNorm[b]);
      for (int i = 0; i < m; ++i) {
        if (fin(L0[i])) rhsNorm[b] += L0[i] * L0[i];
        if (fin(U0[i]) && U0[i] != L0[i]) rhsNorm[b] += U0[i] * U0[i];
      }
      rhsNorm[b] = std::sqrt(rhsNorm[b]);
      double cn = 0, bn = 0;
      for (int j = 0; j < n; ++j) {
        size_t t = static_cast<size_t>(j) * k + b;
        cv[t] = cb[j] * P.C[j] / P.gamma;
        lo[t] = fin(l0[j]) ? l0[j] / P.C[j] / P.beta : -HUGE_VAL;
        up[t] = fin(u0[j]) ? u0[j] / P.C[j] / P.beta : HUGE_VAL;
        x[t] = std::min(std::max(0.0, lo[t]), up[t]);
        cn += cv[t] * cv[t];
      }
      for (int i = 0; i < m; ++i) {
        size_t t = static_cast<size_t>(i) * k + b;
        rl[t] = fin(L0[i]) ? L0[i] * P.R[i] / P.beta : -HUGE_VAL;
        ru[t] = fin(U0[i]) ? U0[i] * P.R[i] / P.beta : HUGE_VAL;
        if (fin(rl[t])) bn += rl[t] * rl[t];
        if (fin(ru[t]) && ru[t] != rl[t]) bn += ru[t] * ru[t];
      }
      if (cn > 1e-20 && bn > 1e-20) omega[b] = std::sqrt(cn) / std::sqrt(bn);
    }
    be.v[C] = cv;
    be.v[LO] = lo;
    be.v[UP] = up;
    be.v[RL] = rl;
    be.v[RU] = ru;
    be.v[X] = x;
    be.v[X0] = x;
    be.v[Y] = y;
    be.v[Y0] = y;
    be.spmv(X, AX);
    be.v[AX0] = be.v[AX];
    const double eta = 0.998 / P.normA;
    auto updSteps = [&]() {
      for (int b = 0; b < k; ++b) {
        be.tau[b] = eta / omega[b];
        be.sigma[b] = eta * omega[b];
      }
    };
    updSteps();

    std::vector<char> done(k, 0);
    std::vector<double> fpR(k, kInf), fpPrev(k, kInf);
    std::vector<long long> since(k, 0);
    std::vector<int> restarts(k, 0);
    int remaining = k;
    long long it = 0;
    std::vector<double> r3, c3, dx2, dy2, ax2, ay2;
    auto extract = [&](int b, int status, double kkt, double pobj, double dobj, double rp, double rd, double g) {
      OracleResult& r = res[b];
      r.status = status;
      r.iterations = it;
      r.relative_kkt = kkt;
      r.objective = pobj + P.offset;
      r.dual_objective = dobj + P.offset;
      r.rel_primal = rp;
      r.rel_dual = rd;
      r.rel_gap = g;
      r.restarts = restarts[b];
      r.primal_weight = omega[b];
      r.spectral_norm = P.normA;
      if (xs)
        for (int j = 0; j < n; ++j) xs[static_cast<size_t>(b) * n + j] = be.v[XP][static_cast<size_t>(j) * k + b] * P.C[j] * P.beta;
      if (ys)
        for (int i = 0; i < m; ++i) ys[static_cast<size_t>(b) * m + i] = be.v[YP][static_cast<size_t>(i) * k + b] * P.R[i] * P.gamma;
    };
    for (; it < maxIter; ++it) {
      be.spmvT(Y, ATY);
      be.primal();
      be.spmv(XP, AXP);
      be.dual();
      const bool check = ((it + 1) % checkEvery) == 0 || it + 1 == maxIter;
      std::vector<int> mask;
      if (check) {
        be.spmvT(YP, ATYP);
        be.kkt(r3, c3);
        be.diff2(X, XP, dx2);
        be.diff2(Y, YP, dy2);
        mask.assign(k, 0);
        bool any = false;
        for (int b = 0; b < k; ++b) {
          if (done[b]) continue;
          double so = P.beta * P.gamma;
          double pobj = c3[3 * b + 1] * so;
          double dobj = (r3[3 * b + 1] + c3[3 * b + 2]) * so;
          double relP = std::sqrt(r3[3 * b]) / (1.0 + rhsNorm[b]);
          double relD = std::sqrt(c3[3 * b]) / (1.0 + cNorm[b]);
          double gap = std::fabs(pobj - dobj) / (1.0 + std::fabs(pobj) + std::fabs(dobj));
          double kkt = std::max(relP, std::max(relD, gap));
          if (b == 0 && trace && *traceLen < traceCap) {
            OracleTrace& tr = trace[(*traceLen)++];
            tr.it = it;
            tr.relP = relP;
            tr.relD = relD;
            tr.gap = gap;
            tr.pobj = pobj + P.offset;
            tr.dobj = dobj + P.offset;
            tr.omega = omega[0];
            tr.restarted = 0;
          }
          if (relP <= tol && relD <= tol && gap <= tol) {
            extract(b, 1, kkt, pobj, dobj, relP, relD, gap);
            done[b] = 1;
            --remaining;
            continue;
          }
          since[b] += checkEvery;
          double fp = std::sqrt(omega[b] * dx2[b] + dy2[b] / omega[b]);
          if (fpR[b] == kInf) fpR[b] = fp;
          bool rs = fp <= 0.2 * fpR[b] || (fp <= 0.8 * fpR[b] && fp > fpPrev[b]) ||
                    since[b] >= static_cast<long long>(0.36 * (it + 1));
          fpPrev[b] = fp;
          if (rs) {
            mask[b] = 1;
            any = true;
          }
        }
        if (any) {
          be.diff2(XP, X0, ax2);
          be.diff2(YP, Y0, ay2);
          for (int b = 0; b < k; ++b) {
            if (!mask[b]) continue;
            double dxn = std::sqrt(ax2[b]), dyn = std::sqrt(ay2[b]);
            if (dxn > 1e-10 && dyn > 1e-10) omega[b] = std::exp(0.5 * std::log(dyn / dxn) + 0.5 * std::log(omega[b]));
            fpR[b] = std::sqrt(omega[b] * dx2[b] + dy2[b] / omega[b]);
            fpPrev[b] = kInf;
            since[b] = 0;
            restarts[b]++;
            if (b == 0 && trace && *traceLen > 0) {
              trace[*traceLen - 1].restarted = 1;
              trace[*traceLen - 1].omega = omega[0];
            }
          }
          updSteps();
        }
        if (remaining == 0) break;
      }
      be.halpern(1.0);
      if (!mask.empty() && std::find(mask.begin(), mask.end(), 1) != mask.end()) be.restart(mask);
    }
    if (remaining > 0) {  // :668-680
      be.spmvT(YP, ATYP);
      be.kkt(r3, c3);
      for (int b = 0; b < k; ++b) {
        if (done[b]) continue;
        double so = P.beta * P.gamma;
        double pobj = c3[3 * b + 1] * so, dobj = (r3[3 * b + 1] + c3[3 * b + 2]) * so;
        double relP = std::sqrt(r3[3 * b]) / (1 + rhsNorm[b]), relD = std::sqrt(c3[3 * b]) / (1 + cNorm[b]);
        double gap = std::fabs(pobj - dobj) / (1 + std::fabs(pobj) + std::fabs(dobj));
        extract(b, 6, std::max(relP, std::max(relD, gap)), pobj, dobj, relP, relD, gap);
      }
    }
    return 0;
  } catch (const std::exception& e) {
    gErr = e.what();
    return -1;
  }
}

// ---------------------------------------------------------------------------
// Safe dual bound (src/cert/interval.h:14-91, src/cert/certifier.cpp:28-118).
// Must be compiled without FMA contraction (-ffp-contract=off).
// ---------------------------------------------------------------------------
static double rdn(double v) { return (std::isinf(v) || std::isnan(v)) ? v : std::nextafter(v, -kInf); }
static double rup(double v) { return (std::isinf(v) || std::isnan(v)) ? v : std::nextafter(v, kInf); }
static double twoSumErr(double a, double b, double s) {
  double bb = s - a;
  return (a - (s - bb)) + (b - bb);
}
static double addDn(double a, double b) {
  double s = a + b;
  if (std::isinf(s) || std::isnan(s)) return s;
  return twoSumErr(a, b, s) < 0 ? rdn(s) : s;
}
static double addUpI(double a, double b) {
  double s = a + b;
  if (std::isinf(s) || std::isnan(s)) return s;
  return twoSumErr(a, b, s) > 0 ? rup(s) : s;
}
static double mulDn(double a, double b) {
  if (a == 0 || b == 0) return 0.0;
  double p = a * b;
  if (std::isinf(p)) return p;
  return std::fma(a, b, -p) < 0 ? rdn(p) : p;
}
static double mulUpI(double a, double b) {
  if (a == 0 || b == 0) return 0.0;
  double p = a * b;
  if (std::isinf(p)) return p;
  return std::fma(a, b, -p) > 0 ? rup(p) : p;
}
static double minProd(double zlo, double zhi, double l, double u) {
  if (zlo == 0 && zhi == 0) return 0.0;
  if (std::isinf(l) && zhi > 0) return -kInf;
  if (std::isinf(u) && zlo < 0) return -kInf;
  double best = kInf;
  const double zs[2] = {zlo, zhi}, xs[2] = {l, u};
  for (double zz : zs)
    for (double xx : xs) {
      double p;
      if (zz == 0 || xx == 0) p = 0.0;
      else if (std::isinf(xx)) p = zz * xx;
      else p = mulDn(zz, xx);
      if (p < best) best = p;
    }
  return best;
}
static double clean(double v) { return fin(v) ? v : (v > 0 ? kInf : -kInf); }

int oracle_safe_dual_bound(int m, int n, const int* start, const int* index, const double* value, const double* c,
                           const double* l, const double* u, const double* L, const double* U, const double* yIn,
                           double* bound, int* infiniteTerms, double* costPerturbation) {
  std::vector<double> y(yIn, yIn + m);
  for (int i = 0; i < m; ++i) {  // projectRowDuals, certifier.cpp:41-56
    bool lo = fin(L[i]), up = fin(U[i]);
    if (!lo && !up) y[i] = 0;
    else if (!lo) y[i] = std::min(y[i], 0.0);
    else if (!up) y[i] = std::max(y[i], 0.0);
  }
  double lb = 0, pert = 0;
  int inf = 0;
  bool minusInf = false;
  for (int i = 0; i < m; ++i) lb = addDn(lb, minProd(y[i], y[i], clean(L[i]), clean(U[i])));
  for (int j = 0; j < n; ++j) {
    double zlo = c[j], zhi = c[j];  // reducedCostIntervals, certifier.cpp:28-37
    for (int e = start[j]; e < start[j + 1]; ++e) {
      double plo = mulDn(value[e], y[index[e]]), phi = mulUpI(value[e], y[index[e]]);
      zlo = addDn(zlo, -phi);
      zhi = addUpI(zhi, -plo);
    }
    double shift = 0;
    if (!fin(l[j]) && zhi > 0) {
      shift = std::max(shift, zhi);
      zhi = 0;
      zlo = std::min(zlo, 0.0);
    }
    if (!fin(u[j]) && zlo < 0) {
      shift = std::max(shift, -zlo);
      zlo = 0;
      zhi = std::max(zhi, 0.0);
    }
    if (shift > 0) {
      double rel = shift / (1.0 + std::fabs(c[j]));
      pert = std::max(pert, rel);
      if (rel > 1e-9) {
        inf++;
        minusInf = true;
        continue;
      }
    }
    double t = minProd(zlo, zhi, clean(l[j]), clean(u[j]));
    if (std::isinf(t)) {
      inf++;
      minusInf = true;
      continue;
    }
    lb = addDn(lb, t);
  }
  *bound = minusInf ? -kInf : lb;
  *infiniteTerms = inf;
  *costPerturbation = pert;
  return 0;
}

// ---- Backend ops one at a time (the Cpu struct above, i.e. CpuBackend,
// src/pdhg/pdhg.cpp:75-214), for op-level comparison with the device backend.
struct OracleOps {
  Csr A, AT;
  Prepared P;  // only rfac / cfac / m / n are used by Cpu::init
  Cpu cpu;
};

void* oracle_ops_create(int m, int n, const int* aPtr, const int* aIdx, const double* aVal, const int* tPtr,
                        const int* tIdx, const double* tVal, const double* rfac, const double* cfac, int k) {
  if (m < 0 || n < 0 || k < 1) {
    gErr = "bad sizes";
    return nullptr;
  }
  auto* h = new OracleOps;
  h->P.m = m;
  h->P.n = n;
  h->P.A.rows = m;
  h->P.A.cols = n;
  h->P.A.ptr.assign(aPtr, aPtr + m + 1);
  h->P.A.idx.assign(aIdx, aIdx + aPtr[m]);
  h->P.A.val.assign(aVal, aVal + aPtr[m]);
  h->P.AT.rows = n;
  h->P.AT.cols = m;
  h->P.AT.ptr.assign(tPtr, tPtr + n + 1);
  h->P.AT.idx.assign(tIdx, tIdx + tPtr[n]);
  h->P.AT.val.assign(tVal, tVal + tPtr[n]);
  h->P.rfac.assign(rfac, rfac + m);
  h->P.cfac.assign(cfac, cfac + n);
  h->cpu.init(h->P, k);
  return h;
}
void oracle_ops_destroy(void* h) { delete static_cast<OracleOps*>(h); }
int oracle_ops_set(void* h, int vec, const double* d) {
  auto& c = static_cast<OracleOps*>(h)->cpu;
  if (vec < 0 || vec >= NV) return -1;
  std::copy(d, d + c.v[vec].size(), c.v[vec].begin());
  return 0;
}
int oracle_ops_get(void* h, int vec, double* d) {
  auto& c = static_cast<OracleOps*>(h)->cpu;
  if (vec < 0 || vec >= NV) return -1;
  std::copy(c.v[vec].begin(), c.v[vec].end(), d);
  return 0;
}
void oracle_ops_spmv(void* h, int in, int out) { static_cast<OracleOps*>(h)->cpu.spmv(in, out); }
void oracle_ops_spmvT(void* h, int in, int out) { static_cast<OracleOps*>(h)->cpu.spmvT(in, out); }
void oracle_ops_set_steps(void* h, const double* tau, const double* sigma) {  // src/pdhg/pdhg.cpp:106-109
  auto& c = static_cast<OracleOps*>(h)->cpu;
  c.tau.assign(tau, tau + c.k);
  c.sigma.assign(sigma, sigma + c.k);
}
void oracle_ops_primal(void* h) { static_cast<OracleOps*>(h)->cpu.primal(); }
void oracle_ops_dual(void* h) { static_cast<OracleOps*>(h)->cpu.dual(); }
void oracle_ops_halpern(void* h, double rho) { static_cast<OracleOps*>(h)->cpu.halpern(rho); }
void oracle_ops_restart(void* h, const int* mask) {
  auto& c = static_cast<OracleOps*>(h)->cpu;
  c.restart(std::vector<int>(mask, mask + c.k));
}
void oracle_ops_kkt(void* h, double* r3, double* c3) {
  auto& c = static_cast<OracleOps*>(h)->cpu;
  std::vector<double> r, cc;
  c.kkt(r, cc);
  std::copy(r.begin(), r.end(), r3);
  std::copy(cc.begin(), cc.end(), c3);
}
void oracle_ops_diff2(void* h, int a, int b, double* out) {
  auto& c = static_cast<OracleOps*>(h)->cpu;
  std::vector<double> o;
  c.diff2(a, b, o);
  std::copy(o.begin(), o.end(), out);
}

}  // extern "C"
