// Test infrastructure only: a flat extern "C" wrapper around the reference
// solver's own PDHG engine and certifier, compiled from the reference sources
// where they lie (see oracle/Makefile).  Only tests/, the smoke check and the
// CPU-baseline leg of bench.py load the resulting oracle/_ref library, and only
// as a checker / CPU baseline, never as part of the product path.
//
// Wraps:  pramana::solvePdhg / solvePdhgBatch   (src/pdhg/pdhg.h:73-77)
//         pramana::safeDualBound                 (src/cert/certifier.h:198-201)
//         pramana::ThreadPool::setThreads        (src/util/parallel.h:24)
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include <cmath>

#include "cert/certifier.h"
#include "mip/propagate.h"
#include "pdhg/pdhg.h"
#include "util/parallel.h"

namespace {

thread_local std::string gErr;

pramana::Model makeModel(int m, int n, const int* start, const int* index, const double* value, const double* c,
                         const double* l, const double* u, const double* L, const double* U, double offset) {
  pramana::Model md;
  md.colCost.assign(c, c + n);
  md.colLower.assign(l, l + n);
  md.colUpper.assign(u, u + n);
  md.rowLower.assign(L, L + m);
  md.rowUpper.assign(U, U + m);
  md.colType.assign(n, pramana::VarType::Continuous);
  md.objOffset = offset;
  md.A = pramana::SparseMatrix(m, n);
  md.A.start.assign(start, start + n + 1);
  const int nnz = start[n];
  md.A.index.assign(index, index + nnz);
  md.A.value.assign(value, value + nnz);
  return md;
}

// Result record, kept flat for ctypes.
struct RefResult {
  int status;
  int restarts;
  long long iterations;
  double objective, dual_objective, relative_kkt;
  double primal_residual, dual_residual, gap, primal_weight, spectral_norm;
  double setup_seconds, solve_seconds;
};

void fill(const pramana::PdhgResult& r, RefResult* out, double* x, double* y) {
  out->status = static_cast<int>(r.status);
  out->restarts = r.stats.restarts;
  out->iterations = r.iterations;
  out->objective = r.objective;
  out->dual_objective = r.dualObjective;
  out->relative_kkt = r.relativeKkt;
  out->primal_residual = r.stats.primalResidual;
  out->dual_residual = r.stats.dualResidual;
  out->gap = r.stats.gap;
  out->primal_weight = r.stats.primalWeight;
  out->spectral_norm = r.stats.spectralNorm;
  out->setup_seconds = r.stats.setupSeconds;
  out->solve_seconds = r.stats.solveSeconds;
  if (x) std::memcpy(x, r.x.data(), r.x.size() * sizeof(double));
  if (y) std::memcpy(y, r.rowDual.data(), r.rowDual.size() * sizeof(double));
}

}  // namespace

extern "C" {

const char* ref_last_error() { return gErr.c_str(); }

int ref_set_threads(int n) {
  // The reference pool is created on first use; this must run before any solve.
  pramana::ThreadPool::setThreads(n);
  return pramana::ThreadPool::instance().threads();
}

int ref_result_size() { return static_cast<int>(sizeof(RefResult)); }

// Single solve on the reference CPU backend (FP64).
int ref_solve_pdhg(int m, int n, const int* start, const int* index, const double* value, const double* c,
                   const double* l, const double* u, const double* L, const double* U, double offset, double tol,
                   long long maxIter, int checkEvery, RefResult* out, double* x, double* y) {
  try {
    setenv("PRAMANA_DISABLE_GPU", "1", 1);
    pramana::Model md = makeModel(m, n, start, index, value, c, l, u, L, U, offset);
    pramana::PdhgOptions po;
    po.tolerance = tol;
    po.maxIterations = maxIter;
    po.checkEvery = checkEvery;
    po.useGpu = false;
    pramana::PdhgResult r = pramana::solvePdhg(md, po);
    fill(r, out, x, y);
    return 0;
  } catch (const std::exception& e) {
    gErr = e.what();
    return -1;
  }
}

// Batched solve: k cases that differ only in cost (costs given k x n, row-major).
int ref_solve_batch_cost(int m, int n, const int* start, const int* index, const double* value, const double* c,
                         const double* l, const double* u, const double* L, const double* U, double offset, int k,
                         const double* costs, double tol, long long maxIter, RefResult* out, double* xs,
                         double* ys) {
  try {
    setenv("PRAMANA_DISABLE_GPU", "1", 1);
    pramana::Model md = makeModel(m, n, start, index, value, c, l, u, L, U, offset);
    std::vector<pramana::LpCase> cases(k);
    for (int b = 0; b < k; ++b) cases[b].cost.assign(costs + static_cast<size_t>(b) * n, costs + static_cast<size_t>(b + 1) * n);
    pramana::PdhgOptions po;
    po.tolerance = tol;
    po.maxIterations = maxIter;
    po.useGpu = false;
    pramana::PdhgStats st;
    std::vector<pramana::PdhgResult> rr = pramana::solvePdhgBatch(md, cases, po, nullptr, &st);
    for (int b = 0; b < k; ++b)
      fill(rr[b], out + b, xs ? xs + static_cast<size_t>(b) * n : nullptr, ys ? ys + static_cast<size_t>(b) * m : nullptr);
    return 0;
  } catch (const std::exception& e) {
    gErr = e.what();
    return -1;
  }
}

// Neumaier-Shcherbina safe dual bound of the reference certifier.
int ref_safe_dual_bound(int m, int n, const int* start, const int* index, const double* value, const double* c,
                        const double* l, const double* u, const double* L, const double* U, const double* y,
                        double* bound, int* infiniteTerms, double* costPerturbation) {
  try {
    pramana::SparseMatrix A(m, n);
    A.start.assign(start, start + n + 1);
    A.index.assign(index, index + start[n]);
    A.value.assign(value, value + start[n]);
    pramana::SafeBoundResult r =
        pramana::safeDualBound(A, std::vector<double>(c, c + n), std::vector<double>(l, l + n),
                               std::vector<double>(u, u + n), std::vector<double>(L, L + m),
                               std::vector<double>(U, U + m), std::vector<double>(y, y + m));
    *bound = r.bound;
    *infiniteTerms = r.infiniteTerms;
    *costPerturbation = r.costPerturbation;
    return 0;
  } catch (const std::exception& e) {
    gErr = e.what();
    return -1;
  }
}

// Family certification bound of the reference: impliedBounds
// (src/parametric/parametric.cpp:62-73, which is file-local there, so its
// four statements are spelled out here around the reference's own Propagator)
// followed by safeDualBound, plus the objective offset (:516-517).
int ref_family_bound(int m, int n, const int* start, const int* index, const double* value, const double* c,
                     const double* l, const double* u, const double* L, const double* U, double offset,
                     const double* y, double* bound) {
  try {
    pramana::Model md = makeModel(m, n, start, index, value, c, l, u, L, U, offset);
    std::vector<double> lo = md.colLower, up = md.colUpper;
    pramana::Propagator p;
    p.setup(md);
    std::vector<double> l2 = lo, u2 = up;
    p.propagate(l2, u2, 30);
    auto fin = [](double v) { return std::fabs(v) < pramana::kInfBoundThreshold; };
    for (size_t j = 0; j < lo.size(); ++j) {
      if (fin(l2[j]) && (!fin(lo[j]) || l2[j] > lo[j])) lo[j] = l2[j] - 1e-6 * (1 + std::fabs(l2[j]));
      if (fin(u2[j]) && (!fin(up[j]) || u2[j] < up[j])) up[j] = u2[j] + 1e-6 * (1 + std::fabs(u2[j]));
    }
    pramana::SafeBoundResult sb = pramana::safeDualBound(md.A, md.colCost, lo, up, md.rowLower, md.rowUpper,
                                                         std::vector<double>(y, y + m));
    *bound = sb.bound + offset;
    return 0;
  } catch (const std::exception& e) {
    gErr = e.what();
    return -1;
  }
}

}  // extern "C"
