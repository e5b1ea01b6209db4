// Certifier handoff: rigorous Neumaier-Shcherbina dual bound with outward-
// rounded FP64 interval arithmetic, activity-based implied bounds, and the
// family acceptance rule of the reference.  Built with -ffp-contract=off: the
// error-free transformations below are exact only without FMA contraction.
#pragma once

#include "core.h"

namespace pdhgamd {

struct SafeBound {
  double bound = -kInf;      // rigorous lower bound on min c'x (any y)
  int infiniteTerms = 0;     // columns that forced -inf
  double costPerturbation = 0;
  double rowProjection = 0;
};

// src/cert/certifier.cpp:41-118 (projectRowDuals, reducedCostIntervals,
// safeBoundInterval with costLo == costHi).
SafeBound safe_dual_bound(const Csc& A, const double* c, const double* l, const double* u, const double* L,
                          const double* U, const double* y);

// Bounds implied by the rows (activity propagation over continuous columns,
// 30 rounds), each relaxed by 1e-6 (1 + |bound|): src/parametric/parametric.cpp:62-73
// with the continuous branch of src/mip/propagate.cpp:22-141.
void implied_bounds(const Csc& A, const double* L, const double* U, std::vector<double>& lo,
                    std::vector<double>& up);

// Largest relative violation of column bounds and (compensated) row activities,
// as certify's primal check (src/cert/certifier.cpp:218-260).
double max_primal_violation(const Csc& A, const double* l, const double* u, const double* L, const double* U,
                            const double* x);

struct FamilyCert {
  double bound = -kInf;
  double relGap = kInf;      // |obj - LB| / (1 + |obj|)
  double primalViolation = 0;
  bool accepted = false;
};

// Acceptance of one family case (src/parametric/parametric.cpp:509-524):
// implied bounds, safe bound on them, |obj - LB| <= 1e-4 (1 + |obj|).
FamilyCert certify_case(const Csc& A, const double* c, const double* l, const double* u, const double* L,
                        const double* U, double offset, const double* x, const double* y, double objective);

}  // namespace pdhgamd
