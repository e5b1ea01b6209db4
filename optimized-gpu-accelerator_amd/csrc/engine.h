// Host engine: the reference's solvePdhg / solvePdhgBatch
// (src/pdhg/pdhg.h:73-77) over the HIP device run.
#pragma once

#include <string>
#include <vector>

#include "../../include/pdhg_hip.h"
#include "core.h"

namespace pdhgamd {

// One case of a family (LpCase, src/pdhg/pdhg.h:68-70); NULL = base model.
struct CaseData {
  const double *c = nullptr, *l = nullptr, *u = nullptr, *L = nullptr, *U = nullptr;
  const double *x0 = nullptr, *y0 = nullptr;  // warm start (original space)
  double omega0 = 0;                            // > 0: initial primal weight
};

// Runs all cases (any count; chunks of at most 64 share one set-up) and fills
// results[k] and stats.
void solve_batch(const Model& base, const std::vector<CaseData>& cases, const pdhg_options& o, pdhg_result* results,
                 pdhg_stats* stats);

// A copy of a C-ABI LP, checked as the reference's Model::validate does
// (src/core/model.cpp:103-116): no NaN in column data, finite coefficients;
// plus the checks of the ABI itself (non-NULL arrays, index ranges).
Model model_from(const pdhg_lp* lp);

// Rejects options outside their documented ranges, before any device work.
void check_options(const pdhg_options& o);

// Unrelated LPs advanced together by the persistent packed kernels.
void solve_packed(const std::vector<Model>& lps, const pdhg_options& o, pdhg_result* results, pdhg_stats* stats);

}  // namespace pdhgamd
