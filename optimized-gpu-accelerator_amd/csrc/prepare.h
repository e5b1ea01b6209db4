// Host set-up of a PDHG run: diagonal preconditioning, objective / bound
// rescaling, CSR copies of the scaled matrix and the spectral-norm estimate.
#pragma once

#include "core.h"

namespace pdhgamd {

struct Prepared {
  int m = 0, n = 0;
  Csr A, AT;                      // scaled A (m rows) and A^T (n rows)
  std::vector<double> R, C;       // Ruiz/Pock-Chambolle factors (powers of two)
  std::vector<double> rfac, cfac; // beta / R, gamma / C (unscaling of residuals)
  double beta = 1, gamma = 1;     // bound / objective rescaling
  double normA = 1;               // 1.01 * power-iteration estimate of ||A_scaled||_2
  double offset = 0;
};

// Scaling of the reference engine (src/pdhg/pdhg.cpp:424-459,
// src/lp/scaling.cpp:9-32,97-139).  The floating-point operations are issued
// in the reference's order so that the factors, beta, gamma and the norm
// estimate are bit-identical to it (this file is built with -ffp-contract=off).
Prepared prepare(const Model& lp, int ruizIterations, bool pockChambolle, bool rescale);

}  // namespace pdhgamd
