#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "device.h"
#include "prepare.h"

namespace pdhgamd {

namespace {

using Clock = std::chrono::steady_clock;
double since_s(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// All cases point at the same data (or all NULL): the array can be stored once.
bool same_pointer(const std::vector<CaseData>& cs, const double* CaseData::*f) {
  for (const auto& c : cs)
    if (c.*f != cs[0].*f) return false;
  return true;
}

}  // namespace

Model model_from(const pdhg_lp* lp) {
  PDHG_REQUIRE(lp && lp->m >= 0 && lp->n >= 0, "bad LP");
  Model md;
  md.m = lp->m;
  md.n = lp->n;
  md.A.rows = lp->m;
  md.A.cols = lp->n;
  md.A.start.assign(lp->col_start, lp->col_start + lp->n + 1);
  const int nnz = md.A.start.back();
  PDHG_REQUIRE(md.A.start[0] == 0 && nnz >= 0, "bad column pointers");
  for (int j = 0; j < lp->n; ++j) PDHG_REQUIRE(md.A.start[j] <= md.A.start[j + 1], "column pointers not monotone");
  md.A.index.assign(lp->row_index, lp->row_index + nnz);
  md.A.value.assign(lp->value, lp->value + nnz);
  for (int e = 0; e < nnz; ++e) {
    PDHG_REQUIRE(md.A.index[e] >= 0 && md.A.index[e] < lp->m, "row index out of range");
    PDHG_REQUIRE(std::isfinite(md.A.value[e]), "non-finite matrix coefficient");
  }
  md.c.assign(lp->cost, lp->cost + lp->n);
  md.l.assign(lp->col_lower, lp->col_lower + lp->n);
  md.u.assign(lp->col_upper, lp->col_upper + lp->n);
  md.L.assign(lp->row_lower, lp->row_lower + lp->m);
  md.U.assign(lp->row_upper, lp->row_upper + lp->m);
  md.offset = lp->obj_offset;
  return md;
}

void solve_batch(const Model& base, const std::vector<CaseData>& cases, const pdhg_options& o, pdhg_result* results,
                 pdhg_stats* stats) {
  const auto tSetup = Clock::now();
  const int k = static_cast<int>(cases.size());
  PDHG_REQUIRE(k >= 1, "no cases");
  PDHG_REQUIRE(o.use_gpu, "use_gpu = 0: this engine has no CPU backend");
  PDHG_REQUIRE(o.check_every >= 2, "check_every must be >= 2");
  PDHG_REQUIRE(o.max_iterations >= 1, "max_iterations must be >= 1");
  const DeviceInfo& di = device_info(o.device);
  if (!di.available) throw Error("GPU unavailable: " + di.reason);
  const int n = base.n, m = base.m;

  // Set-up shared by every chunk (src/pdhg/pdhg.cpp:487).
  const Prepared P = prepare(base, o.ruiz_iterations, o.pock_chambolle != 0, o.rescale != 0);
  const double eta = 0.998 / P.normA;
  double setupSeconds = since_s(tSetup);

  const bool shC = same_pointer(cases, &CaseData::c);
  const bool shB = same_pointer(cases, &CaseData::l) && same_pointer(cases, &CaseData::u);
  const bool shR = same_pointer(cases, &CaseData::L) && same_pointer(cases, &CaseData::U);

  pdhg_stats acc{};
  if (stats) {
    acc.trace = stats->trace;
    acc.trace_capacity = stats->trace_capacity;
  }
  acc.precision = o.precision;
  long long maxItersRun = 0;
  double solveSeconds = 0, kernelSeconds = 0, transfer = 0;
  int chunks = 0;

  for (int first = 0; first < k; first += 64) {
    const auto tChunk = Clock::now();
    const int kc = std::min(64, k - first);
    const int B = next_pow2(kc);
    ++chunks;
    // Case data in scaled space, interleaved (src/pdhg/pdhg.cpp:506-545).
    const size_t nB = static_cast<size_t>(n) * B, mB = static_cast<size_t>(m) * B;
    std::vector<double> c(shC ? n : nB), lo(shB ? n : nB), up(shB ? n : nB), rl(shR ? m : mB), ru(shR ? m : mB);
    std::vector<double> x(nB), y(mB, 0.0), omega(B, 1.0), cNorm(B, 0.0), rhsNorm(B, 0.0);
    std::vector<int> pad(B, 0);
    for (int b = 0; b < B; ++b) {
      const bool real = b < kc;
      pad[b] = real ? 0 : 1;
      const CaseData& cs = cases[first + (real ? b : 0)];
      const double* cc = cs.c ? cs.c : base.c.data();
      const double* cl = cs.l ? cs.l : base.l.data();
      const double* cu = cs.u ? cs.u : base.u.data();
      const double* cL = cs.L ? cs.L : base.L.data();
      const double* cU = cs.U ? cs.U : base.U.data();
      // Norms of the original case vectors (src/pdhg/pdhg.cpp:466-481).
      double cn0 = 0, bn0 = 0;
      for (int j = 0; j < n; ++j) cn0 += cc[j] * cc[j];
      for (int i = 0; i < m; ++i) {
        if (finite_bound(cL[i])) bn0 += cL[i] * cL[i];
        if (finite_bound(cU[i]) && cU[i] != cL[i]) bn0 += cU[i] * cU[i];
      }
      cNorm[b] = std::sqrt(cn0);
      rhsNorm[b] = std::sqrt(bn0);
      double cn = 0, bn = 0;
      for (int j = 0; j < n; ++j) {
        const double cj = cc[j] * P.C[j] / P.gamma;
        const double lj = finite_bound(cl[j]) ? cl[j] / P.C[j] / P.beta : -HUGE_VAL;
        const double uj = finite_bound(cu[j]) ? cu[j] / P.C[j] / P.beta : HUGE_VAL;
        if (shC) {
          if (b == 0) c[j] = cj;
        } else {
          c[static_cast<size_t>(j) * B + b] = cj;
        }
        if (shB) {
          if (b == 0) {
            lo[j] = lj;
            up[j] = uj;
          }
        } else {
          lo[static_cast<size_t>(j) * B + b] = lj;
          up[static_cast<size_t>(j) * B + b] = uj;
        }
        // Cold start x0 = clamp(0) (src/pdhg/pdhg.cpp:519) or the warm start
        // mapped into scaled space (inverse of the unscaling at :581) and clamped.
        const double x0 = cs.x0 ? cs.x0[j] / P.C[j] / P.beta : 0.0;
        x[static_cast<size_t>(j) * B + b] = std::min(std::max(x0, lj), uj);
        cn += cj * cj;
      }
      for (int i = 0; i < m; ++i) {
        const double li = finite_bound(cL[i]) ? cL[i] * P.R[i] / P.beta : -HUGE_VAL;
        const double ui = finite_bound(cU[i]) ? cU[i] * P.R[i] / P.beta : HUGE_VAL;
        if (shR) {
          if (b == 0) {
            rl[i] = li;
            ru[i] = ui;
          }
        } else {
          rl[static_cast<size_t>(i) * B + b] = li;
          ru[static_cast<size_t>(i) * B + b] = ui;
        }
        if (finite_bound(li)) bn += li * li;
        if (finite_bound(ui) && ui != li) bn += ui * ui;
        if (cs.y0) y[static_cast<size_t>(i) * B + b] = cs.y0[i] / P.R[i] / P.gamma;
      }
      if (cn > 1e-20 && bn > 1e-20) omega[b] = std::sqrt(cn) / std::sqrt(bn);
    }

    RunSetup rs;
    rs.m = m;
    rs.n = n;
    rs.B = B;
    rs.precision = o.precision;
    rs.sharedCost = shC;
    rs.sharedColBounds = shB;
    rs.sharedRowBounds = shR;
    rs.rho = o.reflection;
    rs.checkEvery = o.check_every;
    rs.useGraph = o.use_graph != 0;
    for (MapSetup* ms : {&rs.colMap, &rs.rowMap}) {
      ms->nnzPerLane = o.lanes_per_row > 0 ? o.lanes_per_row : 4;
      ms->rowMap = o.row_map == 2 ? 2 : 1;
      ms->xcd = o.xcd_remap == 1 || (o.xcd_remap < 0 && B >= 16);
    }
    auto run = DeviceRun::create(o.device, P, rs);
    run->upload(c.data(), lo.data(), up.data(), rl.data(), ru.data(), x.data(), y.data());
    RunScalars sc;
    sc.eta = eta;
    sc.beta = P.beta;
    sc.gamma = P.gamma;
    sc.offset = P.offset;
    sc.tol = o.tolerance;
    run->init_state(sc, omega.data(), cNorm.data(), rhsNorm.data(), pad.data());
    transfer += since_s(tChunk);
    if (first == 0) setupSeconds = since_s(tSetup);

    // ---- iterations (src/pdhg/pdhg.cpp:588-662), one check per block ----
    const auto tSolve = Clock::now();
    run->start();
    std::vector<CheckRecord> rec(B);
    std::vector<int> done(B), restarts(B);
    std::vector<long long> doneIter(B);
    long long it = 0;
    int checks = 0;
    bool expired = false;
    int remaining = kc;
    while (it < o.max_iterations) {
      const int nIt = static_cast<int>(std::min<long long>(o.check_every, o.max_iterations - it));
      run->run_block(nIt);
      run->fetch(rec.data(), done.data(), doneIter.data(), restarts.data());
      it += nIt;
      ++checks;
      remaining = 0;
      for (int b = 0; b < kc; ++b) remaining += done[b] ? 0 : 1;
      if (first == 0 && acc.trace && acc.trace_length < acc.trace_capacity &&
          static_cast<long long>(rec[0].it) == it - 1) {
        std::memcpy(acc.trace + 8 * acc.trace_length, &rec[0], sizeof(CheckRecord));
        acc.trace_length++;
      }
      if (o.log_interval > 0 && kc == 1 && checks % o.log_interval == 0)
        std::fprintf(stderr, "  pdhg %8lld  pobj % .8e  dobj % .8e  relP %.2e  relD %.2e  gap %.2e  w %.2e  restarts %d\n",
                     it, rec[0].pobj, rec[0].dobj, rec[0].relP, rec[0].relD, rec[0].gap, rec[0].omega, restarts[0]);
      if (remaining == 0) break;
      if (o.time_limit > 0 && since_s(tSetup) >= o.time_limit) {
        expired = true;
        break;
      }
    }
    kernelSeconds += run->stop();
    if (o.kernel_timing_reps > 0 && first == 0) {
      // After the loop: every instance is released so each launch does the
      // full work of an iteration (the results were already reached).
      run->time_kernels(o.kernel_timing_reps, &acc.col_kernel_ms, &acc.row_kernel_ms);
    }
    maxItersRun = std::max(maxItersRun, it);

    // ---- extraction (src/pdhg/pdhg.cpp:571-587, 668-680) ----
    const auto tX = Clock::now();
    std::vector<double> xs(nB), ys(mB);
    run->download(0, xs.data());
    run->download(1, ys.data());
    transfer += since_s(tX);
    for (int b = 0; b < kc; ++b) {
      pdhg_result& r = results[first + b];
      const CheckRecord& q = rec[b];
      if (done[b]) {
        r.status = PDHG_OPTIMAL;
        r.iterations = doneIter[b];
      } else {
        r.status = expired ? PDHG_TIME_LIMIT : PDHG_ITERATION_LIMIT;
        r.iterations = expired ? it - 1 : it;
      }
      r.restarts = restarts[b];
      r.objective = q.pobj;
      r.dual_objective = q.dobj;
      r.rel_primal = q.relP;
      r.rel_dual = q.relD;
      r.rel_gap = q.gap;
      r.relative_kkt = std::max(q.relP, std::max(q.relD, q.gap));
      r.primal_weight = q.omega;
      if (r.x)
        for (int j = 0; j < n; ++j) r.x[j] = xs[static_cast<size_t>(j) * B + b] * P.C[j] * P.beta;
      if (r.row_dual)
        for (int i = 0; i < m; ++i) r.row_dual[i] = ys[static_cast<size_t>(i) * B + b] * P.R[i] * P.gamma;
      acc.restarts += restarts[b];
    }
    solveSeconds += since_s(tSolve);
    acc.kernel_launches += run->launches;
    acc.graph_launches += run->graphLaunches;
    acc.bytes_uploaded += run->bytesUp;
    acc.bytes_downloaded += run->bytesDown;
    double colB = 0, rowB = 0;
    run->kernel_bytes(&colB, &rowB);
    acc.iter_bytes = colB + rowB;
    acc.col_kernel_bytes = colB;
    acc.row_kernel_bytes = rowB;
    const double vs = o.precision == PDHG_FP64 ? 8.0 : 4.0;
    acc.spmv_bytes_per_iter = 2 * ((vs + 4.0) * P.A.nnz()) + 4.0 * (m + 1) + 4.0 * (n + 1) +
                              2 * vs * static_cast<double>(B) * (m + n);
    if (k == 1) {
      acc.primal_residual = rec[0].relP;
      acc.dual_residual = rec[0].relD;
      acc.gap = rec[0].gap;
      acc.primal_weight = rec[0].omega;
    }
    acc.batch = B;
  }
  if (stats) {
    acc.iterations = maxItersRun;
    acc.setup_seconds = setupSeconds;
    acc.solve_seconds = solveSeconds;
    acc.kernel_seconds = kernelSeconds;
    acc.transfer_seconds = transfer;
    acc.spectral_norm = P.normA;
    std::snprintf(acc.backend, sizeof(acc.backend), "hip-gfx950");
    std::snprintf(acc.device_name, sizeof(acc.device_name), "%s", di.name.c_str());
    *stats = acc;
  }
  (void)chunks;
}

}  // namespace pdhgamd
