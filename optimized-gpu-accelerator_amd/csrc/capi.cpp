// C-ABI of the engine, certifier and generators (declared in include/pdhg_hip.h).
// Exceptions never cross this boundary: they become -1 / NULL plus a
// thread-local message, the convention of the reference C API
// (src/api/c_api.cpp:19,275-278).
#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pdhg_hip.h"
#include "certify.h"
#include "device.h"
#include "engine.h"
#include "prepare.h"

namespace pdhgamd {
Model gen_random_lp(uint64_t seed, int m, int n, double perCol, double rowSigma);
Model gen_refinery(uint64_t seed, int T, int S, int K, int P, int Q, double decades);
void gen_family(const Model& base, uint64_t seed, int k, double range, double* thetas, double* costs);
void family_costs(const Model& base, int k, const double* thetas, double* costs);
Model gen_calibration(uint64_t seed, int s);
Model read_mps(const std::string& path, int* sense);
}  // namespace pdhgamd

struct pdhg_model {
  pdhgamd::Model m;
};

namespace {

thread_local std::string gLastError;

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    gLastError = e.what();
    return -1;
  }
}

template <class F>
pdhg_model* guard_model(F&& f) {
  try {
    auto* h = new pdhg_model;
    h->m = f();
    return h;
  } catch (const std::exception& e) {
    gLastError = e.what();
    return nullptr;
  }
}

}  // namespace

extern "C" {

void pdhg_set_last_error(const char* msg) { gLastError = msg ? msg : ""; }
const char* pdhg_last_error(void) { return gLastError.c_str(); }

void pdhg_default_options(pdhg_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->tolerance = 1e-6;
  o->max_iterations = 1000000;
  o->check_every = 64;
  o->use_gpu = 1;
  o->ruiz_iterations = 10;
  o->pock_chambolle = 1;
  o->rescale = 1;
  o->reflection = 1.0;
  o->log_interval = 0;
  o->precision = PDHG_FP64;  // the reference's arithmetic; FP32 iterates are opt-in
  o->use_graph = 1;
  o->lanes_per_row = 0;
  o->row_map = 0;
  o->xcd_remap = -1;
  o->time_limit = 0;
  o->device = 0;
  o->kernel_timing_reps = 0;
  o->compact = 1;
  o->progress_seconds = 0;
  o->warm_fallback = -1;
}

int pdhg_gpu_available(void) { return pdhgamd::device_info(0).available ? 1 : 0; }

int pdhg_gpu_info_json(char* buf, size_t cap) {
  return guard([&] {
    const auto& d = pdhgamd::device_info(0);
    char tmp[1024];
    std::snprintf(tmp, sizeof(tmp),
                  "{\"available\": %s, \"name\": \"%s\", \"arch\": \"%s\", \"reason\": \"%s\", "
                  "\"compute_units\": %d, \"memory_bytes\": %.0f, \"peak_bandwidth_GBs\": %.1f}",
                  d.available ? "true" : "false", d.name.c_str(), d.arch.c_str(), d.reason.c_str(), d.computeUnits,
                  static_cast<double>(d.memoryBytes), d.peakBandwidthGBs);
    PDHG_REQUIRE(buf && cap > std::strlen(tmp), "buffer too small");
    std::memcpy(buf, tmp, std::strlen(tmp) + 1);
  });
}

int pdhg_solve(const pdhg_lp* lp, const pdhg_options* opts, pdhg_result* result, pdhg_stats* stats) {
  return guard([&] {
    PDHG_REQUIRE(lp && result, "null argument");
    pdhg_options o;
    if (opts) o = *opts;
    else pdhg_default_options(&o);
    pdhgamd::Model md = pdhgamd::model_from(lp);
    std::vector<pdhgamd::CaseData> cases(1);
    pdhgamd::solve_batch(md, cases, o, result, stats);
  });
}

int pdhg_solve_batch(const pdhg_lp* base, const pdhg_case* cases, int k, const pdhg_options* opts,
                     pdhg_result* results, pdhg_stats* batch_stats) {
  return guard([&] {
    PDHG_REQUIRE(base && results && k >= 0, "null argument");
    if (k == 0) return;  // solvePdhgBatch returns {} for no cases (src/pdhg/pdhg.cpp:724)
    pdhg_options o;
    if (opts) o = *opts;
    else pdhg_default_options(&o);
    pdhgamd::Model md = pdhgamd::model_from(base);
    std::vector<pdhgamd::CaseData> cs(k);
    if (cases)
      for (int b = 0; b < k; ++b) {
        cs[b].c = cases[b].cost;
        cs[b].l = cases[b].col_lower;
        cs[b].u = cases[b].col_upper;
        cs[b].L = cases[b].row_lower;
        cs[b].U = cases[b].row_upper;
        cs[b].x0 = cases[b].x0;
        cs[b].y0 = cases[b].y0;
        PDHG_REQUIRE(std::isfinite(cases[b].primal_weight) && cases[b].primal_weight >= 0,
                     "primal_weight must be finite and >= 0");
        cs[b].omega0 = cases[b].primal_weight;
      }
    pdhgamd::solve_batch(md, cs, o, results, batch_stats);
  });
}

int pdhg_solve_packed(const pdhg_lp* lps, int count, const pdhg_options* opts, pdhg_result* results,
                      pdhg_stats* stats) {
  return guard([&] {
    PDHG_REQUIRE(lps && results && count >= 1, "null argument");
    pdhg_options o;
    if (opts) o = *opts;
    else pdhg_default_options(&o);
    std::vector<pdhgamd::Model> models;
    models.reserve(count);
    for (int b = 0; b < count; ++b) models.push_back(pdhgamd::model_from(&lps[b]));
    pdhgamd::solve_packed(models, o, results, stats);
  });
}

int pdhg_prepare(const pdhg_lp* lp, int ruiz_iterations, int pock_chambolle, int rescale, double* R, double* C,
                 double* scalars, int* a_ptr, int* a_idx, double* a_val) {
  return guard([&] {
    PDHG_REQUIRE(lp, "null argument");
    const pdhgamd::Model md = pdhgamd::model_from(lp);
    const pdhgamd::Prepared P = pdhgamd::prepare(md, ruiz_iterations, pock_chambolle != 0, rescale != 0);
    if (R) std::copy(P.R.begin(), P.R.end(), R);
    if (C) std::copy(P.C.begin(), P.C.end(), C);
    if (scalars) {
      scalars[0] = P.beta;
      scalars[1] = P.gamma;
      scalars[2] = P.normA;
    }
    if (a_ptr) std::copy(P.A.ptr.begin(), P.A.ptr.end(), a_ptr);
    if (a_idx) std::copy(P.A.idx.begin(), P.A.idx.end(), a_idx);
    if (a_val) std::copy(P.A.val.begin(), P.A.val.end(), a_val);
  });
}

// ---- certifier -----------------------------------------------------------
int pdhg_safe_dual_bound(const pdhg_lp* lp, const double* y, double* bound, int* infinite_terms,
                         double* cost_perturbation) {
  return guard([&] {
    PDHG_REQUIRE(lp && y && bound, "null argument");
    pdhgamd::Model md = pdhgamd::model_from(lp);
    pdhgamd::SafeBound sb =
        pdhgamd::safe_dual_bound(md.A, md.c.data(), md.l.data(), md.u.data(), md.L.data(), md.U.data(), y);
    *bound = sb.bound;
    if (infinite_terms) *infinite_terms = sb.infiniteTerms;
    if (cost_perturbation) *cost_perturbation = sb.costPerturbation;
  });
}

int pdhg_certify(const pdhg_lp* lp, const double* x, const double* y, double objective, double* out) {
  return guard([&] {
    PDHG_REQUIRE(lp && y && out, "null argument");
    pdhgamd::Model md = pdhgamd::model_from(lp);
    pdhgamd::FamilyCert fc = pdhgamd::certify_case(md.A, md.c.data(), md.l.data(), md.u.data(), md.L.data(),
                                                   md.U.data(), md.offset, x, y, objective);
    out[0] = fc.bound;
    out[1] = fc.relGap;
    out[2] = fc.primalViolation;
    out[3] = fc.accepted ? 1.0 : 0.0;
  });
}

// Certifies k cases of a family on `threads` host threads.  Implied bounds
// depend only on A and the bounds, so cases sharing them share one
// propagation.  out[4k] as pdhg_certify.
int pdhg_certify_batch(const pdhg_lp* base, const pdhg_case* cases, int k, const double* xs, const double* ys,
                       const double* objectives, int threads, double* out) {
  return guard([&] {
    PDHG_REQUIRE(base && ys && objectives && out && k >= 0, "null argument");
    const pdhgamd::Model md = pdhgamd::model_from(base);
    const int n = md.n, m = md.m;
    auto pick = [&](int b, const double* pdhg_case::*f, const std::vector<double>& dflt) {
      return (cases && cases[b].*f) ? cases[b].*f : dflt.data();
    };
    bool sharedBounds = true;
    for (int b = 1; b < k && cases; ++b)
      sharedBounds = sharedBounds && cases[b].col_lower == cases[0].col_lower &&
                     cases[b].col_upper == cases[0].col_upper && cases[b].row_lower == cases[0].row_lower &&
                     cases[b].row_upper == cases[0].row_upper;
    std::vector<double> lo0, up0;
    if (sharedBounds && k > 0) {
      lo0.assign(pick(0, &pdhg_case::col_lower, md.l), pick(0, &pdhg_case::col_lower, md.l) + n);
      up0.assign(pick(0, &pdhg_case::col_upper, md.u), pick(0, &pdhg_case::col_upper, md.u) + n);
      pdhgamd::implied_bounds(md.A, pick(0, &pdhg_case::row_lower, md.L), pick(0, &pdhg_case::row_upper, md.U), lo0,
                              up0);
    }
    // At most one thread per case and per hardware thread of the host.
    const int hw = std::max(1, static_cast<int>(std::thread::hardware_concurrency()));
    const int nt = std::max(1, std::min({threads, k, hw}));
    std::vector<std::string> errs(nt);
    auto work = [&](int tid) {
      try {
        for (int b = tid; b < k; b += nt) {
          const double* c = pick(b, &pdhg_case::cost, md.c);
          const double* l = pick(b, &pdhg_case::col_lower, md.l);
          const double* u = pick(b, &pdhg_case::col_upper, md.u);
          const double* L = pick(b, &pdhg_case::row_lower, md.L);
          const double* U = pick(b, &pdhg_case::row_upper, md.U);
          std::vector<double> lo, up;
          if (sharedBounds) {
            lo = lo0;
            up = up0;
          } else {
            lo.assign(l, l + n);
            up.assign(u, u + n);
            pdhgamd::implied_bounds(md.A, L, U, lo, up);
          }
          const double* y = ys + static_cast<size_t>(b) * m;
          pdhgamd::SafeBound sb = pdhgamd::safe_dual_bound(md.A, c, lo.data(), up.data(), L, U, y);
          const double obj = objectives[b];
          double* o = out + 4 * static_cast<size_t>(b);
          o[0] = sb.bound + md.offset;
          o[1] = std::isfinite(sb.bound) ? std::fabs(obj - o[0]) / (1 + std::fabs(obj)) : pdhgamd::kInf;
          o[2] = xs ? pdhgamd::max_primal_violation(md.A, l, u, L, U, xs + static_cast<size_t>(b) * n) : 0.0;
          o[3] = (std::isfinite(sb.bound) && o[1] <= 1e-4) ? 1.0 : 0.0;
        }
      } catch (const std::exception& e) {
        errs[tid] = e.what();
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& t : pool) t.join();
    for (auto& e : errs) PDHG_REQUIRE(e.empty(), e);
  });
}

// ---- generators ------------------------------------------------------------
pdhg_model* pdhg_gen_random_lp(uint64_t seed, int m, int n, double nnz_per_col, double row_sigma) {
  return guard_model([&] { return pdhgamd::gen_random_lp(seed, m, n, nnz_per_col, row_sigma); });
}

pdhg_model* pdhg_gen_refinery(uint64_t seed, int periods, int streams, int crudes, int products, int qualities,
                              double scale_decades) {
  return guard_model(
      [&] { return pdhgamd::gen_refinery(seed, periods, streams, crudes, products, qualities, scale_decades); });
}

pdhg_model* pdhg_gen_calibration(uint64_t seed, int s) {
  return guard_model([&] { return pdhgamd::gen_calibration(seed, s); });
}

int pdhg_gen_family(const pdhg_model* base, uint64_t seed, int k, double range, double* thetas, double* costs) {
  return guard([&] {
    PDHG_REQUIRE(base && thetas && costs && k >= 0, "null argument");
    pdhgamd::gen_family(base->m, seed, k, range, thetas, costs);
  });
}

int pdhg_family_costs(const pdhg_model* base, int k, const double* thetas, double* costs) {
  return guard([&] {
    PDHG_REQUIRE(base && thetas && costs && k >= 0, "null argument");
    pdhgamd::family_costs(base->m, k, thetas, costs);
  });
}

int pdhg_model_view(const pdhg_model* mdl, pdhg_lp* lp, const double** x_star, const double** y_star) {
  return guard([&] {
    PDHG_REQUIRE(mdl && lp, "null argument");
    const pdhgamd::Model& m = mdl->m;
    lp->m = m.m;
    lp->n = m.n;
    lp->col_start = m.A.start.data();
    lp->row_index = m.A.index.data();
    lp->value = m.A.value.data();
    lp->cost = m.c.data();
    lp->col_lower = m.l.data();
    lp->col_upper = m.u.data();
    lp->row_lower = m.L.data();
    lp->row_upper = m.U.data();
    lp->obj_offset = m.offset;
    if (x_star) *x_star = m.xstar.empty() ? nullptr : m.xstar.data();
    if (y_star) *y_star = m.ystar.empty() ? nullptr : m.ystar.data();
  });
}

pdhg_model* pdhg_read_mps(const char* path, int* sense) {
  return guard_model([&] {
    PDHG_REQUIRE(path, "null path");
    return pdhgamd::read_mps(path, sense);
  });
}

void pdhg_model_free(pdhg_model* mdl) { delete mdl; }

}  // extern "C"
