#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "device.h"
#include "tiles.h"
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
  PDHG_REQUIRE(lp->col_start, "null column pointers");
  PDHG_REQUIRE(lp->n == 0 || (lp->cost && lp->col_lower && lp->col_upper), "null column data");
  PDHG_REQUIRE(lp->m == 0 || (lp->row_lower && lp->row_upper), "null row bounds");
  Model md;
  md.m = lp->m;
  md.n = lp->n;
  md.A.rows = lp->m;
  md.A.cols = lp->n;
  md.A.start.assign(lp->col_start, lp->col_start + lp->n + 1);
  const int nnz = md.A.start.back();
  PDHG_REQUIRE(md.A.start[0] == 0 && nnz >= 0, "bad column pointers");
  PDHG_REQUIRE(nnz == 0 || (lp->row_index && lp->value), "null matrix arrays");
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
  for (int j = 0; j < lp->n; ++j)
    PDHG_REQUIRE(!std::isnan(md.c[j]) && !std::isnan(md.l[j]) && !std::isnan(md.u[j]), "NaN in column data");
  md.L.assign(lp->row_lower, lp->row_lower + lp->m);
  md.U.assign(lp->row_upper, lp->row_upper + lp->m);
  md.offset = lp->obj_offset;
  return md;
}

namespace {

// Scaled, interleaved case data of a set of cases occupying lanes 0..B-1
// (src/pdhg/pdhg.cpp:506-545).  Lanes past the cases are padding, marked done.
struct LaneData {
  std::vector<double> c, lo, up, rl, ru, x, y, omega, cNorm, rhsNorm;
  std::vector<int> pad;
};

struct CaseScaled {  // per case: norms of the original vectors and initial omega
  double cNorm = 0, rhsNorm = 0, omega = 1;
  double omegaCold = 1;  // the reference's initial primal weight (the cold start's)
};

// ---- autotuner bookkeeping ------------------------------------------------
struct TuneKey {
  uint64_t shape;
  int B, precision, device;
  int maps;  // 0: autotuned mapping; else the options' explicit mapping
  bool operator<(const TuneKey& r) const {
    return std::tie(shape, B, precision, device, maps) < std::tie(r.shape, r.B, r.precision, r.device, r.maps);
  }
};

// A width's row mapping (column and row kernel) and the measured time of one
// iteration (column + row kernel, ms) under it.
struct Tuned {
  MapSetup col, row;
  double ms = 0;    // one iteration (column + row kernel), every lane live
  double half = 0;  // the same with only the first half of the lanes live
};

std::mutex gTuneMutex;
std::map<TuneKey, Tuned> gTuned;

bool tuned(const TuneKey& k, Tuned* t) {
  std::lock_guard<std::mutex> lk(gTuneMutex);
  auto it = gTuned.find(k);
  if (it == gTuned.end()) return false;
  *t = it->second;
  return true;
}

void remember(const TuneKey& k, const Tuned& t) {
  std::lock_guard<std::mutex> lk(gTuneMutex);
  gTuned[k] = t;
}

// FNV-1a over the sparsity structure of the scaled A (shape, row pointers,
// column indices): runs of the same matrix share their tuned mapping.
uint64_t structure_hash(const Prepared& P) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    for (int i = 0; i < 8; ++i) {
      h ^= (v >> (8 * i)) & 0xff;
      h *= 1099511628211ull;
    }
  };
  mix(static_cast<uint64_t>(P.m));
  mix(static_cast<uint64_t>(P.n));
  for (int v : P.A.ptr) mix(static_cast<uint32_t>(v));
  for (int v : P.A.idx) mix(static_cast<uint32_t>(v));
  return h;
}

// Row mappings the autotuner times.  At B = 64 a wavefront serves one row
// (lane = instance); below, the lanes per row (nnz-per-lane target, which
// sets the row tile S*B) and bucketing are open, and at B = 16 and 32 the
// grouped form (one row per B-lane group).  From B = 16 on the XCD remap is
// always on: without it every XCD's L2 fetches the gathered columns of rows
// spread over the whole matrix (on the refinery matrix at B = 64, 2.8x the
// distinct gathered bytes for the LDS tiles and 6x for one row per
// wavefront, by FETCH_SIZE), a cost that timings inside the Infinity Cache
// do not show but HBM-resident runs pay.
std::vector<MapSetup> map_candidates(int B) {
  std::vector<MapSetup> c;
  for (bool xcd : {false, true}) {
    if (!xcd && B >= 16) continue;
    if (B == 64) {
      for (int rpw : {1, 2}) c.push_back(MapSetup{4, 1, xcd, false, rpw});
      for (int shape = 1; shape <= kTileShapes; ++shape)
        c.push_back(MapSetup{4, 1, xcd, false, 1, shape});  // rows in LDS-staged tiles
      continue;
    }
    for (int per : {2, 4, 8, 16})
      for (int map : {1, 2}) c.push_back(MapSetup{per, map, xcd});
    if (B == 16 || B == 32)  // one row per B-lane group, rows by length or in order
      for (int map : {1, 2}) c.push_back(MapSetup{4, map, xcd, true});
  }
  return c;
}

}  // namespace

void check_options(const pdhg_options& o) {
  PDHG_REQUIRE(o.use_gpu, "use_gpu = 0: this engine has no CPU backend");
  PDHG_REQUIRE(o.check_every >= 2, "check_every must be >= 2");
  PDHG_REQUIRE(o.max_iterations >= 1, "max_iterations must be >= 1");
  PDHG_REQUIRE(!std::isnan(o.tolerance), "tolerance is NaN");
  PDHG_REQUIRE(!std::isnan(o.time_limit), "time_limit is NaN");
  PDHG_REQUIRE(o.precision == PDHG_FP32 || o.precision == PDHG_FP64 || o.precision == PDHG_FP32_AUTO,
               "unknown precision");
  PDHG_REQUIRE(o.lanes_per_row >= 0 && o.lanes_per_row <= 64, "lanes_per_row must be in 0..64");
  PDHG_REQUIRE(o.row_map >= 0 && o.row_map <= 2 + kTileShapes, "row_map must be in 0..5");
  PDHG_REQUIRE(o.xcd_remap >= -1 && o.xcd_remap <= 1, "xcd_remap must be -1, 0 or 1");
  PDHG_REQUIRE(o.device >= 0 && o.device < kMaxDevices, "device ordinal out of range");
  PDHG_REQUIRE(o.warm_fallback >= -1, "warm_fallback must be >= -1");
}

void solve_batch(const Model& base, const std::vector<CaseData>& cases, const pdhg_options& o, pdhg_result* results,
                 pdhg_stats* stats) {
  const auto tSetup = Clock::now();
  const int k = static_cast<int>(cases.size());
  PDHG_REQUIRE(k >= 1, "no cases");
  check_options(o);
  const DeviceInfo& di = device_info(o.device);
  if (!di.available) throw Error("GPU unavailable: " + di.reason);
  const int n = base.n, m = base.m;

  // Set-up shared by every case (src/pdhg/pdhg.cpp:487; beta, gamma from the base model).
  const Prepared P = prepare(base, o.ruiz_iterations, o.pock_chambolle != 0, o.rescale != 0);
  const double eta = 0.998 / P.normA;

  const bool shC = same_pointer(cases, &CaseData::c);
  const bool shB = same_pointer(cases, &CaseData::l) && same_pointer(cases, &CaseData::u);
  const bool shR = same_pointer(cases, &CaseData::L) && same_pointer(cases, &CaseData::U);
  auto cost = [&](int q) { return cases[q].c ? cases[q].c : base.c.data(); };
  auto colL = [&](int q) { return cases[q].l ? cases[q].l : base.l.data(); };
  auto colU = [&](int q) { return cases[q].u ? cases[q].u : base.u.data(); };
  auto rowL = [&](int q) { return cases[q].L ? cases[q].L : base.L.data(); };
  auto rowU = [&](int q) { return cases[q].U ? cases[q].U : base.U.data(); };

  // Norms of the original case vectors (src/pdhg/pdhg.cpp:466-481) and the
  // initial primal weight ||c_hat|| / ||b_hat|| (:513-529).
  std::vector<CaseScaled> cs(k);
  for (int q = 0; q < k; ++q) {
    const double *cc = cost(q), *cL = rowL(q), *cU = rowU(q);
    double cn0 = 0, bn0 = 0, cn = 0, bn = 0;
    for (int j = 0; j < n; ++j) {
      cn0 += cc[j] * cc[j];
      const double cj = cc[j] * P.C[j] / P.gamma;
      cn += cj * cj;
    }
    for (int i = 0; i < m; ++i) {
      if (finite_bound(cL[i])) bn0 += cL[i] * cL[i];
      if (finite_bound(cU[i]) && cU[i] != cL[i]) bn0 += cU[i] * cU[i];
      const double li = finite_bound(cL[i]) ? cL[i] * P.R[i] / P.beta : -HUGE_VAL;
      const double ui = finite_bound(cU[i]) ? cU[i] * P.R[i] / P.beta : HUGE_VAL;
      if (finite_bound(li)) bn += li * li;
      if (finite_bound(ui) && ui != li) bn += ui * ui;
    }
    cs[q].cNorm = std::sqrt(cn0);
    cs[q].rhsNorm = std::sqrt(bn0);
    if (cn > 1e-20 && bn > 1e-20) cs[q].omega = std::sqrt(cn) / std::sqrt(bn);
    cs[q].omegaCold = cs[q].omega;
    if (cases[q].omega0 > 0) cs[q].omega = cases[q].omega0;
  }
  // Cold-start fallback of warm-started cases: the iteration (of the case's
  // own count) at which an unconverged one restarts from the cold start.
  auto is_warm = [&](int q) { return cases[q].x0 || cases[q].y0 || cases[q].omega0 > 0; };
  long long fallbackAt = 0;
  if (o.warm_fallback != 0) {
    const long long w = o.warm_fallback > 0 ? o.warm_fallback : o.max_iterations / 4;
    fallbackAt = std::max<long long>(o.check_every, (w + o.check_every - 1) / o.check_every * o.check_every);
    if (fallbackAt >= o.max_iterations) fallbackAt = 0;
  }
  bool anyFallback = false;
  for (int q = 0; q < k && fallbackAt > 0; ++q) anyFallback = anyFallback || is_warm(q);
  std::vector<char> fellBack(k, 0);
  std::vector<long long> warmIters(k, 0);  // iterations a case ran warm before its fallback
  std::vector<int> warmRestarts(k, 0);
  int fallbacks = 0;

  // Lane data for cases `ids` in a run of width B; `cold` ignores the cases'
  // warm starts (the fallback).
  auto lane_data = [&](const std::vector<int>& ids, int B, bool cold = false) {
    LaneData d;
    const size_t nB = static_cast<size_t>(n) * B, mB = static_cast<size_t>(m) * B;
    d.c.assign(shC ? n : nB, 0.0);
    d.lo.assign(shB ? n : nB, 0.0);
    d.up.assign(shB ? n : nB, 0.0);
    d.rl.assign(shR ? m : mB, 0.0);
    d.ru.assign(shR ? m : mB, 0.0);
    d.x.assign(nB, 0.0);
    d.y.assign(mB, 0.0);
    d.omega.assign(B, 1.0);
    d.cNorm.assign(B, 0.0);
    d.rhsNorm.assign(B, 0.0);
    d.pad.assign(B, 1);
    for (int b = 0; b < B; ++b) {
      const bool real = b < static_cast<int>(ids.size());
      const int q = ids[real ? b : 0];
      d.pad[b] = real ? 0 : 1;
      d.omega[b] = cold ? cs[q].omegaCold : cs[q].omega;
      d.cNorm[b] = cs[q].cNorm;
      d.rhsNorm[b] = cs[q].rhsNorm;
      const double *cc = cost(q), *cl = colL(q), *cu = colU(q), *cL = rowL(q), *cU = rowU(q);
      for (int j = 0; j < n; ++j) {
        const double cj = cc[j] * P.C[j] / P.gamma;
        const double lj = finite_bound(cl[j]) ? cl[j] / P.C[j] / P.beta : -HUGE_VAL;
        const double uj = finite_bound(cu[j]) ? cu[j] / P.C[j] / P.beta : HUGE_VAL;
        if (!shC) d.c[static_cast<size_t>(j) * B + b] = cj;
        else if (b == 0) d.c[j] = cj;
        if (!shB) {
          d.lo[static_cast<size_t>(j) * B + b] = lj;
          d.up[static_cast<size_t>(j) * B + b] = uj;
        } else if (b == 0) {
          d.lo[j] = lj;
          d.up[j] = uj;
        }
        // Cold start x0 = clamp(0) (src/pdhg/pdhg.cpp:519), or the warm start
        // mapped into scaled space (inverse of the unscaling at :581), clamped.
        const double x0 = cases[q].x0 && !cold ? cases[q].x0[j] / P.C[j] / P.beta : 0.0;
        d.x[static_cast<size_t>(j) * B + b] = std::min(std::max(x0, lj), uj);
      }
      for (int i = 0; i < m; ++i) {
        const double li = finite_bound(cL[i]) ? cL[i] * P.R[i] / P.beta : -HUGE_VAL;
        const double ui = finite_bound(cU[i]) ? cU[i] * P.R[i] / P.beta : HUGE_VAL;
        if (!shR) {
          d.rl[static_cast<size_t>(i) * B + b] = li;
          d.ru[static_cast<size_t>(i) * B + b] = ui;
        } else if (b == 0) {
          d.rl[i] = li;
          d.ru[i] = ui;
        }
        if (cases[q].y0 && !cold) d.y[static_cast<size_t>(i) * B + b] = cases[q].y0[i] / P.R[i] / P.gamma;
      }
    }
    return d;
  };

  RunScalars sc;
  sc.eta = eta;
  sc.beta = P.beta;
  sc.gamma = P.gamma;
  sc.offset = P.offset;
  sc.tol = o.tolerance;
  sc.maxIter = o.max_iterations;
  int runPrec = o.precision == PDHG_FP64 ? PDHG_FP64 : PDHG_FP32;
  std::vector<double> refKkt(k, kInf);
  std::vector<int> lastGain(k, 0);
  // Row mapping of the iteration kernels: the options' explicit choice, or
  // (all three on auto) the autotuner's.  Either way the first run of a
  // width times its mapping on the device (scratch state, before the data is
  // uploaded), and mapping and time are cached per matrix structure, width,
  // precision and device for the life of the process: the compaction below
  // uses the times to pick widths.
  const bool autoMap = o.lanes_per_row == 0 && o.row_map == 0 && o.xcd_remap < 0;
  const uint64_t shape = structure_hash(P);
  const int mapsKey = autoMap ? 0 : 1 + o.lanes_per_row * 100 + o.row_map * 10 + (o.xcd_remap + 1);
  bool mapsRecorded = false, tunedFirst = false;
  MapSetup firstCol, firstRow;
  double tuneSeconds = 0;
  auto fixed_map = [&](int B) {
    MapSetup ms;
    ms.nnzPerLane = o.lanes_per_row > 0 ? o.lanes_per_row : 4;
    ms.rowMap = o.row_map == 2 ? 2 : 1;
    ms.tile = B == 64 && o.row_map >= 3 ? o.row_map - 2 : 0;  // 3, 4, 5: tile shapes 1, 2, 3
    ms.xcd = o.xcd_remap == 1 || (o.xcd_remap < 0 && B >= 16);
    return ms;
  };
  // A run of width B (case data not yet uploaded) with its tuned mapping.
  auto create_run = [&](int B, int prec) {
    RunSetup rs;
    rs.m = m;
    rs.n = n;
    rs.B = B;
    rs.precision = prec == PDHG_FP64 ? PDHG_FP64 : PDHG_FP32;
    rs.sharedCost = shC;
    rs.sharedColBounds = shB;
    rs.sharedRowBounds = shR;
    rs.rho = o.reflection;
    rs.checkEvery = o.check_every;
    rs.useGraph = o.use_graph != 0;
    const TuneKey key{shape, B, rs.precision, o.device, mapsKey};
    Tuned t;
    const bool known = tuned(key, &t);
    rs.colMap = known ? t.col : fixed_map(B);
    rs.rowMap = known ? t.row : fixed_map(B);
    auto run = DeviceRun::create(o.device, P, rs);
    if (!known) {
      const auto t0 = Clock::now();
      const std::vector<MapSetup> cand = autoMap ? map_candidates(B) : std::vector<MapSetup>{fixed_map(B)};
      int best[2];
      run->autotune(P, cand, 16, best, &t.ms, &t.half);
      t.col = cand[best[0]];
      t.row = cand[best[1]];
      remember(key, t);
      rs.colMap = t.col;
      rs.rowMap = t.row;
      tuneSeconds += since_s(t0);
    }
    return std::make_pair(std::move(run), rs);
  };
  // Time of one iteration at width W (probing the width once); packed: the
  // live cases sit in the first half of the lanes.
  auto width_cost = [&](int W, int prec, bool packed = false) {
    const TuneKey key{shape, W, prec == PDHG_FP64 ? PDHG_FP64 : PDHG_FP32, o.device, mapsKey};
    Tuned t;
    if (!tuned(key, &t)) {
      create_run(W, prec);
      tuned(key, &t);
    }
    return packed ? std::min(t.ms, t.half) : t.ms;
  };
  auto make_run = [&](const std::vector<int>& ids, int B) {
    auto [run, rs] = create_run(B, runPrec);
    if (!mapsRecorded) {
      mapsRecorded = true;
      tunedFirst = autoMap;
      firstCol = rs.colMap;
      firstRow = rs.rowMap;
      // Tiles need rows short enough for the staging LDS; else the plain plan.
      firstCol.tile = run->tiled(0);
      firstRow.tile = run->tiled(1);
    }
    LaneData d = lane_data(ids, B);
    run->upload(d.c.data(), d.lo.data(), d.up.data(), d.rl.data(), d.ru.data(), d.x.data(), d.y.data());
    run->init_state(sc, d.omega.data(), d.cNorm.data(), d.rhsNorm.data(), d.pad.data());
    return run;
  };

  pdhg_stats acc{};
  if (stats) {
    acc.trace = stats->trace;
    acc.trace_capacity = stats->trace_capacity;
  }
  acc.precision = o.precision;
  acc.fp64_switch_iteration = -1;
  long long maxItersRun = 0;
  double solveSeconds = 0, kernelSeconds = 0, transfer = 0, setupSeconds = 0;
  bool timed = false;

  // More cases than one run holds: continuous batching.  A run of width 64
  // starts with the first 64 cases; a lane whose case finishes (or reaches
  // its own iteration limit) is given the next pending case while the other
  // lanes keep running, so the run's time follows the cases' mean iteration
  // count rather than the slowest case of every group of 64.
  // Lanes run on their own iteration counts (continuous batching) when more
  // cases than one run holds are queued, and when warm-started cases may
  // fall back to the cold start.
  const bool stream = k > 64 || anyFallback;
  for (int first = 0; first < k; first += stream ? k : 64) {
    const auto tChunk = Clock::now();
    std::vector<int> ids;  // lane -> case
    for (int q = first; q < std::min(k, first + 64); ++q) ids.push_back(q);
    std::vector<int> pending;  // cases not yet started (streaming)
    if (k > 64)
      for (int q = k - 1; q >= 64; --q) pending.push_back(q);
    int B = next_pow2(static_cast<int>(ids.size()));
    // Lanes 0..span-1 of the run may hold live cases: span < B after the
    // live cases were re-packed into the first lanes of a run as wide.
    int span = B;
    auto run = make_run(ids, B);
    transfer += since_s(tChunk);
    if (first == 0) setupSeconds = since_s(tSetup);
    if (first == 0) {
      double colB = 0, rowB = 0;
      run->kernel_bytes(&colB, &rowB);
      acc.iter_bytes = colB + rowB;
      acc.col_kernel_bytes = colB;
      acc.row_kernel_bytes = rowB;
      const double vs = o.precision == PDHG_FP64 ? 8.0 : 4.0;
      acc.spmv_bytes_per_iter = 2 * ((vs + 4.0) * P.A.nnz()) + 4.0 * (m + 1) + 4.0 * (n + 1) +
                                2 * vs * static_cast<double>(B) * (m + n);
      acc.batch = B;
    }

    // ---- iterations (src/pdhg/pdhg.cpp:588-662), one check per block ----
    const auto tSolve = Clock::now();
    run->start(true);
    std::vector<CheckRecord> rec(B);
    std::vector<int> done(B), restarts(B);
    std::vector<long long> doneIter(B);
    std::vector<char> extracted(k, 0);
    std::vector<long long> loadIt(B, 0);  // run iteration at which each lane's case started
    std::vector<char> limitHit(B, 0);      // lane reached its own iteration limit
    long long it = 0;
    int checks = 0;
    bool expired = false;
    auto lastProgress = Clock::now();

    // Writes the results of lanes `which` of the current run (src/pdhg/pdhg.cpp:571-587).
    auto extract = [&](const std::vector<int>& which, bool final) {
      const auto tX = Clock::now();
      std::vector<double> xs(static_cast<size_t>(n) * B), ys(static_cast<size_t>(m) * B);
      run->download(0, xs.data());
      run->download(1, ys.data());
      transfer += since_s(tX);
      for (int b : which) {
        const int q = ids[b];
        pdhg_result& r = results[q];
        const CheckRecord& cr = rec[b];
        if (done[b] && !limitHit[b]) {
          r.status = PDHG_OPTIMAL;
          r.iterations = warmIters[q] + doneIter[b];
        } else {
          PDHG_REQUIRE(final || limitHit[b], "internal: extracting a live instance");
          // A lane stopped at its own limit stays so even if the run's time
          // limit expires before it is extracted.
          const bool timedOut = expired && !limitHit[b];
          r.status = timedOut ? PDHG_TIME_LIMIT : PDHG_ITERATION_LIMIT;
          const long long own = limitHit[b] ? o.max_iterations : it - loadIt[b];
          r.iterations = warmIters[q] + (timedOut ? own - 1 : own);
        }
        r.restarts = warmRestarts[q] + restarts[b];
        r.objective = cr.pobj;
        r.dual_objective = cr.dobj;
        r.rel_primal = cr.relP;
        r.rel_dual = cr.relD;
        r.rel_gap = cr.gap;
        r.relative_kkt = std::max(cr.relP, std::max(cr.relD, cr.gap));
        r.primal_weight = cr.omega;
        if (r.x)
          for (int j = 0; j < n; ++j) r.x[j] = xs[static_cast<size_t>(j) * B + b] * P.C[j] * P.beta;
        if (r.row_dual)
          for (int i = 0; i < m; ++i) r.row_dual[i] = ys[static_cast<size_t>(i) * B + b] * P.R[i] * P.gamma;
        acc.restarts += warmRestarts[q] + restarts[b];
        extracted[q] = 1;
      }
    };

    while (stream || it < o.max_iterations) {
      // The block ends at the nearest check point of a live lane, counted on
      // the lane's own iterations: the next multiple of check_every or its
      // iteration limit (src/pdhg/pdhg.cpp:593).  Lanes of one fixed batch
      // share their count; streamed lanes start at block ends.
      long long nextCheck = o.check_every;
      for (int b = 0; b < static_cast<int>(ids.size()); ++b)
        if (!done[b]) {
          const long long own = it - loadIt[b];
          nextCheck = std::min({nextCheck, o.check_every - own % o.check_every, o.max_iterations - own});
          if (fallbackAt > 0 && is_warm(ids[b]) && !fellBack[ids[b]] && own < fallbackAt)
            nextCheck = std::min(nextCheck, fallbackAt - own);
        }
      const int nIt = static_cast<int>(std::max<long long>(1, nextCheck));
      run->run_block(nIt);
      run->fetch(rec.data(), done.data(), doneIter.data(), restarts.data());
      it += nIt;
      ++checks;
      if (stream) {
        // Lanes at their own iteration limit stop (reported as such).
        std::vector<int> capped;
        for (int b = 0; b < static_cast<int>(ids.size()); ++b)
          if (!done[b] && it - loadIt[b] >= o.max_iterations) {
            capped.push_back(b);
            limitHit[b] = 1;
            done[b] = 1;
          }
        run->retire_lanes(capped);
        // Cold-start fallback: warm-started cases still live at their
        // fallback point restart in their lanes from the reference's start.
        std::vector<int> fb, fbIds;
        for (int b = 0; b < static_cast<int>(ids.size()); ++b) {
          const int q = ids[b];
          if (!done[b] && fallbackAt > 0 && is_warm(q) && !fellBack[q] && it - loadIt[b] >= fallbackAt) {
            fb.push_back(b);
            fbIds.push_back(q);
          }
        }
        if (!fb.empty()) {
          const LaneData d = lane_data(fbIds, static_cast<int>(fb.size()), true);
          run->load_lanes(fb, nullptr, nullptr, nullptr, nullptr, nullptr, d.x.data(), d.y.data(), d.omega.data(),
                          d.cNorm.data(), d.rhsNorm.data());
          for (size_t t = 0; t < fb.size(); ++t) {
            const int b = fb[t], q = fbIds[t];
            fellBack[q] = 1;
            warmIters[q] = it - loadIt[b];
            warmRestarts[q] = restarts[b];
            ++fallbacks;
            lastGain[q] = checks;
            loadIt[b] = it;
            doneIter[b] = -1;
            restarts[b] = 0;
          }
        }
        // Refill: finished lanes take pending cases, a batch at a time.
        std::vector<int> freeLanes, newly;
        for (int b = 0; b < B; ++b) {
          const bool used = b < static_cast<int>(ids.size());
          if (used && !done[b]) continue;
          if (used && !extracted[ids[b]]) newly.push_back(b);
          freeLanes.push_back(b);
        }
        const int take = std::min<int>(static_cast<int>(freeLanes.size()), static_cast<int>(pending.size()));
        if (take > 0 && (take >= 8 || take == static_cast<int>(pending.size()))) {
          extract(newly, false);
          std::vector<int> lanes(freeLanes.begin(), freeLanes.begin() + take), newIds;
          for (int q = 0; q < take; ++q) {
            newIds.push_back(pending.back());
            pending.pop_back();
          }
          const LaneData d = lane_data(newIds, take);
          run->load_lanes(lanes, shC ? nullptr : d.c.data(), shB ? nullptr : d.lo.data(),
                          shB ? nullptr : d.up.data(), shR ? nullptr : d.rl.data(), shR ? nullptr : d.ru.data(),
                          d.x.data(), d.y.data(), d.omega.data(), d.cNorm.data(), d.rhsNorm.data());
          for (int q = 0; q < take; ++q) {
            const int b = lanes[q];
            if (b >= static_cast<int>(ids.size())) ids.resize(b + 1, newIds[q]);
            ids[b] = newIds[q];
            lastGain[newIds[q]] = checks;  // FP32_AUTO stall clock starts at the load
            loadIt[b] = it;
            limitHit[b] = 0;
            done[b] = 0;
            doneIter[b] = -1;
            restarts[b] = 0;
          }
        }
      }
      const int real = static_cast<int>(ids.size());
      int remaining = 0;
      for (int b = 0; b < real; ++b) remaining += done[b] ? 0 : 1;
      // Trace of the first case of the first chunk while it is live.
      if (first == 0 && acc.trace && acc.trace_length < acc.trace_capacity) {
        for (int b = 0; b < real; ++b)
          if (ids[b] == 0 && static_cast<long long>(rec[b].it) == it - 1) {
            std::memcpy(acc.trace + 8 * acc.trace_length, &rec[b], sizeof(CheckRecord));
            acc.trace_length++;
          }
      }
      if (o.log_interval > 0 && k == 1 && checks % o.log_interval == 0)
        std::fprintf(stderr, "  pdhg %8lld  pobj % .8e  dobj % .8e  relP %.2e  relD %.2e  gap %.2e  w %.2e  restarts %d\n",
                     it, rec[0].pobj, rec[0].dobj, rec[0].relP, rec[0].relD, rec[0].gap, rec[0].omega, restarts[0]);
      if (o.progress_seconds > 0 && since_s(lastProgress) >= o.progress_seconds) {
        lastProgress = Clock::now();
        std::fprintf(stderr, "  pdhg batch: iteration %lld, %d of %d live in a run of width %d, %.1f s\n", it,
                     remaining, real, B, since_s(tSolve));
      }
      if (remaining == 0) break;
      if ((o.time_limit > 0 && since_s(tSetup) >= o.time_limit) || (o.should_stop && o.should_stop(o.stop_user))) {
        expired = true;
        break;
      }
      // FP32 resolution floor: an instance whose best relative KKT has not
      // halved over the last 256 checks is stalled; once half the live
      // instances are, the run moves to FP64 iterates.
      bool toFp64 = false;
      if (o.precision == PDHG_FP32_AUTO && runPrec == PDHG_FP32) {
        int stalled = 0;
        for (int b = 0; b < real; ++b) {
          if (done[b]) continue;
          const int q = ids[b];
          const double kkt = std::max(rec[b].relP, std::max(rec[b].relD, rec[b].gap));
          if (kkt < 0.5 * refKkt[q]) {
            refKkt[q] = kkt;
            lastGain[q] = checks;
          }
          if (checks - lastGain[q] >= 256) ++stalled;
        }
        toFp64 = stalled > 0 && 2 * stalled >= remaining;
      }
      // Re-pack: finished instances leave (compaction) and/or the run changes
      // precision; live instances keep their iterates, anchors and restart
      // state, moved through an FP64 host copy.
      // Compaction: once at most half the lanes that may be live are, move
      // the live cases into the first lanes of the run (a power of two wide,
      // holding them all) whose measured iteration is cheapest for this
      // matrix and precision.  A narrower run is not always cheaper (the row
      // mapping changes with the width), and a run as wide as this one with
      // its live cases packed into its first half can be cheaper still:
      // lanes of finished cases load nothing, so whole cache lines of the
      // vectors are skipped.
      const int newPrec = toFp64 ? PDHG_FP64 : runPrec;
      int newB = B;
      bool move = false;
      const bool mayNarrow = o.compact && B > 1 && remaining <= span / 2 && pending.empty();
      if (!autoMap) {
        // A fixed mapping keeps runs reproducible: the width follows from the
        // live count alone (the narrowest power of two that holds them), not
        // from timings.
        if (mayNarrow) newB = next_pow2(remaining);
        move = newB < B;
      } else if (mayNarrow || toFp64) {
        double best = toFp64 ? kInf : width_cost(B, runPrec, span < B);
        for (int W = next_pow2(remaining); W <= B; W *= 2) {
          if ((!o.compact || !pending.empty()) && W != B) continue;
          const double c = width_cost(W, newPrec, o.compact && pending.empty() && 2 * remaining <= W);
          if (c < best) {
            best = c;
            newB = W;
            move = true;
          }
        }
      }
      const bool compactNow = move;
      if (compactNow || toFp64) {
        const auto tMove = Clock::now();
        std::vector<int> fin, live;
        for (int b = 0; b < real; ++b) (done[b] ? fin : live).push_back(b);
        std::vector<int> newly;
        for (int b : fin)
          if (!extracted[ids[b]]) newly.push_back(b);
        extract(newly, false);
        LaneState ls;
        run->export_lanes(live, ls);
        kernelSeconds += run->stop();
        acc.kernel_launches += run->launches;
        acc.graph_launches += run->graphLaunches;
        acc.bytes_uploaded += run->bytesUp;
        acc.bytes_downloaded += run->bytesDown;
        std::vector<int> nids;
        std::vector<long long> nload;
        for (int b : live) {
          nids.push_back(ids[b]);
          nload.push_back(loadIt[b]);
        }
        ids = nids;
        B = newB;
        nload.resize(B, it);
        loadIt = nload;
        limitHit.assign(B, 0);
        if (toFp64) {
          runPrec = PDHG_FP64;
          if (acc.fp64_switch_iteration < 0) acc.fp64_switch_iteration = it;
          if (o.progress_seconds > 0)
            std::fprintf(stderr, "  pdhg batch: iteration %lld, FP32 progress stalled, continuing in FP64\n", it);
        }
        span = next_pow2(static_cast<int>(ids.size()));
        run = make_run(ids, B);
        run->import_lanes(ls);
        run->start(false);
        if (o.progress_seconds > 0)
          std::fprintf(stderr,
                       "  pdhg batch: iteration %lld, %d live moved into lanes 0..%d of a run of width %d (%.1f ms)\n",
                       it, static_cast<int>(ids.size()), static_cast<int>(ids.size()) - 1, B, 1e3 * since_s(tMove));
        rec.assign(B, CheckRecord{});
        done.assign(B, 0);
        restarts.assign(B, 0);
        doneIter.assign(B, -1);
      }
    }
    kernelSeconds += run->stop();
    maxItersRun = std::max(maxItersRun, it);
    std::vector<int> rest;
    for (int b = 0; b < static_cast<int>(ids.size()); ++b)
      if (!extracted[ids[b]]) rest.push_back(b);
    extract(rest, true);
    for (int q : pending) {  // never started (time limit)
      results[q].status = PDHG_TIME_LIMIT;
      results[q].iterations = 0;
      results[q].restarts = 0;
      results[q].relative_kkt = kInf;
    }
    if (o.kernel_timing_reps > 0 && !timed) {
      // After the results were read, every instance is released so each
      // timed launch does the full work of an iteration.
      run->time_kernels(o.kernel_timing_reps, &acc.col_kernel_ms, &acc.row_kernel_ms, &acc.iter_kernel_ms);
      double colB = 0, rowB = 0;
      run->kernel_bytes(&colB, &rowB);
      acc.col_kernel_bytes = colB;
      acc.row_kernel_bytes = rowB;
      timed = true;
    }
    solveSeconds += since_s(tSolve);
    acc.kernel_launches += run->launches;
    acc.graph_launches += run->graphLaunches;
    acc.bytes_uploaded += run->bytesUp;
    acc.bytes_downloaded += run->bytesDown;
    if (k == 1) {
      acc.primal_residual = rec[0].relP;
      acc.dual_residual = rec[0].relD;
      acc.gap = rec[0].gap;
      acc.primal_weight = rec[0].omega;
    }
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
    acc.autotuned = tunedFirst ? 1 : 0;
    acc.col_nnz_per_lane = firstCol.group ? 0 : firstCol.nnzPerLane;
    acc.col_row_map = firstCol.tile ? 2 + firstCol.tile : firstCol.rowMap;
    acc.col_xcd = firstCol.xcd ? 1 : 0;
    acc.row_nnz_per_lane = firstRow.group ? 0 : firstRow.nnzPerLane;
    acc.row_row_map = firstRow.tile ? 2 + firstRow.tile : firstRow.rowMap;
    acc.row_xcd = firstRow.xcd ? 1 : 0;
    acc.tune_seconds = tuneSeconds;
    acc.col_rows_per_wave = firstCol.rowsPerWave;
    acc.row_rows_per_wave = firstRow.rowsPerWave;
    acc.warm_fallbacks = fallbacks;
    *stats = acc;
  }
}

}  // namespace pdhgamd
