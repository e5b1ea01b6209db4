// Reference-side binding of the MI355X PDHG engine: the code a maintainer adds
// to the reference solver so that its PDHG entries (src/pdhg/pdhg.h:73-80)
// run on libpdhg_hip.so through the C-ABI include/pdhg_hip.h.
//
// It defines pramana::solvePdhg, solvePdhgBatch, gpuAvailable and gpuInfoJson.
// The reference's own src/pdhg/pdhg.cpp is compiled next to it with those four
// names prefixed by `ref` (a build-line rename, the file itself unchanged):
// its CPU backend keeps serving the requests the reference sends there, as
// runPdhg's backend choice does (src/pdhg/pdhg.cpp:491-500: CPU when useGpu is
// off or no device is present).  Every GPU request goes to the HIP engine;
// the runtime CUDA compiler path is never taken.
//
// Everything above this layer (solve(), crossover, postsolve, the certifier,
// the router, compareFamilyStrategies, calibrateBandwidth, the C API, the CLI)
// is the reference's code, unchanged.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pdhg_hip.h"
#include "pdhg/pdhg.h"
#include "util/common.h"

namespace pramana {

// The reference's entries, compiled from src/pdhg/pdhg.cpp under these names.
PdhgResult refSolvePdhg(const Model& lp, const PdhgOptions& opts, const Deadline* deadline);
std::vector<PdhgResult> refSolvePdhgBatch(const Model& base, const std::vector<LpCase>& cases,
                                          const PdhgOptions& opts, const Deadline* deadline, PdhgStats* batchStats);
bool refGpuAvailable();
Json refGpuInfoJson();

namespace {

// Model is already in minimisation form with a CSC matrix (src/core/model.h:24-64).
pdhg_lp view(const Model& lp) {
  return pdhg_lp{lp.numRows(),          lp.numCols(),         lp.A.start.data(),    lp.A.index.data(),
                 lp.A.value.data(),     lp.colCost.data(),    lp.colLower.data(),   lp.colUpper.data(),
                 lp.rowLower.data(),    lp.rowUpper.data(),   lp.objOffset};
}

// A Deadline (src/util/timer.h:22-40) is polled by the engine at each check,
// as runPdhg polls it (src/pdhg/pdhg.cpp:658): its time limit and the race
// mode's cancellation flag both end the solve with TimeLimit.
int deadline_expired(void* user) { return static_cast<const Deadline*>(user)->expired() ? 1 : 0; }

pdhg_options options(const PdhgOptions& o, const Deadline* d) {  // PdhgOptions, pdhg.h:32-42
  pdhg_options h;
  pdhg_default_options(&h);
  h.tolerance = o.tolerance;
  h.max_iterations = o.maxIterations;
  h.check_every = o.checkEvery;
  h.ruiz_iterations = o.ruizIterations;
  h.pock_chambolle = o.pockChambolle ? 1 : 0;
  h.rescale = o.rescaleObjectiveAndBounds ? 1 : 0;
  h.reflection = o.reflection;
  h.log_interval = o.logInterval;
  h.precision = PDHG_FP64;  // the reference's arithmetic
  if (d) {
    h.should_stop = deadline_expired;
    h.stop_user = const_cast<Deadline*>(d);
  }
  return h;
}

void check(int rc) {  // driver errors become PramanaError, as cuda_driver.cpp:164-169
  if (rc != 0) throw PramanaError(std::string("pdhg_hip: ") + pdhg_last_error());
}

void copy_stats(const pdhg_stats& s, PdhgStats& st) {  // PdhgStats, pdhg.h:44-56
  st.iterations = s.iterations;
  st.restarts = s.restarts;
  st.primalResidual = s.primal_residual;
  st.dualResidual = s.dual_residual;
  st.gap = s.gap;
  st.primalWeight = s.primal_weight;
  st.setupSeconds = s.setup_seconds;
  st.solveSeconds = s.solve_seconds;
  st.kernelSeconds = s.kernel_seconds;
  st.transferSeconds = s.transfer_seconds;
  st.bytesUploaded = s.bytes_uploaded;
  st.bytesDownloaded = s.bytes_downloaded;
  st.kernelLaunches = s.kernel_launches;
  st.spectralNorm = s.spectral_norm;
  st.backend = "gpu";  // the reference's label: solve() reports the engine as pdhg-gpu (solver.cpp:164)
  st.device = s.device_name;
}

bool use_hip(const PdhgOptions& o) { return o.useGpu && pdhg_gpu_available() == 1; }

}  // namespace

std::vector<PdhgResult> solvePdhgBatch(const Model& base, const std::vector<LpCase>& cases, const PdhgOptions& opts,
                                       const Deadline* deadline, PdhgStats* batchStats) {
  if (!use_hip(opts)) {
    PdhgOptions cpu = opts;
    cpu.useGpu = false;
    return refSolvePdhgBatch(base, cases, cpu, deadline, batchStats);
  }
  if (cases.empty()) return {};  // as the reference (pdhg.cpp:724)
  const pdhg_lp lp = view(base);
  const pdhg_options o = options(opts, deadline);
  const int k = static_cast<int>(cases.size());
  auto ptr = [](const std::vector<double>& v) { return v.empty() ? nullptr : v.data(); };  // empty = base model
  std::vector<pdhg_case> hc(k);
  std::vector<pdhg_result> hr(k);
  std::vector<PdhgResult> out(k);
  for (int b = 0; b < k; ++b) {
    hc[b] = pdhg_case{ptr(cases[b].cost), ptr(cases[b].colLower), ptr(cases[b].colUpper), ptr(cases[b].rowLower),
                      ptr(cases[b].rowUpper), nullptr, nullptr, 0.0};
    out[b].x.assign(lp.n, 0.0);
    out[b].rowDual.assign(lp.m, 0.0);
    hr[b] = pdhg_result{};
    hr[b].x = out[b].x.data();
    hr[b].row_dual = out[b].rowDual.data();
  }
  pdhg_stats st{};
  check(pdhg_solve_batch(&lp, hc.data(), k, &o, hr.data(), &st));
  PdhgStats rs;
  copy_stats(st, rs);
  for (int b = 0; b < k; ++b) {
    out[b].status = static_cast<Status>(hr[b].status);  // the same numbering (src/util/common.h:34-46)
    out[b].objective = hr[b].objective;
    out[b].dualObjective = hr[b].dual_objective;
    out[b].iterations = hr[b].iterations;
    out[b].relativeKkt = hr[b].relative_kkt;
    out[b].stats = rs;  // the reference gives every result the batch's stats (pdhg.cpp:685)
  }
  if (batchStats) *batchStats = rs;
  return out;
}

PdhgResult solvePdhg(const Model& lp, const PdhgOptions& opts, const Deadline* deadline) {
  if (!use_hip(opts)) {
    PdhgOptions cpu = opts;
    cpu.useGpu = false;
    return refSolvePdhg(lp, cpu, deadline);
  }
  PdhgStats st;
  std::vector<PdhgResult> r = solvePdhgBatch(lp, {LpCase{}}, opts, deadline, &st);
  r[0].stats = st;
  return std::move(r[0]);
}

bool gpuAvailable() { return pdhg_gpu_available() == 1; }

Json gpuInfoJson() {
  char buf[2048];
  check(pdhg_gpu_info_json(buf, sizeof buf));
  return Json::parse(buf);
}

}  // namespace pramana
