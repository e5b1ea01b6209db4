// Interface between the host engine (plain C++) and the HIP side.  Nothing
// here names a HIP type, so the host translation units build with g++.
#pragma once

#include <memory>
#include <string>

#include "prepare.h"

namespace pdhgamd {

struct DeviceInfo {
  bool available = false;
  std::string reason, name, arch;
  int computeUnits = 0;
  size_t memoryBytes = 0;
  double peakBandwidthGBs = 0;
  int ordinal = 0;
};

// Device ordinals the engine addresses (a fixed table: the probe results keep
// their addresses for the life of the process).
constexpr int kMaxDevices = 64;

// Probes the HIP runtime once per device and process (the analogue of the
// reference's CudaContext singleton, src/gpu/cuda_driver.cpp:148-162).  An
// ordinal outside 0 .. kMaxDevices-1 reads as an unavailable device.
const DeviceInfo& device_info(int device);

// Row-to-lane mapping of one CSR matrix.
struct MapSetup {
  int nnzPerLane = 4;  // target nonzeros per lane: lanes per row = nextpow2(ceil(len / nnzPerLane))
  int rowMap = 1;      // 1: row-length buckets (rows permuted); 2: one mapping for all rows
  bool xcd = false;    // give each XCD a contiguous range of rows
  // B = 16 or 32: one row per B-lane group, 64 / B rows per wavefront, the
  // rows' entries read through the scalar cache (nnzPerLane unused; rowMap 1
  // orders rows by length, 2 keeps their order).
  bool group = false;
  // B = 64: rows per wavefront (1, or 2 with their chunks interleaved).
  int rowsPerWave = 1;
  // B = 64, iterations without checks: rows in locality order, grouped in
  // tiles whose distinct gathered rows a workgroup stages in LDS once
  // (0: no tiles; else the tile shape of tiles.h, 1..kTileShapes).
  int tile = 0;
};

struct RunSetup {
  int m = 0, n = 0;
  int B = 1;           // instances per run: a power of two <= 64
  int precision = 0;   // 0: FP32 iterates, FP64 norms; 1: FP64
  bool sharedCost = false, sharedColBounds = false, sharedRowBounds = false;
  MapSetup colMap, rowMap;
  double rho = 1.0;
  int checkEvery = 64;
  bool useGraph = true;
};

// Per-instance record written at every check (8 doubles, the trace format).
struct CheckRecord {
  double it, relP, relD, gap, pobj, dobj, omega, restarted;
};

struct RunScalars {
  double eta = 1, beta = 1, gamma = 1, offset = 0, tol = 1e-6;
  long long maxIter = 1LL << 62;  // a case's own iteration limit (its last check)
};

// Host copy (FP64) of the iterate and restart state of some lanes of a run,
// k-interleaved over those lanes.  Moves live instances into a narrower run.
struct LaneState {
  int count = 0;
  std::vector<double> x, x0, xp, y, y0, yp, ax, ax0, axp;
  std::vector<double> tau, sigma, omega, fpR, fpPrev;
  std::vector<long long> iter, since, restartAt, applyAt;
  std::vector<int> restarts;
};

// Device state of one batched restarted-PDHG run: B instances sharing A.
class DeviceRun {
 public:
  virtual ~DeviceRun() = default;
  static std::unique_ptr<DeviceRun> create(int device, const Prepared& P, const RunSetup& s);

  // Scaled case data, FP64 on the host: k-interleaved (i*B + b) or, for the
  // arrays flagged shared in RunSetup, one value per row / column.
  virtual void upload(const double* c, const double* lo, const double* up, const double* rl, const double* ru,
                      const double* x, const double* y) = 0;
  // Per-instance scalars; done[b] != 0 marks a padding lane.
  virtual void init_state(const RunScalars& s, const double* omega, const double* cNorm, const double* rhsNorm,
                          const int* done) = 0;
  // initAx: ax = A x, ax0 = ax (a fresh run).  Starts the loop timer.
  virtual void start(bool initAx) = 0;
  // `iters` fused iterations, the last one a check (KKT, restart, primal weight).
  virtual void run_block(int iters) = 0;
  // Waits for the stream, then copies the last check records and done flags.
  virtual void fetch(CheckRecord* rec, int* done, long long* doneIter, int* restarts) = 0;
  // k-interleaved copy of an iterate (0: x+, 1: y+) in FP64.
  virtual void download(int which, double* out) = 0;
  // Stops the loop timer: device seconds since start().
  virtual double stop() = 0;
  // Mean duration (ms) of the column and of the row iteration kernel, each
  // over `reps` back-to-back launches, and of one iteration (the two in
  // turn) over `reps` iterations, HIP events on the run's stream.
  virtual void time_kernels(int reps, double* msCol, double* msRow, double* msIter) = 0;
  // Algorithmic bytes of one launch of the column and row iteration kernels.
  virtual void kernel_bytes(double* colBytes, double* rowBytes) const = 0;
  // Copies the state of `lanes` (in that order) to the host.
  virtual void export_lanes(const std::vector<int>& lanes, LaneState& out) = 0;
  // Overwrites iterates and per-instance state of lanes 0..in.count-1 (after
  // upload() and init_state() of this run).
  virtual void import_lanes(const LaneState& in) = 0;
  // Continuous batching: starts new cases in lanes `lanes` (each finished or
  // free) while the other lanes keep running.  Arrays are interleaved over
  // the K = lanes.size() new cases (element (i, q) at i*K + q), in the same
  // scaled space as upload(); case arrays flagged shared in RunSetup are
  // NULL.  The lanes get ax = A x, fresh anchors and restart state, and
  // their iteration counters start at 0.
  virtual void load_lanes(const std::vector<int>& lanes, const double* c, const double* lo, const double* up,
                          const double* rl, const double* ru, const double* x, const double* y, const double* omega,
                          const double* cNorm, const double* rhsNorm) = 0;
  // Stops lanes (their kernels skip them from the next iteration on).
  virtual void retire_lanes(const std::vector<int>& lanes) = 0;
  // Autotuner: times iterations (column and row kernel in turn, `reps` of
  // them, HIP events on the run's stream) with each candidate row mapping,
  // first for the column kernel (with the run's row mapping), then for the
  // row kernel (with the chosen column mapping), and keeps the fastest per
  // direction.  Runs on scratch state, so it must come before upload() /
  // init_state(); best[0] / best[1] get the chosen candidate index for the
  // column / row kernel, msIter the mean iteration under both choices and
  // msIterHalf the same with only the first half of the lanes live (lanes
  // B/2..B-1 done), the state of a run into which half as many cases as
  // lanes were re-packed.
  virtual void autotune(const Prepared& P, const std::vector<MapSetup>& cand, int reps, int* best, double* msIter,
                        double* msIterHalf) = 0;
  // The tile shape (MapSetup::tile) in which the iterations without checks
  // run, 0 without tiles, for the column (0) or row (1) kernel.
  virtual int tiled(int side) const = 0;

  long long launches = 0, graphLaunches = 0, bytesUp = 0, bytesDown = 0;
};

}  // namespace pdhgamd
