// HIP kernels of the restarted-PDHG engine for gfx950 (MI355X, CDNA4) and the
// device runtime that drives them.
//
// Layout: B instances share the scaled matrix; an iterate element (i, b) sits
// at i*B + b, so for every nonzero a wavefront reads B consecutive values.
// A row is served by a group of B*S lanes (lane = instance + B * sub-lane); S
// (a power of two) is chosen per row-length bucket so that each lane streams a
// few consecutive nonzeros.  Rows of one bucket are packed into 256-thread
// workgroups; a workgroup only ever holds rows of one bucket, so shuffle
// reductions are uniform inside it.  At B = 64 a wavefront is one row (lane =
// instance); the iterations without checks may instead run over row tiles
// whose distinct gathered rows a 1024-thread workgroup stages in LDS once
// (k_col_tiled / k_row_tiled, same sums).  No atomics anywhere: every
// cross-lane or cross-block sum is a fixed-order tree.
//
// One PDHG iteration is two kernels (reference src/pdhg/pdhg.cpp:588-662):
//   k_col : aty = A^T y, x+ = proj(x - tau (c - aty)), Halpern update of x
//   k_row : ax+ = A x+, reflected dual step, Halpern update of y and ax
// and a check iteration adds k_colcheck (FP64 A^T y+ and the column KKT terms)
// and k_finalize (FP64 KKT / gap / restart / primal weight, one block per
// instance, at each instance's own check points).  Restarts are applied by
// the first iteration after the check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pdhg_hip.h"
#include "device.h"

#include "kcommon.h"

namespace pdhgamd {

namespace {

// ---------------------------------------------------------------------------
// Iteration kernels
// ---------------------------------------------------------------------------

// Column side of an iteration over the rows of A^T (reference pdhg.cpp:589-590,
// 116-125, 138-152): x+ = clamp(x - tau (c - A^T y), lo, up), then
// x <- lam((1+rho) x+ - rho x) + (1-lam) x0 with lam = (k+1)/(k+2).
// CHECK adds ||x - x+||^2 and ||x+ - x0||^2 (quantities 0, 1).
// col_update is that step for column j of lane b, given acc = (A^T y)_j;
// col_load / col_apply split it around the dot product (operands first).
template <typename T>
struct ColOps {
  T xin, xo, c, lo, up;
};

template <typename T, int LOGB>
__device__ inline ColOps<T> col_load(int j, int b, bool rst, const T* __restrict__ x, const T* __restrict__ xp,
                                     const T* __restrict__ x0, const T* __restrict__ c, const Ld& lc,
                                     const T* __restrict__ lo, const T* __restrict__ up, const Ld& lb) {
  constexpr bool SPL = kSplit<T, LOGB>;
  ColOps<T> o;
  o.xin = vld<SPL, LOGB>(rst ? xp : x, j, b);
  o.xo = rst ? o.xin : vld<SPL, LOGB>(x0, j, b);
  o.c = cld<SPL, LOGB>(c, lc, j, b);
  o.lo = cld<SPL, LOGB>(lo, lb, j, b);
  o.up = cld<SPL, LOGB>(up, lb, j, b);
  return o;
}

template <typename T, int LOGB, bool CHECK>
__device__ inline void col_apply(int j, T acc, const ColOps<T>& o, int b, bool rst, long long it, T* __restrict__ x,
                                 T* __restrict__ xp, T* __restrict__ x0, const Inst& st, double rho,
                                 double (&q)[2]) {
  constexpr bool SPL = kSplit<T, LOGB>;
  if (rst) vst<SPL, LOGB>(x0, j, b, o.xin);
  const T tau = static_cast<T>(st.tau[b]);
  const T v = clamp_to<T>(o.xin - tau * (o.c - acc), o.lo, o.up);
  const double kk = static_cast<double>(it - st.restartAt[b]);
  const double lam = (kk + 1.0) / (kk + 2.0);
  vst<SPL, LOGB>(x, j, b,
                 static_cast<T>(lam) * (static_cast<T>(1.0 + rho) * v - static_cast<T>(rho) * o.xin) +
                     static_cast<T>(1.0 - lam) * o.xo);
  vst<SPL, LOGB>(xp, j, b, v);
  if (CHECK) {
    const double d = static_cast<double>(o.xin) - static_cast<double>(v);
    const double e = static_cast<double>(v) - static_cast<double>(o.xo);
    q[0] += d * d;
    q[1] += e * e;
  }
}

template <typename T, int LOGB, bool CHECK>
__device__ inline void col_update(int j, T acc, int b, bool rst, long long it, T* __restrict__ x,
                                  T* __restrict__ xp, T* __restrict__ x0, const T* __restrict__ c, const Ld& lc,
                                  const T* __restrict__ lo, const T* __restrict__ up, const Ld& lb, const Inst& st,
                                  double rho, double (&q)[2]) {
  const ColOps<T> o = col_load<T, LOGB>(j, b, rst, x, xp, x0, c, lc, lo, up, lb);
  col_apply<T, LOGB, CHECK>(j, acc, o, b, rst, it, x, xp, x0, st, rho, q);
}

template <typename T, int LOGB, bool CHECK, int RPW = 1>
__global__ __launch_bounds__(kWG) void k_col(RowPlan plan, const int* __restrict__ ptr, const int* __restrict__ idx,
                                             const T* __restrict__ val, const T* __restrict__ y,
                                             const T* __restrict__ yp, T* __restrict__ x, T* __restrict__ xp,
                                             T* __restrict__ x0, const T* __restrict__ c, Ld lc,
                                             const T* __restrict__ lo, const T* __restrict__ up, Ld lb, Inst st,
                                             int tloc, double rho, double* __restrict__ part, int nblk) {
  constexpr int B = 1 << LOGB;
  constexpr bool SPL = kSplit<T, LOGB>;
  const int bid = block_id(plan);
  const Slot s = locate<LOGB>(plan, bid);
  const int b = threadIdx.x & (B - 1);
  const long long it = st.iter[b] + tloc;
  // An instance restarted at the last check continues from (x+, y+): its A^T y
  // of this iteration reads y+ (reference pdhg.cpp:160-162 then :589).
  const bool rst = st.applyAt[b] == it;
  const T* __restrict__ ysrc = rst ? yp : y;
  int rows[2];
  T accs[2];
  lane_rows<LOGB, RPW, T, SPL>(plan, bid, s, ptr, idx, val, ysrc, b, rows, accs);
  double q[2] = {0.0, 0.0};
  for (int rr = 0; rr < 2; ++rr) {
    if (rows[rr] < 0 || s.sub != 0 || st.done[b]) continue;
    col_update<T, LOGB, CHECK>(rows[rr], accs[rr], b, rst, it, x, xp, x0, c, lc, lo, up, lb, st, rho, q);
  }
  if (CHECK) block_partials<LOGB, 2>(q, part, 0, nblk, bid);
}

// Row side of an iteration over the rows of A (reference pdhg.cpp:591-592,
// 126-137, 138-152): ax+ = A x+, v = y - sigma (2 ax+ - ax),
// y+ = v + sigma clamp(-v / sigma, L, U), Halpern updates of y and ax.
// CHECK also stores y+ and ax+, recomputes A x+ with FP64 values and FP64
// accumulation, and adds (quantities 0..3): sum res^2, the row part of the
// dual objective, ||y - y+||^2, ||y+ - y0||^2.
// row_update is that step for row i of lane b, given acc = (A x+)_i (and its
// FP64 recomputation acc64 when CHECK); row_load / row_apply split it around
// the dot product (operands first).
template <typename T>
struct RowOps {
  T yin, axin, yo, axo, rl, ru;
};

template <typename T, int LOGB>
__device__ inline RowOps<T> row_load(int i, int b, bool rst, const T* __restrict__ y, const T* __restrict__ yp,
                                     const T* __restrict__ y0, const T* __restrict__ ax, const T* __restrict__ axp,
                                     const T* __restrict__ ax0, const T* __restrict__ rl, const T* __restrict__ ru,
                                     const Ld& lr) {
  constexpr bool SPL = kSplit<T, LOGB>;
  RowOps<T> o;
  o.yin = vld<SPL, LOGB>(rst ? yp : y, i, b);
  o.axin = vld<SPL, LOGB>(rst ? axp : ax, i, b);
  o.yo = rst ? o.yin : vld<SPL, LOGB>(y0, i, b);
  o.axo = rst ? o.axin : vld<SPL, LOGB>(ax0, i, b);
  o.rl = cld<SPL, LOGB>(rl, lr, i, b);
  o.ru = cld<SPL, LOGB>(ru, lr, i, b);
  return o;
}

template <typename T, int LOGB, bool CHECK>
__device__ inline void row_apply(int i, T acc, double acc64, const RowOps<T>& o, int b, bool rst, long long it,
                                 T* __restrict__ y, T* __restrict__ yp, T* __restrict__ y0, T* __restrict__ ax,
                                 T* __restrict__ axp, T* __restrict__ ax0, const double* __restrict__ rl64,
                                 const double* __restrict__ ru64, const Ld& lr, const double* __restrict__ rfac,
                                 const Inst& st, double rho, double (&q)[4]) {
  constexpr bool SPL = kSplit<T, LOGB>;
  if (rst) {
    vst<SPL, LOGB>(y0, i, b, o.yin);
    vst<SPL, LOGB>(ax0, i, b, o.axin);
  }
  // Reflected dual step y+ = v + sigma clamp(-v/sigma, L, U), v = y - sigma g,
  // g = 2 ax+ - ax (reference pdhg.cpp:126-137), evaluated as
  // y+ = y - sigma (g - w) with g - w in {g - L, g - U, y / sigma}: the
  // large activity g only meets the bound it is compared with, so y keeps
  // its low-order bits in FP32 (v + sigma w would cancel them away).
  const T sg = static_cast<T>(st.sigma[b]);
  const T g = static_cast<T>(2) * acc - o.axin;
  const T tq = o.yin / sg;
  const T dlo = g - o.rl, dhi = g - o.ru;
  const T yplus = dlo < tq ? o.yin - sg * dlo : (dhi > tq ? o.yin - sg * dhi : static_cast<T>(0));
  const double kk = static_cast<double>(it - st.restartAt[b]);
  const double lam = (kk + 1.0) / (kk + 2.0);
  const T lamT = static_cast<T>(lam), omT = static_cast<T>(1.0 - lam);
  const T r1 = static_cast<T>(1.0 + rho), r0 = static_cast<T>(rho);
  vst<SPL, LOGB>(y, i, b, lamT * (r1 * yplus - r0 * o.yin) + omT * o.yo);
  vst<SPL, LOGB>(ax, i, b, lamT * (r1 * acc - r0 * o.axin) + omT * o.axo);
  if (CHECK) {
    vst<SPL, LOGB>(yp, i, b, yplus);
    vst<SPL, LOGB>(axp, i, b, acc);
    const double L = rl64[lr.at(i, b)], U = ru64[lr.at(i, b)];
    const double res = (fmax(L - acc64, 0.0) + fmax(acc64 - U, 0.0)) * rfac[i];
    const double yy = static_cast<double>(yplus);
    q[0] += res * res;
    if (yy > 0 && L > -1e300) q[1] += yy * L;
    else if (yy < 0 && U < 1e300) q[1] += yy * U;
    const double d = static_cast<double>(o.yin) - yy, e = yy - static_cast<double>(o.yo);
    q[2] += d * d;
    q[3] += e * e;
  }
}

template <typename T, int LOGB, bool CHECK>
__device__ inline void row_update(int i, T acc, double acc64, int b, long long it, T* __restrict__ y,
                                  T* __restrict__ yp, T* __restrict__ y0, T* __restrict__ ax, T* __restrict__ axp,
                                  T* __restrict__ ax0, const T* __restrict__ rl, const T* __restrict__ ru,
                                  const double* __restrict__ rl64, const double* __restrict__ ru64, const Ld& lr,
                                  const double* __restrict__ rfac, const Inst& st, double rho, double (&q)[4]) {
  const bool rst = st.applyAt[b] == it;
  const RowOps<T> o = row_load<T, LOGB>(i, b, rst, y, yp, y0, ax, axp, ax0, rl, ru, lr);
  row_apply<T, LOGB, CHECK>(i, acc, acc64, o, b, rst, it, y, yp, y0, ax, axp, ax0, rl64, ru64, lr, rfac, st, rho, q);
}

template <typename T, int LOGB, bool CHECK, int RPW = 1>
__global__ __launch_bounds__(kWG) void k_row(RowPlan plan, const int* __restrict__ ptr, const int* __restrict__ idx,
                                             const T* __restrict__ val, const double* __restrict__ val64,
                                             const T* __restrict__ xp, T* __restrict__ y, T* __restrict__ yp,
                                             T* __restrict__ y0, T* __restrict__ ax, T* __restrict__ axp,
                                             T* __restrict__ ax0, const T* __restrict__ rl, const T* __restrict__ ru,
                                             const double* __restrict__ rl64, const double* __restrict__ ru64, Ld lr,
                                             const double* __restrict__ rfac, Inst st, int tloc, double rho,
                                             double* __restrict__ part, int nblk) {
  constexpr int B = 1 << LOGB;
  constexpr bool SPL = kSplit<T, LOGB>;
  const int bid = block_id(plan);
  const Slot s = locate<LOGB>(plan, bid);
  const int b = threadIdx.x & (B - 1);
  int rows[2];
  T accs[2];
  double accs64[2] = {0.0, 0.0};
  lane_rows<LOGB, RPW, T, SPL>(plan, bid, s, ptr, idx, val, xp, b, rows, accs);
  if constexpr (CHECK) {
    if constexpr (sizeof(T) == 8) {  // FP64 iterates: val is val64, the same sums
      accs64[0] = accs[0];
      accs64[1] = accs[1];
    } else {
      lane_rows<LOGB, RPW, double, SPL>(plan, bid, s, ptr, idx, val64, xp, b, rows, accs64);
    }
  }
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  const long long it = st.iter[b] + tloc;
  for (int rr = 0; rr < 2; ++rr) {
    if (rows[rr] < 0 || s.sub != 0 || st.done[b]) continue;
    row_update<T, LOGB, CHECK>(rows[rr], accs[rr], accs64[rr], b, it, y, yp, y0, ax, axp, ax0, rl, ru, rl64, ru64,
                               lr, rfac, st, rho, q);
  }
  if (CHECK) block_partials<LOGB, 4>(q, part, 0, nblk, bid);
}

// B = 64 tiled forms of k_col / k_row (no check quantities): a workgroup
// takes one tile (up to tile_waves(SHAPE) rows in locality order), each
// wavefront one row of it.  The wavefront's record (row, entry range, staged columns)
// arrives in one round of scalar loads; one round of vector loads then
// brings the tile's distinct gathered rows (staged in LDS) and the row's
// update operands; after one barrier the dot products read LDS.  loc holds
// tile-local columns.
template <typename T, int SHAPE>
__global__ __launch_bounds__(64 * tile_waves(SHAPE)) __attribute__((amdgpu_waves_per_eu(8))) void k_col_tiled(RowPlan plan, const int* __restrict__ loc,
                                                   const T* __restrict__ val, const T* __restrict__ y,
                                                   const T* __restrict__ yp, T* __restrict__ x, T* __restrict__ xp,
                                                   T* __restrict__ x0, const T* __restrict__ c, Ld lc,
                                                   const T* __restrict__ lo, const T* __restrict__ up, Ld lb, Inst st,
                                                   int tloc, double rho) {
  constexpr bool SPL = kSplit<T, 6>;
  constexpr int WAVES = tile_waves(SHAPE);
  constexpr int NS = kStageCols<T, SHAPE>;
  __shared__ T tile[kTileCols<T, SHAPE> * 64];
  __shared__ int entL[WAVES * 64];
  __shared__ T entV[WAVES * 64];
  const int t = block_id(plan);
  const int b = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const TileRec<NS> r = tile_record<NS, WAVES>(plan, t);
  const long long it = st.iter[b] + tloc;
  const bool rst = st.applyAt[b] == it;
  const bool live = !st.done[b];  // a finished instance's lane loads nothing
  // A row of at most 64 entries brings them in the same round of loads,
  // one per lane, and reads them back from LDS.
  const int len = r.e1 - r.e0;
  const bool ent = len <= 64;
  int eL = 0;
  T eV = 0;
  if (ent && b < len) {
    eL = loc[r.e0 + b];
    eV = val[r.e0 + b];
  }
  T stage[NS];
  tile_load<SPL>(r, rst ? yp : y, b, live, stage);
  const int j = r.row;
  ColOps<T> ops;
  if (j >= 0 && live) ops = col_load<T, 6>(j, b, rst, x, xp, x0, c, lc, lo, up, lb);
  if (ent && b < len) {
    entL[wave * 64 + b] = eL;
    entV[wave * 64 + b] = eV;
  }
  tile_store<WAVES>(r, stage, tile, b);
  if (j < 0) return;
  const T acc = ent ? row_dot_lds_staged<T>(entL + wave * 64, entV + wave * 64, tile, len, b)
                    : row_dot_lds<T>(loc, val, tile, r.e0, r.e1, b);
  double q[2] = {0.0, 0.0};
  if (live) col_apply<T, 6, false>(j, acc, ops, b, rst, it, x, xp, x0, st, rho, q);
}

template <typename T, int SHAPE>
__global__ __launch_bounds__(64 * tile_waves(SHAPE)) __attribute__((amdgpu_waves_per_eu(8))) void k_row_tiled(RowPlan plan, const int* __restrict__ loc,
                                                   const T* __restrict__ val, const T* __restrict__ xp,
                                                   T* __restrict__ y, T* __restrict__ yp, T* __restrict__ y0,
                                                   T* __restrict__ ax, T* __restrict__ axp, T* __restrict__ ax0,
                                                   const T* __restrict__ rl, const T* __restrict__ ru,
                                                   const double* __restrict__ rl64, const double* __restrict__ ru64,
                                                   Ld lr, const double* __restrict__ rfac, Inst st, int tloc,
                                                   double rho) {
  constexpr bool SPL = kSplit<T, 6>;
  constexpr int WAVES = tile_waves(SHAPE);
  constexpr int NS = kStageCols<T, SHAPE>;
  __shared__ T tile[kTileCols<T, SHAPE> * 64];
  __shared__ int entL[WAVES * 64];
  __shared__ T entV[WAVES * 64];
  const int t = block_id(plan);
  const int b = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const TileRec<NS> r = tile_record<NS, WAVES>(plan, t);
  const long long it = st.iter[b] + tloc;
  const bool rst = st.applyAt[b] == it;
  const bool live = !st.done[b];
  const int len = r.e1 - r.e0;
  const bool ent = len <= 64;
  int eL = 0;
  T eV = 0;
  if (ent && b < len) {
    eL = loc[r.e0 + b];
    eV = val[r.e0 + b];
  }
  T stage[NS];
  tile_load<SPL>(r, xp, b, live, stage);
  const int i = r.row;
  RowOps<T> ops;
  if (i >= 0 && live) ops = row_load<T, 6>(i, b, rst, y, yp, y0, ax, axp, ax0, rl, ru, lr);
  if (ent && b < len) {
    entL[wave * 64 + b] = eL;
    entV[wave * 64 + b] = eV;
  }
  tile_store<WAVES>(r, stage, tile, b);
  if (i < 0) return;
  const T acc = ent ? row_dot_lds_staged<T>(entL + wave * 64, entV + wave * 64, tile, len, b)
                    : row_dot_lds<T>(loc, val, tile, r.e0, r.e1, b);
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  if (live)
    row_apply<T, 6, false>(i, acc, 0.0, ops, b, rst, it, y, yp, y0, ax, axp, ax0, rl64, ru64, lr, rfac, st, rho, q);
}

// Column KKT terms at (x+, y+) in FP64 (reference pdhg.cpp:181-196):
// d = c - A^T y+ ; dual residual (by bound pattern) * cfac, c'x+, and the
// column part of the dual objective (quantities 2, 3, 4).
template <typename T, int LOGB, int RPW = 1>
__global__ __launch_bounds__(kWG) void k_colcheck(RowPlan plan, const int* __restrict__ ptr,
                                                  const int* __restrict__ idx, const double* __restrict__ val64,
                                                  const T* __restrict__ yp, const T* __restrict__ xp,
                                                  const double* __restrict__ c64, Ld lc,
                                                  const double* __restrict__ lo64, const double* __restrict__ up64,
                                                  Ld lb, const double* __restrict__ cfac, Inst st,
                                                  double* __restrict__ part, int nblk) {
  constexpr int B = 1 << LOGB;
  const int bid = block_id(plan);
  const Slot s = locate<LOGB>(plan, bid);
  const int b = threadIdx.x & (B - 1);
  constexpr bool SPL = kSplit<T, LOGB>;
  int rows[2];
  double accs[2];
  lane_rows<LOGB, RPW, double, SPL>(plan, bid, s, ptr, idx, val64, yp, b, rows, accs);
  double q[3] = {0.0, 0.0, 0.0};
  for (int rr = 0; rr < 2; ++rr) {
    if (rows[rr] < 0 || s.sub != 0 || st.done[b]) continue;
    const int j = rows[rr];
    const double acc = accs[rr];
    const double cj = c64[lc.at(j, b)], l = lo64[lb.at(j, b)], u = up64[lb.at(j, b)];
    const double d = cj - acc;
    const bool hl = l > -1e300, hu = u < 1e300;
    double res = 0;
    if (!hl && !hu) res = fabs(d);
    else if (!hl) res = fmax(d, 0.0);
    else if (!hu) res = fmax(-d, 0.0);
    res *= cfac[j];
    q[0] += res * res;
    q[1] += cj * static_cast<double>(vld<SPL, LOGB>(xp, j, b));
    if (d > 0 && hl) q[2] += d * l;
    else if (d < 0 && hu) q[2] += d * u;
  }
  block_partials<LOGB, 3>(q, part, 2, nblk, bid);
}

// Check logic of one instance per workgroup, all FP64 (reference
// pdhg.cpp:604-656): fixed-order sums of the partials, relative KKT,
// convergence, the restart rule, the primal-weight update and new steps.
__global__ __launch_bounds__(kWG) void k_finalize(Inst st, const double* __restrict__ colPart, int nblkC,
                                                  const double* __restrict__ rowPart, int nblkR, RunScalars sc,
                                                  int blockIters, int checkEvery) {
  const int b = blockIdx.x, B = gridDim.x;
  if (st.done[b]) return;
  constexpr int NQ = 9;  // col 0..4, row 0..3
  // Each thread sums partials k = tid, tid + kWG, ... of every quantity in
  // that order; the quantities share one loop so their loads are in flight
  // together.
  double v[NQ] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double* pc = colPart + static_cast<size_t>(b) * nblkC;
  const size_t sc5 = static_cast<size_t>(B) * nblkC;
#pragma unroll 2
  for (int k = threadIdx.x; k < nblkC; k += kWG) {
#pragma unroll
    for (int q = 0; q < 5; ++q) v[q] += pc[q * sc5 + k];
  }
  const double* pr = rowPart + static_cast<size_t>(b) * nblkR;
  const size_t sr4 = static_cast<size_t>(B) * nblkR;
#pragma unroll 2
  for (int k = threadIdx.x; k < nblkR; k += kWG) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[5 + q] += pr[q * sr4 + k];
  }
  __shared__ double sh[NQ][kWG / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double s = v[q];
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) sh[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double S[NQ];
  for (int q = 0; q < NQ; ++q) {
    S[q] = sh[q][0];
    for (int w = 1; w < kWaves; ++w) S[q] += sh[q][w];
  }
  finalize_instance(st, b, S, sc, blockIters, checkEvery);
}

// ---------------------------------------------------------------------------
// Lane loading (continuous batching)
// ---------------------------------------------------------------------------
// dst(i, lanes[q]) = src[i*K + q] for i < len, q < K, in the run's layout
// (split words when `split`: FP64 at B >= 16).
template <typename U>
__global__ void k_scatter_lanes(U* __restrict__ dst, const U* __restrict__ src, const int* __restrict__ lanes,
                                long long len, int K, int logB, int split) {
  const long long total = len * K;
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long i = t / K;
    const int b = lanes[t % K];
    const U v = src[t];
    if constexpr (sizeof(U) == 8) {
      if (split) {
        unsigned* u = reinterpret_cast<unsigned*>(dst) + (i << (logB + 1));
        const unsigned long long bits = __builtin_bit_cast(unsigned long long, v);
        u[b] = static_cast<unsigned>(bits);
        u[(1 << logB) + b] = static_cast<unsigned>(bits >> 32);
        continue;
      }
    }
    dst[(i << logB) + b] = v;
  }
}

// dst(i, lanes[q]) = src(i, lanes[q]) for i < len (same layout; split FP64
// words are copied as two dwords).
template <typename U>
__global__ void k_copy_lanes(U* __restrict__ dst, const U* __restrict__ src, const int* __restrict__ lanes,
                             long long len, int K, int logB, int split) {
  const long long total = len * K;
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long i = t / K;
    const int b = lanes[t % K];
    if constexpr (sizeof(U) == 8) {
      if (split) {
        const size_t o = static_cast<size_t>(i) << (logB + 1);
        unsigned* ud = reinterpret_cast<unsigned*>(dst) + o;
        const unsigned* us = reinterpret_cast<const unsigned*>(src) + o;
        ud[b] = us[b];
        ud[(1 << logB) + b] = us[(1 << logB) + b];
        continue;
      }
    }
    const size_t e = (static_cast<size_t>(i) << logB) + b;
    dst[e] = src[e];
  }
}

// ---------------------------------------------------------------------------
// Op-level kernels (the Backend vtable, reference pdhg.cpp:51-70)
// ---------------------------------------------------------------------------
template <typename T, int LOGB, bool SPL, int RPW = 1>
__global__ __launch_bounds__(kWG) void k_spmv(RowPlan plan, const int* __restrict__ ptr, const int* __restrict__ idx,
                                              const T* __restrict__ val, const T* __restrict__ in,
                                              T* __restrict__ out) {
  constexpr int B = 1 << LOGB;
  const int bid = block_id(plan);
  const Slot s = locate<LOGB>(plan, bid);
  const int b = threadIdx.x & (B - 1);
  int rows[2];
  T accs[2];
  lane_rows<LOGB, RPW, T, SPL>(plan, bid, s, ptr, idx, val, in, b, rows, accs);
  for (int rr = 0; rr < 2; ++rr)
    if (rows[rr] >= 0 && s.sub == 0) vst<SPL, LOGB>(out, rows[rr], b, accs[rr]);
}

// B = 64 SpMV over LDS-staged row tiles of shape SHAPE (tiles.h): the column
// and row kernels' tiled product without their update, the same sums as one
// row per wavefront.
template <typename T, int SHAPE>
__global__ __launch_bounds__(64 * tile_waves(SHAPE)) __attribute__((amdgpu_waves_per_eu(8))) void k_spmv_tiled(
    RowPlan plan, const int* __restrict__ loc, const T* __restrict__ val, const T* __restrict__ in,
    T* __restrict__ out) {
  constexpr bool SPL = kSplit<T, 6>;
  constexpr int WAVES = tile_waves(SHAPE);
  constexpr int NS = kStageCols<T, SHAPE>;
  __shared__ T tile[kTileCols<T, SHAPE> * 64];
  __shared__ int entL[WAVES * 64];
  __shared__ T entV[WAVES * 64];
  const int t = block_id(plan);
  const int b = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const TileRec<NS> r = tile_record<NS, WAVES>(plan, t);
  const int len = r.e1 - r.e0;
  const bool ent = len <= 64;
  int eL = 0;
  T eV = 0;
  if (ent && b < len) {
    eL = loc[r.e0 + b];
    eV = val[r.e0 + b];
  }
  T stage[NS];
  tile_load<SPL>(r, in, b, true, stage);
  if (ent && b < len) {
    entL[wave * 64 + b] = eL;
    entV[wave * 64 + b] = eV;
  }
  tile_store<WAVES>(r, stage, tile, b);
  if (r.row < 0) return;
  const T acc = ent ? row_dot_lds_staged<T>(entL + wave * 64, entV + wave * 64, tile, len, b)
                    : row_dot_lds<T>(loc, val, tile, r.e0, r.e1, b);
  vst<SPL, 6>(out, r.row, b, acc);
}

// Width-1 SpMV over a sliced ELLPACK copy of the matrix (slices of 64 rows,
// SellDev): a wavefront takes a slice, lane i its row perm[64 s + i].  The
// slice stores entry k of its 64 rows at soff[s] + 64 k + i (rows shorter
// than the slice padded with weight 0), so every index and value load is
// one contiguous run of 64 elements, and neighbouring rows, whose entries
// share columns, gather together.  Each lane sums its row in entry order
// over four chains (entry k into chain k mod 4); eight entries' loads are
// issued before their multiply-adds.
template <typename T>
__global__ __launch_bounds__(kWG) void k_spmv_sell(int rows, int slices, const int* __restrict__ soff,
                                                  const int* __restrict__ sidx, const T* __restrict__ sval,
                                                  const int* __restrict__ perm, const T* __restrict__ in,
                                                  T* __restrict__ out) {
  const int s = blockIdx.x * (kWG / 64) + static_cast<int>(threadIdx.x >> 6);
  if (s >= slices) return;
  const int lane = threadIdx.x & 63;
  const int e0 = __builtin_amdgcn_readfirstlane(soff[s]);
  const int w = __builtin_amdgcn_readfirstlane((soff[s + 1] - e0) >> 6);
  const int* __restrict__ ip = sidx + e0 + lane;
  const T* __restrict__ vp = sval + e0 + lane;
  T a[4] = {0, 0, 0, 0};
  int k = 0;
  for (; k + 8 <= w; k += 8) {
    int c[8];
    T v[8], g[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      c[u] = ip[(k + u) * 64];
      v[u] = vp[(k + u) * 64];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) g[u] = in[c[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u & 3] += v[u] * g[u];
  }
  for (; k < w; ++k) a[k & 3] += vp[k * 64] * in[ip[k * 64]];
  const int r = s * 64 + lane;
  if (r < rows) out[perm[r]] = (a[0] + a[1]) + (a[2] + a[3]);
}

// Vector layout of the op-level kernels: element (i, b) of rows of `ld`
// lanes at i*ld + b, or, for FP64 with `split`, the run layout of the
// batched kernels (the low words of row i's ld lanes, then their high words).
struct OpLay {
  int ld;
  int split;
};

template <typename T>
__device__ inline T oget(const T* __restrict__ p, size_t i, int b, OpLay L) {
  if constexpr (sizeof(T) == 8) {
    if (L.split) {
      const unsigned* u = reinterpret_cast<const unsigned*>(p) + i * 2 * L.ld;
      return __builtin_bit_cast(T, (static_cast<unsigned long long>(u[L.ld + b]) << 32) | u[b]);
    }
  }
  return p[i * L.ld + b];
}

template <typename T>
__device__ inline void oput(T* __restrict__ p, size_t i, int b, OpLay L, T v) {
  if constexpr (sizeof(T) == 8) {
    if (L.split) {
      unsigned* u = reinterpret_cast<unsigned*>(p) + i * 2 * L.ld;
      const unsigned long long bits = __builtin_bit_cast(unsigned long long, v);
      u[b] = static_cast<unsigned>(bits);
      u[L.ld + b] = static_cast<unsigned>(bits >> 32);
      return;
    }
  }
  p[i * L.ld + b] = v;
}

// Elementwise ops over rows x ld elements (element (i, b) with b = t mod ld).
template <typename T>
__global__ void k_op_primal(long long len, OpLay L, const T* x, const T* aty, const T* c, const T* lo, const T* up,
                            const double* tau, T* xp) {
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < len;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const size_t i = static_cast<size_t>(t / L.ld);
    const int b = static_cast<int>(t % L.ld);
    const T tv = static_cast<T>(tau[b]);
    oput(xp, i, b, L,
         clamp_to<T>(oget(x, i, b, L) - tv * (oget(c, i, b, L) - oget(aty, i, b, L)), oget(lo, i, b, L),
                     oget(up, i, b, L)));
  }
}

template <typename T>
__global__ void k_op_dual(long long len, OpLay L, const T* y, const T* axp, const T* ax, const T* rl, const T* ru,
                          const double* sigma, T* yp) {
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < len;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const size_t i = static_cast<size_t>(t / L.ld);
    const int b = static_cast<int>(t % L.ld);
    const T s = static_cast<T>(sigma[b]);
    const T v = oget(y, i, b, L) - s * (static_cast<T>(2) * oget(axp, i, b, L) - oget(ax, i, b, L));
    oput(yp, i, b, L, v + s * clamp_to<T>(-v / s, oget(rl, i, b, L), oget(ru, i, b, L)));
  }
}

template <typename T>
__global__ void k_op_halpern(long long len, OpLay L, const int* iter, double rho, T* z, const T* zp, const T* z0) {
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < len;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const size_t i = static_cast<size_t>(t / L.ld);
    const int b = static_cast<int>(t % L.ld);
    const double kk = iter[b];
    const double lam = (kk + 1.0) / (kk + 2.0);
    oput(z, i, b, L,
         static_cast<T>(lam) * (static_cast<T>(1.0 + rho) * oget(zp, i, b, L) - static_cast<T>(rho) * oget(z, i, b, L)) +
             static_cast<T>(1.0 - lam) * oget(z0, i, b, L));
  }
}

__global__ void k_op_advance(int k, int* iter, const int* mask, int reset) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= k) return;
  if (reset) {
    if (mask[b]) iter[b] = 0;
  } else {
    iter[b] += 1;
  }
}

template <typename T>
__global__ void k_op_restart(long long len, OpLay L, const int* mask, const T* src, T* d1, T* d2) {
  for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < len;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const size_t i = static_cast<size_t>(t / L.ld);
    const int b = static_cast<int>(t % L.ld);
    if (mask[b]) {
      const T v = oget(src, i, b, L);
      oput(d1, i, b, L, v);
      oput(d2, i, b, L, v);
    }
  }
}

// Partial sums partial[(b * g + blk) * 3 + s] over a grid (g, k), as the
// reference's kkt_rows / kkt_cols / reduce_diff2 layout (kernels.cpp:112-192).
__device__ inline void op_block_sum3(double a, double bq, double cq, double* out) {
  __shared__ double sh[3][kWG / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double v[3] = {a, bq, cq};
  for (int q = 0; q < 3; ++q) {
    double s = v[q];
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) sh[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) {
      double t = sh[q][0];
      for (int w = 1; w < kWaves; ++w) t += sh[q][w];
      out[q] = t;
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_op_kkt_rows(int m, OpLay Y, const T* axp, const T* rl, const T* ru,
                                                     const T* yp, const double* rfac, double* partial) {
  const int b = blockIdx.y;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = blockIdx.x * kWG + threadIdx.x; i < m; i += gridDim.x * kWG) {
    const double a = oget(axp, i, b, Y), L = oget(rl, i, b, Y), U = oget(ru, i, b, Y), y = oget(yp, i, b, Y);
    const double res = (fmax(L - a, 0.0) + fmax(a - U, 0.0)) * rfac[i];
    s0 += res * res;
    if (y > 0 && L > -1e300) s1 += y * L;
    else if (y < 0 && U < 1e300) s1 += y * U;
    s2 += y * y;
  }
  op_block_sum3(s0, s1, s2, partial + (static_cast<size_t>(b) * gridDim.x + blockIdx.x) * 3);
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_op_kkt_cols(int n, OpLay X, const T* xp, const T* c, const T* atyp,
                                                     const T* lo, const T* up, const double* cfac,
                                                     double* partial) {
  const int b = blockIdx.y;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int j = blockIdx.x * kWG + threadIdx.x; j < n; j += gridDim.x * kWG) {
    const double cj = oget(c, j, b, X), l = oget(lo, j, b, X), u = oget(up, j, b, X);
    const double d = cj - static_cast<double>(oget(atyp, j, b, X));
    const bool hl = l > -1e300, hu = u < 1e300;
    double res = 0;
    if (!hl && !hu) res = fabs(d);
    else if (!hl) res = fmax(d, 0.0);
    else if (!hu) res = fmax(-d, 0.0);
    res *= cfac[j];
    s0 += res * res;
    s1 += cj * static_cast<double>(oget(xp, j, b, X));
    if (d > 0 && hl) s2 += d * l;
    else if (d < 0 && hu) s2 += d * u;
  }
  op_block_sum3(s0, s1, s2, partial + (static_cast<size_t>(b) * gridDim.x + blockIdx.x) * 3);
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_op_diff2(int len, OpLay L, const T* a, const T* bv, double* partial) {
  const int b = blockIdx.y;
  double s = 0;
  for (int i = blockIdx.x * kWG + threadIdx.x; i < len; i += gridDim.x * kWG) {
    const double d = static_cast<double>(oget(a, i, b, L)) - static_cast<double>(oget(bv, i, b, L));
    s += d * d;
  }
  op_block_sum3(s, 0.0, 0.0, partial + (static_cast<size_t>(b) * gridDim.x + blockIdx.x) * 3);
}

struct PlanHost {
  RowPlan plan{};
  std::vector<int> perm;
  int* dPerm = nullptr;
  int blocks = 0;
  // Tiles for the iterations without checks (B = 64, MapSetup::tile); the
  // check kernels keep `plan`.
  bool tiled = false;
  int tileShape = 0;  // 1..kTileShapes (tiles.h) when tiled
  RowPlan tplan{};
  TileSet tiles;
  std::vector<int> desc;  // tile_descriptors, kDescInts ints per (tile, wavefront)
  int* dLoc = nullptr;
};

// valueBytes: bytes of an iterate element (the tiles' LDS budget in columns).
PlanHost build_plan(const Csr& M, int logB, const MapSetup& ms, int valueBytes = 0) {
  PlanHost h;
  const int rows = M.rows;
  const int maxLogS = 6 - logB;  // B * S <= 64 lanes (one wavefront per row at most)
  auto logs_of = [&](int len) {
    const int per = std::max(1, ms.nnzPerLane);
    int want = (len + per - 1) / per;
    int l = ilog2(std::max(1, want));
    return std::min(l, maxLogS);
  };
  int nb = 0;
  if (ms.group && (logB == 4 || logB == 5)) {
    // One row per B-lane group; rowMap 1 orders the rows by length so the
    // rows sharing a wavefront have similar lengths.
    h.plan.logS[0] = 0;
    h.plan.rowStart[0] = 0;
    h.plan.rowCount[0] = rows;
    h.plan.grp = 1;
    if (ms.rowMap == 1) {
      h.perm.resize(rows);
      for (int i = 0; i < rows; ++i) h.perm[i] = i;
      std::stable_sort(h.perm.begin(), h.perm.end(), [&](int a, int b) {
        return M.ptr[a + 1] - M.ptr[a] > M.ptr[b + 1] - M.ptr[b];
      });
    }
    nb = 1;
  } else if (ms.rowMap == 2 || maxLogS == 0) {
    // One mapping: S from the mean row length.
    const double mean = rows > 0 ? static_cast<double>(M.nnz()) / rows : 0;
    const int l = logs_of(static_cast<int>(std::ceil(mean)));
    h.plan.logS[0] = l;
    h.plan.rowStart[0] = 0;
    h.plan.rowCount[0] = rows;
    nb = 1;
  } else {
    std::vector<int> cnt(maxLogS + 1, 0), bucket(rows);
    for (int i = 0; i < rows; ++i) {
      bucket[i] = logs_of(M.ptr[i + 1] - M.ptr[i]);
      cnt[bucket[i]]++;
    }
    std::vector<int> start(maxLogS + 2, 0);
    for (int l = 0; l <= maxLogS; ++l) start[l + 1] = start[l] + cnt[l];
    h.perm.resize(rows);
    std::vector<int> pos(start.begin(), start.end() - 1);
    for (int i = 0; i < rows; ++i) h.perm[pos[bucket[i]]++] = i;
    for (int l = 0; l <= maxLogS; ++l) {
      if (cnt[l] == 0) continue;
      h.plan.logS[nb] = l;
      h.plan.rowStart[nb] = start[l];
      h.plan.rowCount[nb] = cnt[l];
      ++nb;
    }
    if (nb == 1) {  // a single bucket keeps the natural order
      h.perm.clear();
      h.plan.rowStart[0] = 0;
    }
    if (nb == 0) {
      h.plan.logS[0] = 0;
      h.plan.rowStart[0] = 0;
      h.plan.rowCount[0] = 0;
      nb = 1;
    }
  }
  h.plan.nb = nb;
  h.plan.rpw = logB == 6 && ms.rowsPerWave == 2 ? 2 : 1;
  int blk = 0;
  for (int q = 0; q < nb; ++q) {
    h.plan.blkStart[q] = blk;
    const int rowsPerWg = (kWG >> (logB + h.plan.logS[q])) * h.plan.rpw;
    blk += (h.plan.rowCount[q] + rowsPerWg - 1) / rowsPerWg;
  }
  h.plan.blkStart[nb] = blk;
  h.blocks = std::max(blk, 1);
  h.plan.xcd = ms.xcd ? 1 : 0;
  h.plan.perm = nullptr;
  const int shape = ms.tile, waves = tile_waves(shape);
  const int tileCols = shape > 0 && valueBytes > 0 ? tile_bytes(shape) / (64 * valueBytes) : 0;
  if (shape > 0 && logB == 6 && tileCols > 0 && build_tiles(M, tileCols, h.tiles, waves)) {
    h.tiled = true;
    h.tileShape = shape;
    h.tplan = h.plan;
    h.tplan.xcd = ms.xcd ? 1 : 0;
    h.desc = tile_descriptors(M, h.tiles, (tileCols + waves - 1) / waves, waves);
  }
  return h;
}

std::mutex gInfoMutex;

}  // namespace

// ---------------------------------------------------------------------------
// Device probe (gpuAvailable / gpuInfoJson)
// ---------------------------------------------------------------------------
const DeviceInfo& device_info(int device) {
  static DeviceInfo cache[kMaxDevices];
  static bool probed[kMaxDevices] = {};
  static const DeviceInfo outOfRange = [] {
    DeviceInfo d;
    d.ordinal = -1;
    d.reason = "device ordinal outside 0.." + std::to_string(kMaxDevices - 1);
    return d;
  }();
  if (device < 0 || device >= kMaxDevices) return outOfRange;
  std::lock_guard<std::mutex> lk(gInfoMutex);
  DeviceInfo& d = cache[device];
  if (probed[device]) return d;
  probed[device] = 1;
  d.ordinal = device;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= device) {
    d.reason = e != hipSuccess ? std::string("hipGetDeviceCount: ") + hipGetErrorString(e) : "no such HIP device";
    return d;
  }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) {
    d.reason = "hipGetDeviceProperties failed";
    return d;
  }
  d.arch = p.gcnArchName;
  d.name = p.name[0] ? std::string(p.name) : "AMD " + d.arch.substr(0, d.arch.find(':'));
  d.computeUnits = p.multiProcessorCount;
  d.memoryBytes = p.totalGlobalMem;
  // memoryClockRate (kHz) x bus width x 2 is the DDR formula; it reads half
  // of HBM3E's rate on gfx950 (4.1 TB/s), so that part reports its published
  // peak (MI350 series, 8 stacks: 8.0 TB/s).
  d.peakBandwidthGBs = 2.0 * p.memoryClockRate * 1e3 * (p.memoryBusWidth / 8.0) / 1e9;
  if (d.arch.rfind("gfx950", 0) == 0) d.peakBandwidthGBs = 8000.0;
  if (d.arch.rfind("gfx950", 0) != 0) {
    d.reason = "device " + d.arch + " is not gfx950 (this build carries gfx950 code objects only)";
    return d;
  }
  d.available = true;
  return d;
}

// ---------------------------------------------------------------------------
// DeviceRun implementation
// ---------------------------------------------------------------------------
namespace {

template <typename T>
struct Kern {
  using ColFn = void (*)(RowPlan, const int*, const int*, const T*, const T*, const T*, T*, T*, T*, const T*, Ld,
                         const T*, const T*, Ld, Inst, int, double, double*, int);
  using RowFn = void (*)(RowPlan, const int*, const int*, const T*, const double*, const T*, T*, T*, T*, T*, T*, T*,
                         const T*, const T*, const double*, const double*, Ld, const double*, Inst, int, double,
                         double*, int);
  using ColCheckFn = void (*)(RowPlan, const int*, const int*, const double*, const T*, const T*, const double*, Ld,
                              const double*, const double*, Ld, const double*, Inst, double*, int);
  using SpmvFn = void (*)(RowPlan, const int*, const int*, const T*, const T*, T*);
  // [rpw - 1]: one row per wavefront, or (B = 64) two with their chunks interleaved
  ColFn col[2], colChk[2];
  RowFn row[2], rowChk[2];
  ColCheckFn colcheck[2];
  SpmvFn spmv[2];  // the run's vector layout
  SpmvFn spmvStd;  // element (i, b) at i*B + b (op-level backend)
  using ColTiledFn = void (*)(RowPlan, const int*, const T*, const T*, const T*, T*, T*, T*, const T*, Ld, const T*,
                              const T*, Ld, Inst, int, double);
  using RowTiledFn = void (*)(RowPlan, const int*, const T*, const T*, T*, T*, T*, T*, T*, T*, const T*, const T*,
                              const double*, const double*, Ld, const double*, Inst, int, double);
  ColTiledFn colTiled[kTileShapes] = {};  // B = 64 tiles, by shape - 1
  RowTiledFn rowTiled[kTileShapes] = {};
  using SpmvTiledFn = void (*)(RowPlan, const int*, const T*, const T*, T*);
  SpmvTiledFn spmvTiled[kTileShapes] = {};  // op-level SpMV, B = 64 tiles
};

template <typename T, int LOGB>
Kern<T> kern_for() {
  Kern<T> k;
  constexpr int R2 = LOGB == 6 ? 2 : 1;
  k.col[0] = k_col<T, LOGB, false>;
  k.colChk[0] = k_col<T, LOGB, true>;
  k.row[0] = k_row<T, LOGB, false>;
  k.rowChk[0] = k_row<T, LOGB, true>;
  k.colcheck[0] = k_colcheck<T, LOGB>;
  k.spmv[0] = k_spmv<T, LOGB, kSplit<T, LOGB>>;
  k.col[1] = k_col<T, LOGB, false, R2>;
  k.colChk[1] = k_col<T, LOGB, true, R2>;
  k.row[1] = k_row<T, LOGB, false, R2>;
  k.rowChk[1] = k_row<T, LOGB, true, R2>;
  k.colcheck[1] = k_colcheck<T, LOGB, R2>;
  k.spmv[1] = k_spmv<T, LOGB, kSplit<T, LOGB>, R2>;
  k.spmvStd = k_spmv<T, LOGB, false>;
  if constexpr (LOGB == 6) {
    k.colTiled[0] = k_col_tiled<T, 1>;
    k.rowTiled[0] = k_row_tiled<T, 1>;
    k.colTiled[1] = k_col_tiled<T, 2>;
    k.rowTiled[1] = k_row_tiled<T, 2>;
    k.colTiled[2] = k_col_tiled<T, 3>;
    k.rowTiled[2] = k_row_tiled<T, 3>;
    k.spmvTiled[0] = k_spmv_tiled<T, 1>;
    k.spmvTiled[1] = k_spmv_tiled<T, 2>;
    k.spmvTiled[2] = k_spmv_tiled<T, 3>;
  }
  return k;
}

template <typename T>
Kern<T> kern_table(int logB) {
  switch (logB) {
    case 0: return kern_for<T, 0>();
    case 1: return kern_for<T, 1>();
    case 2: return kern_for<T, 2>();
    case 3: return kern_for<T, 3>();
    case 4: return kern_for<T, 4>();
    case 5: return kern_for<T, 5>();
    case 6: return kern_for<T, 6>();
  }
  throw Error("batch width must be a power of two <= 64");
}

template <typename T>
class RunImpl final : public DeviceRun {
 public:
  RunImpl(int device, const Prepared& P, const RunSetup& s) : s_(s), dev_(device) {
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&ev0_));
    HIP_OK(hipEventCreate(&ev1_));
    m_ = P.m;
    n_ = P.n;
    B_ = s.B;
    logB_ = ilog2(B_);
    PDHG_REQUIRE((1 << logB_) == B_ && B_ <= 64, "batch width must be a power of two <= 64");
    split_ = sizeof(T) == 8 && B_ >= 16;  // kSplit
    k_ = kern_table<T>(logB_);
    nnz_ = P.A.nnz();
    // Matrices: FP32 (or FP64) values for the iterations, FP64 for the checks.
    aPtr_ = upload_int(P.A.ptr);
    aIdx_ = upload_int(csr_padded(P.A.idx));
    tPtr_ = upload_int(P.AT.ptr);
    tIdx_ = upload_int(csr_padded(P.AT.idx));
    aVal64_ = upload_dbl(csr_padded(P.A.val));
    tVal64_ = upload_dbl(csr_padded(P.AT.val));
    if (sizeof(T) == 8) {
      aVal_ = reinterpret_cast<T*>(aVal64_);
      tVal_ = reinterpret_cast<T*>(tVal64_);
    } else {
      aVal_ = upload_conv(csr_padded(P.A.val));
      tVal_ = upload_conv(csr_padded(P.AT.val));
    }
    rfac_ = upload_dbl(P.rfac);
    cfac_ = upload_dbl(P.cfac);
    colPlan_ = build_plan(P.AT, logB_, s.colMap, static_cast<int>(sizeof(T)));
    rowPlan_ = build_plan(P.A, logB_, s.rowMap, static_cast<int>(sizeof(T)));
    for (PlanHost* h : {&colPlan_, &rowPlan_}) upload_plan(*h);
    // Iterates (per instance) and case data (per instance or shared).
    const size_t nB = static_cast<size_t>(n_) * B_, mB = static_cast<size_t>(m_) * B_;
    lc_ = s.sharedCost ? Ld{1, 0} : Ld{B_, B_ - 1};
    lb_ = s.sharedColBounds ? Ld{1, 0} : Ld{B_, B_ - 1};
    lr_ = s.sharedRowBounds ? Ld{1, 0} : Ld{B_, B_ - 1};
    const size_t nc = s.sharedCost ? n_ : nB, nbnd = s.sharedColBounds ? n_ : nB, mr = s.sharedRowBounds ? m_ : mB;
    x_ = buf_.alloc<T>(nB);
    xp_ = buf_.alloc<T>(nB);
    x0_ = buf_.alloc<T>(nB);
    y_ = buf_.alloc<T>(mB);
    yp_ = buf_.alloc<T>(mB);
    y0_ = buf_.alloc<T>(mB);
    ax_ = buf_.alloc<T>(mB);
    axp_ = buf_.alloc<T>(mB);
    ax0_ = buf_.alloc<T>(mB);
    c_ = buf_.alloc<T>(nc);
    lo_ = buf_.alloc<T>(nbnd);
    up_ = buf_.alloc<T>(nbnd);
    rl_ = buf_.alloc<T>(mr);
    ru_ = buf_.alloc<T>(mr);
    c64_ = buf_.alloc<double>(nc);
    lo64_ = buf_.alloc<double>(nbnd);
    up64_ = buf_.alloc<double>(nbnd);
    rl64_ = buf_.alloc<double>(mr);
    ru64_ = buf_.alloc<double>(mr);
    nc_ = nc;
    nbnd_ = nbnd;
    mr_ = mr;
    colPart_ = buf_.alloc<double>(static_cast<size_t>(5) * B_ * colPlan_.blocks);
    rowPart_ = buf_.alloc<double>(static_cast<size_t>(4) * B_ * rowPlan_.blocks);
    // Instance state.
    st_.tau = buf_.alloc<double>(B_);
    st_.sigma = buf_.alloc<double>(B_);
    st_.omega = buf_.alloc<double>(B_);
    st_.fpR = buf_.alloc<double>(B_);
    st_.fpPrev = buf_.alloc<double>(B_);
    st_.iter = buf_.alloc<long long>(B_);
    st_.since = buf_.alloc<long long>(B_);
    st_.restartAt = buf_.alloc<long long>(B_);
    st_.applyAt = buf_.alloc<long long>(B_);
    st_.doneIter = buf_.alloc<long long>(B_);
    st_.done = buf_.alloc<int>(B_);
    st_.restarts = buf_.alloc<int>(B_);
    cNorm_ = buf_.alloc<double>(B_);
    rhsNorm_ = buf_.alloc<double>(B_);
    st_.cNorm = cNorm_;
    st_.rhsNorm = rhsNorm_;
    st_.rec = buf_.alloc<CheckRecord>(B_);
    HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&hostRec_), B_ * (sizeof(CheckRecord) + 2 * sizeof(int) + sizeof(long long)),
                         hipHostMallocDefault));
  }
  ~RunImpl() override {
    if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (hostRec_) (void)hipHostFree(hostRec_);
    (void)hipEventDestroy(ev0_);
    (void)hipEventDestroy(ev1_);
    (void)hipStreamDestroy(stream_);
  }

  void upload(const double* c, const double* lo, const double* up, const double* rl, const double* ru, const double* x,
              const double* y) override {
    const size_t nB = static_cast<size_t>(n_) * B_, mB = static_cast<size_t>(m_) * B_;
    put_both(c_, c64_, c, nc_, lc_.mask != 0);
    put_both(lo_, lo64_, lo, nbnd_, lb_.mask != 0);
    put_both(up_, up64_, up, nbnd_, lb_.mask != 0);
    put_both(rl_, rl64_, rl, mr_, lr_.mask != 0);
    put_both(ru_, ru64_, ru, mr_, lr_.mask != 0);
    put(x_, x, nB, true);
    put(x0_, x, nB, true);
    put(xp_, x, nB, true);
    put(y_, y, mB, true);
    put(y0_, y, mB, true);
    put(yp_, y, mB, true);
    HIP_OK(hipStreamSynchronize(stream_));
  }

  void init_state(const RunScalars& sc, const double* omega, const double* cNorm, const double* rhsNorm,
                  const int* done) override {
    sc_ = sc;
    std::vector<double> tau(B_), sigma(B_), inf(B_, INFINITY);
    for (int b = 0; b < B_; ++b) {
      tau[b] = sc.eta / omega[b];
      sigma[b] = sc.eta * omega[b];
    }
    std::vector<long long> zero(B_, 0), neg(B_, -1);
    std::vector<int> zi(B_, 0);
    copy_h2d(st_.tau, tau.data(), B_);
    copy_h2d(st_.sigma, sigma.data(), B_);
    copy_h2d(st_.omega, omega, B_);
    copy_h2d(st_.fpR, inf.data(), B_);
    copy_h2d(st_.fpPrev, inf.data(), B_);
    copy_h2d(st_.iter, zero.data(), B_);
    copy_h2d(st_.since, zero.data(), B_);
    copy_h2d(st_.restartAt, zero.data(), B_);
    copy_h2d(st_.applyAt, neg.data(), B_);
    copy_h2d(st_.doneIter, neg.data(), B_);
    copy_h2d(st_.done, done, B_);
    copy_h2d(st_.restarts, zi.data(), B_);
    copy_h2d(cNorm_, cNorm, B_);
    copy_h2d(rhsNorm_, rhsNorm, B_);
    HIP_OK(hipStreamSynchronize(stream_));
  }

  void start(bool initAx) override {
    if (initAx) {
      launch_spmv(false, x_, ax_);
      HIP_OK(hipMemcpyAsync(ax0_, ax_, sizeof(T) * m_ * static_cast<size_t>(B_), hipMemcpyDeviceToDevice, stream_));
      HIP_OK(hipMemcpyAsync(axp_, ax_, sizeof(T) * m_ * static_cast<size_t>(B_), hipMemcpyDeviceToDevice, stream_));
    }
    HIP_OK(hipEventRecord(ev0_, stream_));
  }

  void run_block(int iters) override {
    PDHG_REQUIRE(iters >= 1, "empty block");
    if (s_.useGraph && iters == s_.checkEvery) {
      if (!graphExec_) capture(iters);
      HIP_OK(hipGraphLaunch(graphExec_, stream_));
      graphLaunches++;
      launches += 2 * iters + 2;
    } else {
      enqueue(iters);
    }
  }

  void fetch(CheckRecord* rec, int* done, long long* doneIter, int* restarts) override {
    char* h = reinterpret_cast<char*>(hostRec_);
    CheckRecord* hr = hostRec_;
    int* hd = reinterpret_cast<int*>(h + B_ * sizeof(CheckRecord));
    int* hrs = hd + B_;
    long long* hdi = reinterpret_cast<long long*>(hrs + B_);
    HIP_OK(hipMemcpyAsync(hr, st_.rec, B_ * sizeof(CheckRecord), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipMemcpyAsync(hd, st_.done, B_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipMemcpyAsync(hrs, st_.restarts, B_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipMemcpyAsync(hdi, st_.doneIter, B_ * sizeof(long long), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    std::memcpy(rec, hr, B_ * sizeof(CheckRecord));
    std::memcpy(done, hd, B_ * sizeof(int));
    std::memcpy(restarts, hrs, B_ * sizeof(int));
    std::memcpy(doneIter, hdi, B_ * sizeof(long long));
    bytesDown += B_ * (sizeof(CheckRecord) + 2 * sizeof(int) + sizeof(long long));
  }

  void download(int which, double* out) override {
    const size_t len = static_cast<size_t>(which == 0 ? n_ : m_) * B_;
    std::vector<T> tmp(len);
    HIP_OK(hipMemcpyAsync(tmp.data(), which == 0 ? xp_ : yp_, len * sizeof(T), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    from_device_layout(tmp);
    for (size_t t = 0; t < len; ++t) out[t] = static_cast<double>(tmp[t]);
    bytesDown += len * sizeof(T);
  }

  double stop() override {
    HIP_OK(hipEventRecord(ev1_, stream_));
    HIP_OK(hipEventSynchronize(ev1_));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, ev0_, ev1_));
    return ms / 1000.0;
  }

  void time_kernels(int reps, double* msCol, double* msRow, double* msIter) override {
    // Called after the results were read: every instance is released (no
    // done flag, no pending restart) so each launch streams and writes its
    // full data.  Each kernel alone over `reps` back-to-back launches, then
    // `reps` iterations (column and row kernel in turn), HIP events at the
    // ends of each series only.
    std::vector<int> zeros(B_, 0);
    std::vector<long long> none(B_, -1);
    copy_h2d(st_.done, zeros.data(), B_);
    copy_h2d(st_.applyAt, none.data(), B_);
    launch_col(false, 0);
    launch_row(false, 0);
    auto series = [&](bool col, bool row) {
      HIP_OK(hipEventRecord(ev0_, stream_));
      for (int r = 0; r < reps; ++r) {
        if (col) launch_col(false, 0);
        if (row) launch_row(false, 0);
      }
      HIP_OK(hipEventRecord(ev1_, stream_));
      HIP_OK(hipEventSynchronize(ev1_));
      float ms = 0;
      HIP_OK(hipEventElapsedTime(&ms, ev0_, ev1_));
      return static_cast<double>(ms) / reps;
    };
    *msCol = series(true, false);
    *msRow = series(false, true);
    *msIter = series(true, true);
  }

  void kernel_bytes(double* colBytes, double* rowBytes) const override {
    // Algorithmic bytes: matrix (value + column index + row pointer), one
    // compulsory read of the gathered vector, and every streamed vector once.
    const double vs = sizeof(T), nB = static_cast<double>(n_) * B_, mB = static_cast<double>(m_) * B_;
    const double mat = (vs + 4.0) * nnz_;
    const double nc = static_cast<double>(nc_), nb = static_cast<double>(nbnd_), mr = static_cast<double>(mr_);
    *colBytes = mat + 4.0 * (n_ + 1) + vs * mB + vs * (2 * nB /*x, x0*/ + nc + 2 * nb) + vs * 2 * nB /*x, xp*/;
    *rowBytes = mat + 4.0 * (m_ + 1) + vs * nB + vs * (4 * mB /*y, y0, ax, ax0*/ + 2 * mr) + vs * 2 * mB /*y, ax*/;
  }

  void export_lanes(const std::vector<int>& lanes, LaneState& o) override {
    const int k = static_cast<int>(lanes.size());
    o.count = k;
    auto pick = [&](const T* d, int rows, std::vector<double>& out) {
      std::vector<T> full(static_cast<size_t>(rows) * B_);
      HIP_OK(hipMemcpyAsync(full.data(), d, full.size() * sizeof(T), hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      from_device_layout(full);
      bytesDown += full.size() * sizeof(T);
      out.resize(static_cast<size_t>(rows) * k);
      for (int i = 0; i < rows; ++i)
        for (int q = 0; q < k; ++q)
          out[static_cast<size_t>(i) * k + q] = full[static_cast<size_t>(i) * B_ + lanes[q]];
    };
    pick(x_, n_, o.x);
    pick(x0_, n_, o.x0);
    pick(xp_, n_, o.xp);
    pick(y_, m_, o.y);
    pick(y0_, m_, o.y0);
    pick(yp_, m_, o.yp);
    pick(ax_, m_, o.ax);
    pick(ax0_, m_, o.ax0);
    pick(axp_, m_, o.axp);
    auto sel = [&](const auto* d, auto& out) {
      using U = std::remove_const_t<std::remove_pointer_t<decltype(d)>>;
      std::vector<U> full(B_);
      HIP_OK(hipMemcpyAsync(full.data(), d, B_ * sizeof(U), hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      out.resize(k);
      for (int q = 0; q < k; ++q) out[q] = full[lanes[q]];
    };
    sel(st_.tau, o.tau);
    sel(st_.sigma, o.sigma);
    sel(st_.omega, o.omega);
    sel(st_.fpR, o.fpR);
    sel(st_.fpPrev, o.fpPrev);
    sel(st_.iter, o.iter);
    sel(st_.since, o.since);
    sel(st_.restartAt, o.restartAt);
    sel(st_.applyAt, o.applyAt);
    sel(st_.restarts, o.restarts);
  }

  void load_lanes(const std::vector<int>& lanes, const double* c, const double* lo, const double* up,
                  const double* rl, const double* ru, const double* x, const double* y, const double* omega,
                  const double* cNorm, const double* rhsNorm) override {
    const int K = static_cast<int>(lanes.size());
    if (K == 0) return;
    for (int b : lanes) PDHG_REQUIRE(b >= 0 && b < B_, "load_lanes: lane out of range");
    const size_t maxLen = static_cast<size_t>(std::max(n_, m_));
    if (!stageT_) {
      stageT_ = buf_.alloc<T>(maxLen * B_);
      stage64_ = buf_.alloc<double>(maxLen * B_);
      laneIdx_ = buf_.alloc<int>(B_);
      axScratch_ = buf_.alloc<T>(static_cast<size_t>(m_) * B_);
    }
    copy_h2d(laneIdx_, lanes.data(), K);
    const int split = split_ ? 1 : 0;
    auto grid = [&](size_t total) { return dim3(static_cast<unsigned>(std::min<size_t>(4096, (total + 255) / 256))); };
    // Iteration copy (T, run layout) of a K-interleaved host array.
    auto scatterT = [&](T* const* dsts, int ndst, const double* h, int len) {
      std::vector<T> t(h, h + static_cast<size_t>(len) * K);
      copy_h2d(stageT_, t.data(), t.size());
      for (int d = 0; d < ndst; ++d)
        hipLaunchKernelGGL(k_scatter_lanes<T>, grid(t.size()), dim3(256), 0, stream_, dsts[d], stageT_, laneIdx_,
                           static_cast<long long>(len), K, logB_, split);
    };
    // FP64 copy (element order) read by the check kernels.
    auto scatter64 = [&](double* dst, const double* h, int len) {
      copy_h2d(stage64_, h, static_cast<size_t>(len) * K);
      hipLaunchKernelGGL(k_scatter_lanes<double>, grid(static_cast<size_t>(len) * K), dim3(256), 0, stream_, dst,
                         stage64_, laneIdx_, static_cast<long long>(len), K, logB_, 0);
    };
    T* xs[3] = {x_, x0_, xp_};
    scatterT(xs, 3, x, n_);
    T* ys[3] = {y_, y0_, yp_};
    scatterT(ys, 3, y, m_);
    if (c) {
      scatterT(&c_, 1, c, n_);
      scatter64(c64_, c, n_);
    }
    if (lo) {
      scatterT(&lo_, 1, lo, n_);
      scatterT(&up_, 1, up, n_);
      scatter64(lo64_, lo, n_);
      scatter64(up64_, up, n_);
    }
    if (rl) {
      scatterT(&rl_, 1, rl, m_);
      scatterT(&ru_, 1, ru, m_);
      scatter64(rl64_, rl, m_);
      scatter64(ru64_, ru, m_);
    }
    // ax = A x for the new lanes (all lanes computed into scratch, only the
    // new lanes copied, so the live lanes keep their own ax).
    launch_spmv(false, x_, axScratch_);
    for (T* d : {ax_, ax0_, axp_})
      hipLaunchKernelGGL(k_copy_lanes<T>, grid(static_cast<size_t>(m_) * K), dim3(256), 0, stream_, d, axScratch_,
                         laneIdx_, static_cast<long long>(m_), K, logB_, split);
    HIP_OK(hipGetLastError());
    // Per-lane restart state (src/pdhg/pdhg.cpp:506-545 for a fresh case).
    auto put_lanes = [&](auto* d, auto value_of) {
      using U = std::remove_pointer_t<decltype(d)>;
      std::vector<U> full(B_);
      HIP_OK(hipMemcpyAsync(full.data(), d, B_ * sizeof(U), hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      for (int q = 0; q < K; ++q) full[lanes[q]] = static_cast<U>(value_of(q));
      copy_h2d(d, full.data(), B_);
    };
    put_lanes(st_.tau, [&](int q) { return sc_.eta / omega[q]; });
    put_lanes(st_.sigma, [&](int q) { return sc_.eta * omega[q]; });
    put_lanes(st_.omega, [&](int q) { return omega[q]; });
    put_lanes(st_.fpR, [&](int) { return INFINITY; });
    put_lanes(st_.fpPrev, [&](int) { return INFINITY; });
    put_lanes(st_.iter, [&](int) { return 0; });
    put_lanes(st_.since, [&](int) { return 0; });
    put_lanes(st_.restartAt, [&](int) { return 0; });
    put_lanes(st_.applyAt, [&](int) { return -1; });
    put_lanes(st_.doneIter, [&](int) { return -1; });
    put_lanes(st_.restarts, [&](int) { return 0; });
    put_lanes(cNorm_, [&](int q) { return cNorm[q]; });
    put_lanes(rhsNorm_, [&](int q) { return rhsNorm[q]; });
    put_lanes(st_.done, [&](int) { return 0; });
  }

  void retire_lanes(const std::vector<int>& lanes) override {
    if (lanes.empty()) return;
    std::vector<int> full(B_);
    HIP_OK(hipMemcpyAsync(full.data(), st_.done, B_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    for (int b : lanes) full[b] = 1;
    copy_h2d(st_.done, full.data(), B_);
  }

  void autotune(const Prepared& P, const std::vector<MapSetup>& cand, int reps, int* best, double* msIter,
                double* msIterHalf) override {
    PDHG_REQUIRE(!cand.empty() && reps >= 1, "autotune: no candidates");
    // Scratch state: zero vectors, every lane live, no restart pending.
    const size_t nB = static_cast<size_t>(n_) * B_, mB = static_cast<size_t>(m_) * B_;
    for (T* v : {x_, xp_, x0_}) HIP_OK(hipMemsetAsync(v, 0, nB * sizeof(T), stream_));
    for (T* v : {y_, yp_, y0_, ax_, axp_, ax0_}) HIP_OK(hipMemsetAsync(v, 0, mB * sizeof(T), stream_));
    HIP_OK(hipMemsetAsync(c_, 0, nc_ * sizeof(T), stream_));
    for (T* v : {lo_, up_}) HIP_OK(hipMemsetAsync(v, 0, nbnd_ * sizeof(T), stream_));
    for (T* v : {rl_, ru_}) HIP_OK(hipMemsetAsync(v, 0, mr_ * sizeof(T), stream_));
    std::vector<int> zeros(B_, 0);
    std::vector<long long> zl(B_, 0), none(B_, -1);
    std::vector<double> step(B_, 1e-3);
    copy_h2d(st_.done, zeros.data(), B_);
    copy_h2d(st_.applyAt, none.data(), B_);
    copy_h2d(st_.iter, zl.data(), B_);
    copy_h2d(st_.restartAt, zl.data(), B_);
    copy_h2d(st_.tau, step.data(), B_);
    copy_h2d(st_.sigma, step.data(), B_);
    // Every candidate is timed inside iterations, column and row kernel in
    // turn, since a kernel's time depends on what ran before it: the column
    // kernel's candidates with the run's row mapping, then the row kernel's
    // with the chosen column mapping.
    auto iterations = [&] {
      launch_col(false, 0);
      launch_row(false, 0);
      HIP_OK(hipEventRecord(ev0_, stream_));
      for (int r = 0; r < reps; ++r) {
        launch_col(false, 0);
        launch_row(false, 0);
      }
      HIP_OK(hipEventRecord(ev1_, stream_));
      HIP_OK(hipEventSynchronize(ev1_));
      HIP_OK(hipGetLastError());
      float el = 0;
      HIP_OK(hipEventElapsedTime(&el, ev0_, ev1_));
      return static_cast<double>(el) / reps;
    };
    int* scratch = buf_.alloc<int>(static_cast<size_t>(std::max(m_, n_)));
    for (int dir = 0; dir < 2; ++dir) {
      PlanHost& cur = dir == 0 ? colPlan_ : rowPlan_;
      const Csr& M = dir == 0 ? P.AT : P.A;
      double bestMs = 0;
      int bestI = -1;
      for (size_t i = 0; i < cand.size(); ++i) {
        PlanHost h = build_plan(M, logB_, cand[i], static_cast<int>(sizeof(T)));
        if (h.tiled) {
          upload_plan(h);  // tiles: their own arrays, kept with the run's buffers
        } else {
          if (!h.perm.empty()) copy_h2d(scratch, h.perm.data(), h.perm.size());
          h.plan.perm = h.perm.empty() ? nullptr : scratch;
        }
        const PlanHost keep = cur;
        cur = h;
        const double t = iterations();
        cur = keep;
        if (bestI < 0 || t < bestMs) {
          bestMs = t;
          bestI = static_cast<int>(i);
        }
      }
      // Install the winner with its own permutation and partial buffers.
      PlanHost h = build_plan(M, logB_, cand[bestI], static_cast<int>(sizeof(T)));
      upload_plan(h);
      cur = h;
      best[dir] = bestI;
    }
    *msIter = iterations();
    // The chosen mappings again with lanes B/2..B-1 done.
    *msIterHalf = *msIter;
    if (B_ >= 2) {
      std::vector<int> upper(B_, 0);
      for (int b = B_ / 2; b < B_; ++b) upper[b] = 1;
      copy_h2d(st_.done, upper.data(), B_);
      *msIterHalf = iterations();
      copy_h2d(st_.done, zeros.data(), B_);
    }
    colPart_ = buf_.alloc<double>(static_cast<size_t>(5) * B_ * colPlan_.blocks);
    rowPart_ = buf_.alloc<double>(static_cast<size_t>(4) * B_ * rowPlan_.blocks);
    launches = 0;
  }

  int tiled(int side) const override { return side == 0 ? colPlan_.tileShape : rowPlan_.tileShape; }

  void import_lanes(const LaneState& in) override {
    const int k = in.count;
    PDHG_REQUIRE(k <= B_, "import: more lanes than the run holds");
    auto place = [&](T* d, int rows, const std::vector<double>& src) {
      std::vector<T> full(static_cast<size_t>(rows) * B_);
      HIP_OK(hipMemcpyAsync(full.data(), d, full.size() * sizeof(T), hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      from_device_layout(full);
      for (int i = 0; i < rows; ++i)
        for (int q = 0; q < k; ++q)
          full[static_cast<size_t>(i) * B_ + q] = static_cast<T>(src[static_cast<size_t>(i) * k + q]);
      to_device_layout(full);
      copy_h2d(d, full.data(), full.size());
    };
    place(x_, n_, in.x);
    place(x0_, n_, in.x0);
    place(xp_, n_, in.xp);
    place(y_, m_, in.y);
    place(y0_, m_, in.y0);
    place(yp_, m_, in.yp);
    place(ax_, m_, in.ax);
    place(ax0_, m_, in.ax0);
    place(axp_, m_, in.axp);
    auto put_state = [&](auto* d, const auto& src) {
      using U = std::remove_pointer_t<decltype(d)>;
      std::vector<U> full(B_);
      HIP_OK(hipMemcpyAsync(full.data(), d, B_ * sizeof(U), hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      for (int q = 0; q < k; ++q) full[q] = static_cast<U>(src[q]);
      copy_h2d(d, full.data(), B_);
    };
    put_state(st_.tau, in.tau);
    put_state(st_.sigma, in.sigma);
    put_state(st_.omega, in.omega);
    put_state(st_.fpR, in.fpR);
    put_state(st_.fpPrev, in.fpPrev);
    put_state(st_.iter, in.iter);
    put_state(st_.since, in.since);
    put_state(st_.restartAt, in.restartAt);
    put_state(st_.applyAt, in.applyAt);
    put_state(st_.restarts, in.restarts);
  }

 private:
  template <typename U>
  void copy_h2d(U* d, const U* h, size_t count) {
    HIP_OK(hipMemcpyAsync(d, h, count * sizeof(U), hipMemcpyHostToDevice, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    bytesUp += count * sizeof(U);
  }
  int* upload_int(const std::vector<int>& v) {
    int* d = buf_.alloc<int>(v.size());
    copy_h2d(d, v.data(), v.size());
    return d;
  }
  // Device copies of a plan's permutation and tiles.
  void upload_plan(PlanHost& h) {
    if (!h.perm.empty()) h.dPerm = upload_int(h.perm);
    h.plan.perm = h.dPerm;
    if (h.tiled) {
      h.tplan.perm = nullptr;
      h.tplan.desc = upload_int(h.desc);
      h.dLoc = upload_int(h.tiles.loc);
    }
  }
  double* upload_dbl(const std::vector<double>& v) {
    double* d = buf_.alloc<double>(v.size());
    copy_h2d(d, v.data(), v.size());
    return d;
  }
  T* upload_conv(const std::vector<double>& v) {
    std::vector<T> t(v.begin(), v.end());
    T* d = buf_.alloc<T>(v.size());
    copy_h2d(d, t.data(), t.size());
    return d;
  }
  // Interleaved rows x B vectors between element order (i*B + b) and the
  // device layout (the split layout for FP64 at B >= 16, kSplit).
  void to_device_layout(std::vector<T>& v) const {
    if (!split_) return;
    std::vector<uint32_t> w(2 * static_cast<size_t>(B_));
    for (size_t r = 0; r + B_ <= v.size(); r += B_) {
      for (int b = 0; b < B_; ++b) {
        uint64_t bits;
        std::memcpy(&bits, &v[r + b], 8);
        w[b] = static_cast<uint32_t>(bits);
        w[B_ + b] = static_cast<uint32_t>(bits >> 32);
      }
      std::memcpy(&v[r], w.data(), w.size() * sizeof(uint32_t));
    }
  }
  void from_device_layout(std::vector<T>& v) const {
    if (!split_) return;
    std::vector<uint32_t> w(2 * static_cast<size_t>(B_));
    for (size_t r = 0; r + B_ <= v.size(); r += B_) {
      std::memcpy(w.data(), &v[r], w.size() * sizeof(uint32_t));
      for (int b = 0; b < B_; ++b) {
        const uint64_t bits = (static_cast<uint64_t>(w[B_ + b]) << 32) | w[b];
        std::memcpy(&v[r + b], &bits, 8);
      }
    }
  }
  void put(T* d, const double* h, size_t count, bool interleaved) {
    std::vector<T> t(h, h + count);
    if (interleaved) to_device_layout(t);
    copy_h2d(d, t.data(), count);
  }
  // Case data: the iteration copy in the run's layout, the FP64 copy (read by
  // the check kernels) in element order.
  void put_both(T* d, double* d64, const double* h, size_t count, bool interleaved) {
    put(d, h, count, interleaved);
    copy_h2d(d64, h, count);
  }

  void launch_spmv(bool transpose, const T* in, T* out) {
    const PlanHost& p = transpose ? colPlan_ : rowPlan_;
    hipLaunchKernelGGL(k_.spmv[p.plan.rpw - 1], dim3(p.blocks), dim3(kWG), 0, stream_, p.plan, transpose ? tPtr_ : aPtr_,
                       transpose ? tIdx_ : aIdx_, transpose ? tVal_ : aVal_, in, out);
    HIP_OK(hipGetLastError());
    launches++;
  }
  void launch_col(bool check, int tloc) {
    if (!check && colPlan_.tiled) {
      hipLaunchKernelGGL(k_.colTiled[colPlan_.tileShape - 1], dim3(colPlan_.tiles.tiles),
                         dim3(64 * tile_waves(colPlan_.tileShape)), 0, stream_, colPlan_.tplan,
                         colPlan_.dLoc, tVal_, y_, yp_, x_, xp_, x0_, c_, lc_, lo_, up_, lb_, st_, tloc, s_.rho);
      return;
    }
    const int cr = colPlan_.plan.rpw - 1;
    hipLaunchKernelGGL(check ? k_.colChk[cr] : k_.col[cr], dim3(colPlan_.blocks), dim3(kWG), 0, stream_, colPlan_.plan, tPtr_,
                       tIdx_, tVal_, y_, yp_, x_, xp_, x0_, c_, lc_, lo_, up_, lb_, st_, tloc, s_.rho, colPart_,
                       colPlan_.blocks);
  }
  void launch_row(bool check, int tloc) {
    if (!check && rowPlan_.tiled) {
      hipLaunchKernelGGL(k_.rowTiled[rowPlan_.tileShape - 1], dim3(rowPlan_.tiles.tiles),
                         dim3(64 * tile_waves(rowPlan_.tileShape)), 0, stream_, rowPlan_.tplan,
                         rowPlan_.dLoc, aVal_, xp_, y_, yp_, y0_, ax_, axp_, ax0_, rl_, ru_, rl64_, ru64_, lr_, rfac_,
                         st_, tloc, s_.rho);
      return;
    }
    const int rr = rowPlan_.plan.rpw - 1;
    hipLaunchKernelGGL(check ? k_.rowChk[rr] : k_.row[rr], dim3(rowPlan_.blocks), dim3(kWG), 0, stream_, rowPlan_.plan, aPtr_,
                       aIdx_, aVal_, aVal64_, xp_, y_, yp_, y0_, ax_, axp_, ax0_, rl_, ru_, rl64_, ru64_, lr_, rfac_,
                       st_, tloc, s_.rho, rowPart_, rowPlan_.blocks);
  }
  void enqueue(int iters) {
    for (int t = 0; t < iters; ++t) {
      const bool chk = t == iters - 1;
      launch_col(chk, t);
      launch_row(chk, t);
    }
    hipLaunchKernelGGL(k_.colcheck[colPlan_.plan.rpw - 1], dim3(colPlan_.blocks), dim3(kWG), 0, stream_, colPlan_.plan, tPtr_, tIdx_,
                       tVal64_, yp_, xp_, c64_, lc_, lo64_, up64_, lb_, cfac_, st_, colPart_, colPlan_.blocks);
    hipLaunchKernelGGL(k_finalize, dim3(B_), dim3(kWG), 0, stream_, st_, colPart_, colPlan_.blocks, rowPart_,
                       rowPlan_.blocks, sc_, iters, s_.checkEvery);
    HIP_OK(hipGetLastError());
    launches += 2 * iters + 2;
    // FP32 iterates: the activity ax follows the Halpern recursion of x
    // (equal to A x in exact arithmetic), whose rounding drifts over long
    // stretches without a restart; once per block it is recomputed as A x.
    // FP64 keeps the recursion alone (the reference's arithmetic).
    if (sizeof(T) == 4) launch_spmv(false, x_, ax_);
  }
  void capture(int iters) {
    HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    enqueue(iters);
    launches -= 2 * iters + 2;
    HIP_OK(hipStreamEndCapture(stream_, &graph_));
    HIP_OK(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0));
  }

  RunSetup s_;
  int dev_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graphExec_ = nullptr;
  Buffers buf_;
  Kern<T> k_{};
  int m_ = 0, n_ = 0, B_ = 1, logB_ = 0, nnz_ = 0;
  bool split_ = false;
  T *stageT_ = nullptr, *axScratch_ = nullptr;  // load_lanes staging
  double* stage64_ = nullptr;
  int* laneIdx_ = nullptr;
  int *aPtr_ = nullptr, *aIdx_ = nullptr, *tPtr_ = nullptr, *tIdx_ = nullptr;
  T *aVal_ = nullptr, *tVal_ = nullptr;
  double *aVal64_ = nullptr, *tVal64_ = nullptr, *rfac_ = nullptr, *cfac_ = nullptr;
  PlanHost colPlan_, rowPlan_;
  Ld lc_{1, 0}, lb_{1, 0}, lr_{1, 0};
  size_t nc_ = 0, nbnd_ = 0, mr_ = 0;
  T *x_, *xp_, *x0_, *y_, *yp_, *y0_, *ax_, *axp_, *ax0_, *c_, *lo_, *up_, *rl_, *ru_;
  double *c64_, *lo64_, *up64_, *rl64_, *ru64_, *colPart_, *rowPart_, *cNorm_, *rhsNorm_;
  Inst st_{};
  RunScalars sc_{};
  CheckRecord* hostRec_ = nullptr;
};

}  // namespace

std::unique_ptr<DeviceRun> DeviceRun::create(int device, const Prepared& P, const RunSetup& s) {
  const DeviceInfo& di = device_info(device);
  if (!di.available) throw Error("GPU unavailable: " + di.reason);
  if (s.precision == PDHG_FP64) return std::make_unique<RunImpl<double>>(device, P, s);
  return std::make_unique<RunImpl<float>>(device, P, s);
}

// ---------------------------------------------------------------------------
// Op-level backend: the reference Backend vtable over HIP (C-ABI below).
// ---------------------------------------------------------------------------
namespace {

class OpBackendBase {
 public:
  virtual ~OpBackendBase() = default;
  virtual void set(int v, const double* d) = 0;
  virtual void get(int v, double* d) = 0;
  virtual void spmv(int in, int out, bool transpose) = 0;
  virtual void setSteps(const double* tau, const double* sigma) = 0;
  virtual void primal() = 0;
  virtual void dual() = 0;
  virtual void halpern(double rho) = 0;
  virtual void restart(const int* mask) = 0;
  virtual void kkt(double* r3, double* c3) = 0;
  virtual void diff2(int a, int b, double* out) = 0;
  virtual double timeSpmv(bool transpose, int reps) = 0;
  virtual int spmvLayout(bool transpose) const = 0;
  virtual void setSpmvLayout(bool transpose, int layout) = 0;
};

// Any k from 1 to 64 instances, as the reference's Backend::init takes
// (src/pdhg/pdhg.cpp:53-55).  The device vectors are interleaved over
// B = nextpow2(k) lanes (element (i, b) at i*B + b, the layout of the SpMV
// kernels); set / get convert from and to the caller's k-interleaved arrays,
// and the B - k padding lanes (zero data, unit steps) are never reported.
template <typename T>
class OpBackend final : public OpBackendBase {
 public:
  OpBackend(int m, int n, const Csr& A, const Csr& AT, const double* rfac, const double* cfac, int k, int device)
      : m_(m), n_(n), k_(k), A_(A), AT_(AT) {
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&e0_));
    HIP_OK(hipEventCreate(&e1_));
    PDHG_REQUIRE(k >= 1 && k <= 64, "backend batch width must be in 1..64");
    logB_ = ilog2(k);
    B_ = 1 << logB_;
    split_ = sizeof(T) == 8 && logB_ >= 4;  // the batched kernels' layout (kSplit)
    lay_ = OpLay{B_, split_ ? 1 : 0};
    kern_ = kern_table<T>(logB_);
    aPtr_ = up(A.ptr);
    aIdx_ = up(csr_padded(A.idx));  // whole-chunk reads past a row stay in bounds
    tPtr_ = up(AT.ptr);
    tIdx_ = up(csr_padded(AT.idx));
    aVal_ = upT(csr_padded(A.val));
    tVal_ = upT(csr_padded(AT.val));
    rfac_ = buf_.alloc<double>(m);
    cfac_ = buf_.alloc<double>(n);
    put(rfac_, rfac, m);
    put(cfac_, cfac, n);
    for (int v = 0; v < PDHG_NV; ++v) {
      const size_t len = dev_len(v);
      vec_[v] = buf_.alloc<T>(len);
      HIP_OK(hipMemsetAsync(vec_[v], 0, len * sizeof(T), stream_));
    }
    iter_ = buf_.alloc<int>(B_);
    mask_ = buf_.alloc<int>(B_);
    HIP_OK(hipMemsetAsync(iter_, 0, B_ * sizeof(int), stream_));
    HIP_OK(hipMemsetAsync(mask_, 0, B_ * sizeof(int), stream_));
    tau_ = buf_.alloc<double>(B_);
    sigma_ = buf_.alloc<double>(B_);
    const std::vector<double> one(B_, 1.0);
    put(tau_, one.data(), B_);
    put(sigma_, one.data(), B_);
    partial_ = buf_.alloc<double>(static_cast<size_t>(kRed) * k * 3);
    // Row mapping of the two SpMVs: the fastest candidate of the batched
    // engine's autotuner list (LDS tiles aside) for this matrix and width,
    // and at width 1 the sliced ELLPACK copy when it is faster still.
    double msA = 0, msAT = 0;
    rowPlan_ = tune_spmv(A, aPtr_, aIdx_, aVal_, vec_[PDHG_V_X], vec_[PDHG_V_AX], &msA);
    colPlan_ = tune_spmv(AT, tPtr_, tIdx_, tVal_, vec_[PDHG_V_Y], vec_[PDHG_V_ATY], &msAT);
    if (B_ == 1) {
      sellA_ = try_sell(A, vec_[PDHG_V_X], vec_[PDHG_V_AX], msA);
      sellAT_ = try_sell(AT, vec_[PDHG_V_Y], vec_[PDHG_V_ATY], msAT);
    }
    HIP_OK(hipStreamSynchronize(stream_));
  }
  ~OpBackend() override {
    (void)hipEventDestroy(e0_);
    (void)hipEventDestroy(e1_);
    (void)hipStreamDestroy(stream_);
  }
  int rows_of(int v) const { return v >= PDHG_V_Y ? m_ : n_; }
  size_t len_of(int v) const { return static_cast<size_t>(rows_of(v)) * k_; }   // caller's layout
  size_t dev_len(int v) const { return static_cast<size_t>(rows_of(v)) * B_; }  // device layout
  // Every host copy is ordered on the backend's own stream, after the
  // kernels already queued there (the stream does not synchronise with the
  // null stream), and waited for before the host buffer is released.
  void set(int v, const double* d) override {
    check_vec(v);
    PDHG_REQUIRE(d != nullptr, "null vector data");
    std::vector<T> t(dev_len(v), static_cast<T>(0));
    const int rows = rows_of(v);
    for (int i = 0; i < rows; ++i)
      for (int b = 0; b < k_; ++b) t[static_cast<size_t>(i) * B_ + b] = static_cast<T>(d[static_cast<size_t>(i) * k_ + b]);
    if (split_) to_split(t, rows);
    put(vec_[v], t.data(), t.size());
  }
  void get(int v, double* d) override {
    check_vec(v);
    PDHG_REQUIRE(d != nullptr, "null output");
    std::vector<T> t(dev_len(v));
    HIP_OK(hipMemcpyAsync(t.data(), vec_[v], t.size() * sizeof(T), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    const int rows = rows_of(v);
    if (split_) from_split(t, rows);
    for (int i = 0; i < rows; ++i)
      for (int b = 0; b < k_; ++b) d[static_cast<size_t>(i) * k_ + b] = t[static_cast<size_t>(i) * B_ + b];
  }
  void spmv(int in, int out, bool transpose) override {
    check_vec(in);
    check_vec(out);
    const SellDev& sd = transpose ? sellAT_ : sellA_;
    if (sd.use) {
      launch_sell(sd, vec_[in], vec_[out]);
      return;
    }
    const PlanHost& p = transpose ? colPlan_ : rowPlan_;
    launch_csr(p, transpose ? tPtr_ : aPtr_, transpose ? tIdx_ : aIdx_, transpose ? tVal_ : aVal_, vec_[in],
               vec_[out]);
  }
  // One SpMV under plan p: LDS-staged tiles (B = 64, p.tiled) or the row mapping.
  void launch_csr(const PlanHost& p, const int* ptr, const int* idx, const T* val, const T* in, T* out) {
    if (p.tiled) {
      hipLaunchKernelGGL(kern_.spmvTiled[p.tileShape - 1], dim3(p.tiles.tiles), dim3(64 * tile_waves(p.tileShape)), 0,
                         stream_, p.tplan, p.dLoc, val, in, out);
    } else {
      hipLaunchKernelGGL(kern_.spmv[p.plan.rpw - 1], dim3(p.blocks), dim3(kWG), 0, stream_, p.plan, ptr, idx, val, in,
                         out);
    }
    HIP_OK(hipGetLastError());
  }
  void setSteps(const double* tau, const double* sigma) override {
    PDHG_REQUIRE(tau && sigma, "null step sizes");
    put(tau_, tau, k_);  // padding lanes keep unit steps
    put(sigma_, sigma, k_);
  }
  void primal() override {
    const long long len = static_cast<long long>(n_) * B_;
    hipLaunchKernelGGL(k_op_primal<T>, dim3(grid(len)), dim3(kWG), 0, stream_, len, lay_, vec_[PDHG_V_X],
                       vec_[PDHG_V_ATY], vec_[PDHG_V_C], vec_[PDHG_V_LO], vec_[PDHG_V_UP], tau_, vec_[PDHG_V_XP]);
    HIP_OK(hipGetLastError());
  }
  void dual() override {
    const long long len = static_cast<long long>(m_) * B_;
    hipLaunchKernelGGL(k_op_dual<T>, dim3(grid(len)), dim3(kWG), 0, stream_, len, lay_, vec_[PDHG_V_Y],
                       vec_[PDHG_V_AXP], vec_[PDHG_V_AX], vec_[PDHG_V_RL], vec_[PDHG_V_RU], sigma_, vec_[PDHG_V_YP]);
    HIP_OK(hipGetLastError());
  }
  void halpern(double rho) override {
    auto one = [&](int rows, int z, int zp, int z0) {
      const long long len = static_cast<long long>(rows) * B_;
      hipLaunchKernelGGL(k_op_halpern<T>, dim3(grid(len)), dim3(kWG), 0, stream_, len, lay_, iter_, rho, vec_[z],
                         vec_[zp], vec_[z0]);
    };
    one(n_, PDHG_V_X, PDHG_V_XP, PDHG_V_X0);
    one(m_, PDHG_V_Y, PDHG_V_YP, PDHG_V_Y0);
    one(m_, PDHG_V_AX, PDHG_V_AXP, PDHG_V_AX0);
    hipLaunchKernelGGL(k_op_advance, dim3(1), dim3(64), 0, stream_, B_, iter_, mask_, 0);
    HIP_OK(hipGetLastError());
  }
  void restart(const int* mask) override {
    PDHG_REQUIRE(mask != nullptr, "null restart mask");
    put(mask_, mask, k_);
    auto cp = [&](int rows, int src, int d1, int d2) {
      const long long len = static_cast<long long>(rows) * B_;
      hipLaunchKernelGGL(k_op_restart<T>, dim3(grid(len)), dim3(kWG), 0, stream_, len, lay_, mask_, vec_[src],
                         vec_[d1], vec_[d2]);
    };
    cp(n_, PDHG_V_XP, PDHG_V_X, PDHG_V_X0);
    cp(m_, PDHG_V_YP, PDHG_V_Y, PDHG_V_Y0);
    cp(m_, PDHG_V_AXP, PDHG_V_AX, PDHG_V_AX0);
    hipLaunchKernelGGL(k_op_advance, dim3(1), dim3(64), 0, stream_, B_, iter_, mask_, 1);
    HIP_OK(hipGetLastError());
  }
  void kkt(double* r3, double* c3) override {
    std::fill(r3, r3 + 3 * k_, 0.0);
    std::fill(c3, c3 + 3 * k_, 0.0);
    std::vector<double> host;
    if (m_ > 0) {
      const int g = red_blocks(m_);
      hipLaunchKernelGGL(k_op_kkt_rows<T>, dim3(g, k_), dim3(kWG), 0, stream_, m_, lay_, vec_[PDHG_V_AXP],
                         vec_[PDHG_V_RL], vec_[PDHG_V_RU], vec_[PDHG_V_YP], rfac_, partial_);
      fetch_sum3(g, r3);
    }
    const int g = red_blocks(n_);
    hipLaunchKernelGGL(k_op_kkt_cols<T>, dim3(g, k_), dim3(kWG), 0, stream_, n_, lay_, vec_[PDHG_V_XP],
                       vec_[PDHG_V_C], vec_[PDHG_V_ATYP], vec_[PDHG_V_LO], vec_[PDHG_V_UP], cfac_, partial_);
    fetch_sum3(g, c3);
  }
  void diff2(int a, int b, double* out) override {
    check_vec(a);
    check_vec(b);
    const int len = a >= PDHG_V_Y ? m_ : n_;
    std::fill(out, out + k_, 0.0);
    if (len == 0) return;
    const int g = red_blocks(len);
    hipLaunchKernelGGL(k_op_diff2<T>, dim3(g, k_), dim3(kWG), 0, stream_, len, lay_, vec_[a], vec_[b], partial_);
    std::vector<double> r3(3 * k_, 0.0);
    fetch_sum3(g, r3.data());
    for (int q = 0; q < k_; ++q) out[q] = r3[3 * q];
  }
  int spmvLayout(bool transpose) const override {
    if ((transpose ? sellAT_ : sellA_).use) return 1;
    return (transpose ? colPlan_ : rowPlan_).tiled ? 2 : 0;
  }
  void setSpmvLayout(bool transpose, int layout) override {
    PDHG_REQUIRE(layout >= 0 && layout <= 2, "spmv layout must be 0 (CSR rows), 1 (sliced ELLPACK) or 2 (LDS tiles)");
    SellDev& d = transpose ? sellAT_ : sellA_;
    PDHG_REQUIRE(layout != 1 || d.built, "sliced ELLPACK needs a backend of width 1");
    PDHG_REQUIRE(layout != 2 || B_ == 64, "LDS tiles need a backend of width 33..64");
    d.use = layout == 1 && d.rows > 0;
    if (layout == 1) return;
    // The storage's default plan: one row per wavefront (rows bucketed by
    // length below width 64), or 8-row tiles.
    const Csr& M = transpose ? AT_ : A_;
    PlanHost& cur = transpose ? colPlan_ : rowPlan_;
    MapSetup ms{4, 1, B_ >= 16, false, 1, layout == 2 ? 2 : 0};
    PlanHost h = build_plan(M, logB_, ms, static_cast<int>(sizeof(T)));
    PDHG_REQUIRE(layout != 2 || h.tiled, "a row is longer than the tiles' column budget");
    if (!h.perm.empty()) h.dPerm = up(h.perm);
    h.plan.perm = h.dPerm;
    if (h.tiled) up_tiles(h);
    cur = h;
  }
  double timeSpmv(bool transpose, int reps) override {
    const int in = transpose ? PDHG_V_Y : PDHG_V_X, out = transpose ? PDHG_V_ATY : PDHG_V_AX;
    spmv(in, out, transpose);  // warm
    HIP_OK(hipEventRecord(e0_, stream_));
    for (int r = 0; r < reps; ++r) spmv(in, out, transpose);
    HIP_OK(hipEventRecord(e1_, stream_));
    HIP_OK(hipEventSynchronize(e1_));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0_, e1_));
    return ms / reps;
  }

 private:
  static constexpr int kRed = 256;
  void check_vec(int v) const { PDHG_REQUIRE(v >= 0 && v < PDHG_NV, "vector id out of range"); }
  static int grid(long long len) { return static_cast<int>(std::max(1LL, std::min(8192LL, (len + kWG - 1) / kWG))); }
  static int red_blocks(int len) { return std::max(1, std::min(kRed, (len + kWG - 1) / kWG)); }
  void fetch_sum3(int g, double* acc) {
    std::vector<double> host(static_cast<size_t>(g) * k_ * 3);
    HIP_OK(hipMemcpyAsync(host.data(), partial_, host.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    for (int b = 0; b < k_; ++b)  // fixed order, as the reference host sum (pdhg.cpp:346-348)
      for (int q = 0; q < g; ++q)
        for (int s = 0; s < 3; ++s) acc[3 * b + s] += host[(static_cast<size_t>(b) * g + q) * 3 + s];
  }
  template <typename U>
  void put(U* d, const U* h, size_t count) {
    if (count == 0) return;
    HIP_OK(hipMemcpyAsync(d, h, count * sizeof(U), hipMemcpyHostToDevice, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
  }
  int* up(const std::vector<int>& v) {
    int* d = buf_.alloc<int>(v.size());
    put(d, v.data(), v.size());
    return d;
  }
  // B-strided FP64 rows <-> the split layout (low words of a row's B lanes,
  // then their high words).
  void to_split(std::vector<T>& t, int rows) const {
    std::vector<unsigned> u(t.size() * 2);
    for (size_t i = 0; i < static_cast<size_t>(rows); ++i)
      for (int b = 0; b < B_; ++b) {
        unsigned long long bits;
        std::memcpy(&bits, &t[i * B_ + b], sizeof(bits));
        u[i * 2 * B_ + b] = static_cast<unsigned>(bits);
        u[i * 2 * B_ + B_ + b] = static_cast<unsigned>(bits >> 32);
      }
    std::memcpy(t.data(), u.data(), t.size() * sizeof(T));
  }
  void from_split(std::vector<T>& t, int rows) const {
    std::vector<unsigned> u(t.size() * 2);
    std::memcpy(u.data(), t.data(), t.size() * sizeof(T));
    for (size_t i = 0; i < static_cast<size_t>(rows); ++i)
      for (int b = 0; b < B_; ++b) {
        const unsigned long long bits =
            (static_cast<unsigned long long>(u[i * 2 * B_ + B_ + b]) << 32) | u[i * 2 * B_ + b];
        std::memcpy(&t[i * B_ + b], &bits, sizeof(bits));
      }
  }
  // Times every candidate mapping of one SpMV (8 launches after a warm one,
  // HIP events on the backend's stream, zero vectors) and keeps the fastest
  // (its mean launch time in *bestMsOut).  From B = 16 on the XCD remap
  // stays on, as in the engine's tuner.
  PlanHost tune_spmv(const Csr& M, const int* ptr, const int* idx, const T* val, const T* in, T* out,
                     double* bestMsOut) {
    std::vector<MapSetup> cand;
    for (bool xcd : {false, true}) {
      if (!xcd && B_ >= 16) continue;
      if (B_ == 64) {
        for (int rpw : {1, 2}) cand.push_back(MapSetup{4, 1, xcd, false, rpw});
        for (int shape = 1; shape <= kTileShapes; ++shape)
          cand.push_back(MapSetup{4, 1, xcd, false, 1, shape});  // rows in LDS-staged tiles
        continue;
      }
      for (int per : {2, 4, 8, 16})
        for (int map : {1, 2}) cand.push_back(MapSetup{per, map, xcd});
      if (B_ == 16 || B_ == 32)
        for (int map : {1, 2}) cand.push_back(MapSetup{4, map, xcd, true});
    }
    int* scratch = buf_.alloc<int>(static_cast<size_t>(std::max(1, M.rows)));
    double bestMs = 0;
    size_t best = 0;
    for (size_t c = 0; c < cand.size(); ++c) {
      PlanHost h = build_plan(M, logB_, cand[c], static_cast<int>(sizeof(T)));
      if (h.tiled) {
        up_tiles(h);  // their own arrays, kept with the backend's buffers
      } else {
        if (!h.perm.empty()) put(scratch, h.perm.data(), h.perm.size());
        h.plan.perm = h.perm.empty() ? nullptr : scratch;
      }
      auto launch = [&] { launch_csr(h, ptr, idx, val, in, out); };
      launch();
      HIP_OK(hipEventRecord(e0_, stream_));
      for (int r = 0; r < 8; ++r) launch();
      HIP_OK(hipEventRecord(e1_, stream_));
      HIP_OK(hipEventSynchronize(e1_));
      HIP_OK(hipGetLastError());
      float ms = 0;
      HIP_OK(hipEventElapsedTime(&ms, e0_, e1_));
      if (c == 0 || ms < bestMs) {
        bestMs = ms;
        best = c;
      }
    }
    PlanHost h = build_plan(M, logB_, cand[best], static_cast<int>(sizeof(T)));
    if (!h.perm.empty()) h.dPerm = up(h.perm);
    h.plan.perm = h.dPerm;
    if (h.tiled) up_tiles(h);
    *bestMsOut = bestMs / 8;
    return h;
  }
  // Device copies of a tiled plan's records and tile-local columns.
  void up_tiles(PlanHost& h) {
    h.tplan.perm = nullptr;
    h.tplan.desc = up(h.desc);
    h.dLoc = up(h.tiles.loc);
  }
  // Width 1: the sliced ELLPACK copy of M, timed like the candidates above;
  // used when faster than the best of them (csrMs per launch).
  struct SellDev {
    int* soff = nullptr;
    int* sidx = nullptr;
    T* sval = nullptr;
    int* perm = nullptr;
    int slices = 0, rows = 0;
    bool built = false, use = false;
  };
  void launch_sell(const SellDev& d, const T* in, T* out) {
    constexpr int kSlicesPerWG = kWG / 64;
    hipLaunchKernelGGL(k_spmv_sell<T>, dim3((d.slices + kSlicesPerWG - 1) / kSlicesPerWG), dim3(kWG), 0, stream_,
                       d.rows, d.slices, d.soff, d.sidx, d.sval, d.perm, in, out);
    HIP_OK(hipGetLastError());
  }
  SellDev try_sell(const Csr& M, const T* in, T* out, double csrMs) {
    SellDev d;
    d.built = true;
    if (M.rows == 0) return d;
    const SellHost h = build_sell(M);
    d.soff = up(h.soff);
    d.sidx = up(h.sidx);
    d.sval = upT(h.sval);
    d.perm = up(h.perm);
    d.slices = h.slices;
    d.rows = M.rows;
    launch_sell(d, in, out);
    HIP_OK(hipEventRecord(e0_, stream_));
    for (int r = 0; r < 8; ++r) launch_sell(d, in, out);
    HIP_OK(hipEventRecord(e1_, stream_));
    HIP_OK(hipEventSynchronize(e1_));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0_, e1_));
    d.use = ms / 8 < csrMs;
    return d;
  }
  T* upT(const std::vector<double>& v) {
    std::vector<T> t(v.begin(), v.end());
    T* d = buf_.alloc<T>(v.size());
    put(d, t.data(), t.size());
    return d;
  }

  int m_, n_, k_, logB_ = 0, B_ = 1;
  Csr A_, AT_;  // host copies, for plans built on request (setSpmvLayout)
  bool split_ = false;
  OpLay lay_{1, 0};
  hipStream_t stream_ = nullptr;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  Buffers buf_;
  Kern<T> kern_{};
  int *aPtr_, *aIdx_, *tPtr_, *tIdx_;
  T *aVal_, *tVal_;
  PlanHost rowPlan_, colPlan_;
  SellDev sellA_, sellAT_;
  double *rfac_, *cfac_, *tau_, *sigma_, *partial_;
  T* vec_[PDHG_NV];
  int *iter_, *mask_;
};

// ---- KKT sums without allocation ------------------------------------------
// The block partials of the KKT kernels live in one fixed device array of the
// library, `scratch`: the row partials in its first half, the column partials
// in its second, at most kNoAllocBlocks blocks x 64 instances x 3 sums each.
// Each GPU holds its own copy until the process ends.  Host code never uses
// the array's name as an address (it is a host-side stand-in there): it uses
// the per-device address hipGetSymbolAddress returns.
constexpr int kNoAllocBlocks = 256;
constexpr int kNoAllocMaxK = 64;
constexpr size_t kScratchHalf = static_cast<size_t>(kNoAllocBlocks) * kNoAllocMaxK * 3;
static __device__ __align__(256) double scratch[2 * kScratchHalf];

// out[3b + s] = sum over blocks q = 0 .. g-1 of partial[(b g + q) 3 + s], in
// that order from 0: the order of the host sum of the op-level backend.
__global__ void k_sum3_fixed(int g, int k, const double* partial, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * k) return;
  const int b = t / 3, s = t % 3;
  double acc = 0.0;
  for (int q = 0; q < g; ++q) acc += partial[(static_cast<size_t>(b) * g + q) * 3 + s];
  out[t] = acc;
}

std::mutex gScratchMutex;
double* gScratchPtr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
double* scratch_on_current_device() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  PDHG_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(gScratchMutex);
  if (!gScratchPtr[dev]) {
    void* ptr = nullptr;
    HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    PDHG_REQUIRE(ptr != nullptr, "no device address for the scratch array");
    gScratchPtr[dev] = static_cast<double*>(ptr);
  }
  return gScratchPtr[dev];
}

template <typename T>
void kkt_noalloc(int m, int n, int k, const void* axp, const void* rl, const void* ru, const void* yp,
                 const double* rfac, const void* xp, const void* c, const void* atyp, const void* lo, const void* up,
                 const double* cfac, double* rows3k, double* cols3k, hipStream_t s) {
  double* part = scratch_on_current_device();
  auto blocks = [](int len) { return std::max(1, std::min(kNoAllocBlocks, (len + kWG - 1) / kWG)); };
  const dim3 sumGrid((3 * k + 63) / 64), sumBlock(64);
  const int gr = m > 0 ? blocks(m) : 0, gc = n > 0 ? blocks(n) : 0;
  if (gr > 0)
    hipLaunchKernelGGL(k_op_kkt_rows<T>, dim3(gr, k), dim3(kWG), 0, s, m, OpLay{k, 0}, static_cast<const T*>(axp),
                       static_cast<const T*>(rl), static_cast<const T*>(ru), static_cast<const T*>(yp), rfac, part);
  hipLaunchKernelGGL(k_sum3_fixed, sumGrid, sumBlock, 0, s, gr, k, part, rows3k);
  if (gc > 0)
    hipLaunchKernelGGL(k_op_kkt_cols<T>, dim3(gc, k), dim3(kWG), 0, s, n, OpLay{k, 0}, static_cast<const T*>(xp),
                       static_cast<const T*>(c), static_cast<const T*>(atyp), static_cast<const T*>(lo),
                       static_cast<const T*>(up), cfac, part + kScratchHalf);
  hipLaunchKernelGGL(k_sum3_fixed, sumGrid, sumBlock, 0, s, gc, k, part + kScratchHalf, cols3k);
  HIP_OK(hipGetLastError());
}

thread_local std::string tErr;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    tErr = e.what();
    return -1;
  }
}

}  // namespace

const char* device_last_error() { return tErr.c_str(); }

}  // namespace pdhgamd

// ---------------------------------------------------------------------------
// C-ABI: op-level backend
// ---------------------------------------------------------------------------
struct pdhg_backend {
  std::unique_ptr<pdhgamd::OpBackendBase> impl;
};

extern "C" {

void pdhg_set_last_error(const char* msg);

pdhg_backend* pdhg_backend_create(int m, int n, const int* a_ptr, const int* a_idx, const double* a_val,
                                  const int* t_ptr, const int* t_idx, const double* t_val, const double* rfac,
                                  const double* cfac, int k, int precision, int device) {
  try {
    using namespace pdhgamd;
    // The inputs are checked before any device work, as model_from checks a
    // C-ABI LP: the kernels trust row pointers and indices.
    PDHG_REQUIRE(m >= 0 && n >= 0 && k >= 1, "bad sizes");
    auto csr = [](int rows, int cols, const int* ptr, const int* idx, const double* val, const char* what) {
      PDHG_REQUIRE(ptr != nullptr, std::string(what) + ": null row pointers");
      PDHG_REQUIRE(ptr[0] == 0, std::string(what) + ": row pointers must start at 0");
      for (int i = 0; i < rows; ++i)
        PDHG_REQUIRE(ptr[i] <= ptr[i + 1], std::string(what) + ": row pointers not monotone");
      Csr M;
      M.rows = rows;
      M.cols = cols;
      M.ptr.assign(ptr, ptr + rows + 1);
      const int nnz = M.ptr.back();
      PDHG_REQUIRE(nnz == 0 || (idx && val), std::string(what) + ": null index or value array");
      M.idx.assign(idx, idx + nnz);
      M.val.assign(val, val + nnz);
      for (int e = 0; e < nnz; ++e) {
        PDHG_REQUIRE(M.idx[e] >= 0 && M.idx[e] < cols, std::string(what) + ": column index out of range");
        PDHG_REQUIRE(std::isfinite(M.val[e]), std::string(what) + ": non-finite value");
      }
      return M;
    };
    Csr A = csr(m, n, a_ptr, a_idx, a_val, "A");
    Csr AT = csr(n, m, t_ptr, t_idx, t_val, "A^T");
    PDHG_REQUIRE(A.nnz() == AT.nnz(), "A and A^T nonzero counts differ");
    PDHG_REQUIRE((m == 0 || rfac) && (n == 0 || cfac), "null scale factors");
    for (int i = 0; i < m; ++i) PDHG_REQUIRE(std::isfinite(rfac[i]), "non-finite row factor");
    for (int j = 0; j < n; ++j) PDHG_REQUIRE(std::isfinite(cfac[j]), "non-finite column factor");
    PDHG_REQUIRE(precision == PDHG_FP32 || precision == PDHG_FP64, "backend precision must be FP32 or FP64");
    const DeviceInfo& di = device_info(device);
    if (!di.available) throw Error("GPU unavailable: " + di.reason);
    auto* h = new pdhg_backend;
    if (precision == PDHG_FP64) h->impl = std::make_unique<OpBackend<double>>(m, n, A, AT, rfac, cfac, k, device);
    else h->impl = std::make_unique<OpBackend<float>>(m, n, A, AT, rfac, cfac, k, device);
    return h;
  } catch (const std::exception& e) {
    pdhg_set_last_error(e.what());
    return nullptr;
  }
}

void pdhg_backend_destroy(pdhg_backend* be) { delete be; }

#define PDHG_BE_CALL(body)                                                  \
  do {                                                                      \
    if (!be) {                                                              \
      pdhg_set_last_error("null backend");                                  \
      return -1;                                                            \
    }                                                                       \
    if (pdhgamd::guarded([&] { body; }) != 0) {                             \
      pdhg_set_last_error(pdhgamd::device_last_error());                    \
      return -1;                                                            \
    }                                                                       \
    return 0;                                                               \
  } while (0)

int pdhg_backend_set(pdhg_backend* be, int vec, const double* data) { PDHG_BE_CALL(be->impl->set(vec, data)); }
int pdhg_backend_get(pdhg_backend* be, int vec, double* out) { PDHG_BE_CALL(be->impl->get(vec, out)); }
int pdhg_backend_spmv(pdhg_backend* be, int in, int out) { PDHG_BE_CALL(be->impl->spmv(in, out, false)); }
int pdhg_backend_spmvT(pdhg_backend* be, int in, int out) { PDHG_BE_CALL(be->impl->spmv(in, out, true)); }
int pdhg_backend_set_steps(pdhg_backend* be, const double* tau, const double* sigma) {
  PDHG_BE_CALL(be->impl->setSteps(tau, sigma));
}
int pdhg_backend_primal_step(pdhg_backend* be) { PDHG_BE_CALL(be->impl->primal()); }
int pdhg_backend_dual_step(pdhg_backend* be) { PDHG_BE_CALL(be->impl->dual()); }
int pdhg_backend_halpern(pdhg_backend* be, double rho) { PDHG_BE_CALL(be->impl->halpern(rho)); }
int pdhg_backend_restart(pdhg_backend* be, const int* mask) { PDHG_BE_CALL(be->impl->restart(mask)); }
int pdhg_backend_kkt(pdhg_backend* be, double* rows3k, double* cols3k) { PDHG_BE_CALL(be->impl->kkt(rows3k, cols3k)); }
int pdhg_backend_diff2(pdhg_backend* be, int a, int b, double* out) { PDHG_BE_CALL(be->impl->diff2(a, b, out)); }
int pdhg_backend_time_spmv(pdhg_backend* be, int transpose, int reps, double* ms) {
  PDHG_BE_CALL(*ms = be->impl->timeSpmv(transpose != 0, reps < 1 ? 1 : reps));
}

int pdhg_backend_spmv_layout(pdhg_backend* be, int transpose, int* layout) {
  PDHG_BE_CALL(*layout = be->impl->spmvLayout(transpose != 0));
}

int pdhg_backend_set_spmv_layout(pdhg_backend* be, int transpose, int layout) {
  PDHG_BE_CALL(be->impl->setSpmvLayout(transpose != 0, layout));
}

int pdhg_kkt_noalloc(int m, int n, int k, int precision, const void* axp, const void* rl, const void* ru,
                     const void* yp, const double* rfac, const void* xp, const void* c, const void* atyp,
                     const void* lo, const void* up, const double* cfac, double* rows3k, double* cols3k,
                     void* stream) {
  try {
    using namespace pdhgamd;
    PDHG_REQUIRE(m >= 0 && n >= 0, "bad sizes");
    PDHG_REQUIRE(k >= 1 && k <= kNoAllocMaxK, "pdhg_kkt_noalloc: k must be in 1..64 (the scratch array's limit)");
    PDHG_REQUIRE(precision == PDHG_FP32 || precision == PDHG_FP64, "precision must be FP32 or FP64");
    PDHG_REQUIRE(rows3k && cols3k, "null output");
    PDHG_REQUIRE(m == 0 || (axp && rl && ru && yp && rfac), "null row array");
    PDHG_REQUIRE(n == 0 || (xp && c && atyp && lo && up && cfac), "null column array");
    auto s = static_cast<hipStream_t>(stream);
    if (precision == PDHG_FP64)
      kkt_noalloc<double>(m, n, k, axp, rl, ru, yp, rfac, xp, c, atyp, lo, up, cfac, rows3k, cols3k, s);
    else
      kkt_noalloc<float>(m, n, k, axp, rl, ru, yp, rfac, xp, c, atyp, lo, up, cfac, rows3k, cols3k, s);
    return 0;
  } catch (const std::exception& e) {
    pdhg_set_last_error(e.what());
    return -1;
  }
}

}  // extern "C"
