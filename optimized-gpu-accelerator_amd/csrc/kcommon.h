// Device-side building blocks shared by the batched and the packed PDHG
// kernels (gfx950): row-to-lane mapping, the row dot products, fixed-order
// workgroup reductions, per-instance state, and small host helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "device.h"
#include "tiles.h"

namespace pdhgamd {

#define HIP_OK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess)                                                                              \
      throw ::pdhgamd::Error(std::string("HIP error '") + hipGetErrorString(_e) + "' at " #expr);      \
  } while (0)

namespace {

#ifndef PDHG_LOG_WG
#define PDHG_LOG_WG 8
#endif
constexpr int kLogWG = PDHG_LOG_WG;  // threads per workgroup: 256 (4 wave64)
constexpr int kWG = 1 << kLogWG;
constexpr int kWaves = kWG / 64;
constexpr int kMaxBuckets = 8;

// ---------------------------------------------------------------------------
// Row mapping
// ---------------------------------------------------------------------------
struct RowPlan {
  int nb;                          // buckets in use
  int blkStart[kMaxBuckets + 1];   // first workgroup of each bucket
  int rowStart[kMaxBuckets];       // offset of the bucket in perm (or its first row)
  int rowCount[kMaxBuckets];
  int logS[kMaxBuckets];           // log2 lanes per row (per instance)
  const int* perm;                 // row order, NULL = identity
  int xcd;                         // 1: contiguous row ranges per XCD
  int grp;                         // 1: one row per B-lane group, entries via the scalar cache
  int rpw;                         // B = 64: rows per wavefront (1 or 2)
  // B = 64 tiles (one workgroup each): one record per (tile, wavefront)
  // with the wavefront's row, its entry range and the tile columns it
  // stages in LDS (tile_descriptors in tiles.h).
  const int* desc;
};

// Workgroup ids are dispatched round-robin over the 8 XCDs; this bijection
// gives XCD x a contiguous range of logical blocks, so neighbouring rows (which
// share columns) meet in the same L2.
__device__ inline int remap_xcd(int bid, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int x = bid & 7, l = bid >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + l;
}

struct Slot {
  int row;   // -1: no row for this lane
  int sub;   // sub-lane within the row group
  int logS;
};

template <int LOGB>
__device__ inline Slot locate(const RowPlan& p, int bid) {
  int q = 0;
  while (q + 1 < p.nb && bid >= p.blkStart[q + 1]) ++q;
  Slot s;
  s.logS = p.logS[q];
  const int logG = LOGB + s.logS;
  const int t = threadIdx.x;
  const int r = ((bid - p.blkStart[q]) << (kLogWG - logG)) + (t >> logG);
  s.sub = (t & ((1 << logG) - 1)) >> LOGB;
  s.row = r < p.rowCount[q] ? (p.perm ? p.perm[p.rowStart[q] + r] : p.rowStart[q] + r) : -1;
  return s;
}

__device__ inline int block_id(const RowPlan& p) { return p.xcd ? remap_xcd(blockIdx.x, gridDim.x) : blockIdx.x; }

// Sum over the S sub-lanes of a row group (lanes b, b+B, b+2B, ...).
template <int LOGB, typename T>
__device__ inline T reduce_sub(T v, int logS) {
  for (int o = 1 << LOGB; o < (1 << (LOGB + logS)); o <<= 1) v += __shfl_xor(v, o);
  return v;
}

// Vector layout of a B-lane run (B = 1 << LOGB): element (i, b) sits at
// i*B + b, except for FP64 at B >= 16 ("split"): the B values of row i take
// 2B dwords, the low words of lanes 0..B-1 followed by their high words.  A
// 64-lane access is then two dword accesses instead of one dwordx2 access,
// which costs the vector L1 twice the cycles per byte.
template <typename T, int LOGB>
constexpr bool kSplit = sizeof(T) == 8 && LOGB >= 4;

template <bool SPLIT, int LOGB, typename T>
__device__ inline T vld(const T* __restrict__ p, size_t i, int b) {
  if constexpr (SPLIT) {
    static_assert(sizeof(T) == 8, "split layout is FP64");
    const unsigned* u = reinterpret_cast<const unsigned*>(p) + (i << (LOGB + 1));
    const unsigned lo = u[b], hi = u[(1 << LOGB) + b];
    return __builtin_bit_cast(T, (static_cast<unsigned long long>(hi) << 32) | lo);
  } else {
    return p[(i << LOGB) + b];
  }
}

template <bool SPLIT, int LOGB, typename T>
__device__ inline void vst(T* __restrict__ p, size_t i, int b, T v) {
  if constexpr (SPLIT) {
    static_assert(sizeof(T) == 8, "split layout is FP64");
    unsigned* u = reinterpret_cast<unsigned*>(p) + (i << (LOGB + 1));
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, v);
    u[b] = static_cast<unsigned>(bits);
    u[(1 << LOGB) + b] = static_cast<unsigned>(bits >> 32);
  } else {
    p[(i << LOGB) + b] = v;
  }
}

// sum_e val[e] * x[idx[e]*B + b] over e = e0, e0+step, ... < e1, with four
// independent chains so the index, value and gather loads of four nonzeros are
// in flight together.
template <int LOGB, typename Acc, bool SPLIT = false, typename V, typename X>
__device__ inline Acc row_dot(const int* __restrict__ idx, const V* __restrict__ val, const X* __restrict__ x,
                              int e0, int e1, int step, int b) {
  Acc a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int e = e0;
  for (; e + 3 * step < e1; e += 4 * step) {
    const int c0 = idx[e], c1 = idx[e + step], c2 = idx[e + 2 * step], c3 = idx[e + 3 * step];
    const Acc v0 = val[e], v1 = val[e + step], v2 = val[e + 2 * step], v3 = val[e + 3 * step];
    a0 += v0 * static_cast<Acc>(vld<SPLIT, LOGB>(x, c0, b));
    a1 += v1 * static_cast<Acc>(vld<SPLIT, LOGB>(x, c1, b));
    a2 += v2 * static_cast<Acc>(vld<SPLIT, LOGB>(x, c2, b));
    a3 += v3 * static_cast<Acc>(vld<SPLIT, LOGB>(x, c3, b));
  }
  for (; e < e1; e += step) a0 += static_cast<Acc>(val[e]) * static_cast<Acc>(vld<SPLIT, LOGB>(x, idx[e], b));
  return (a0 + a1) + (a2 + a3);
}

// B = 64: a whole wavefront serves one row (lane = instance), so the row's
// bounds are wave-uniform.  Column indices and values then come through the
// scalar cache (s_load) eight at a time, and the eight 64-lane gathers of
// a chunk are issued back to back before the first multiply-add waits.  The
// last partial chunk is predicated: its spare slots re-read the row's first
// column (a cache hit) with a zero weight, so no serial tail remains.
template <typename Acc, bool SPLIT, typename V, typename X>
__device__ inline Acc row_dot_wave(const int* __restrict__ idx, const V* __restrict__ val, const X* __restrict__ x,
                                   int e0, int e1, int b) {
  constexpr int CH = 8;
  e0 = __builtin_amdgcn_readfirstlane(e0);
  e1 = __builtin_amdgcn_readfirstlane(e1);
  Acc a[4] = {0, 0, 0, 0};
  int e = e0;
  for (; e + CH <= e1; e += CH) {
    int c[CH];
    Acc v[CH], g[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      c[u] = idx[e + u];
      v[u] = static_cast<Acc>(val[e + u]);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) g[u] = static_cast<Acc>(vld<SPLIT, 6>(x, c[u], b));
#pragma unroll
    for (int u = 0; u < CH; ++u) a[u & 3] += v[u] * g[u];
  }
  if (e < e1) {
    const int last = e1 - 1;
    int c[CH];
    Acc v[CH], g[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int p = min(e + u, last);  // always a valid entry of this row
      c[u] = idx[p];
      const Acc w = static_cast<Acc>(val[p]);
      v[u] = e + u <= last ? w : static_cast<Acc>(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) g[u] = static_cast<Acc>(vld<SPLIT, 6>(x, c[u], b));
#pragma unroll
    for (int u = 0; u < CH; ++u) a[u & 3] += v[u] * g[u];
  }
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// B = 64 over a staged tile: row_dot_wave with the gathers read from LDS,
// where tile column u of lane b sits at tile[u * 64 + b], and loc[e] the tile
// column of entry e.  Same chunks, predication and summation order.
template <typename Acc, typename V, typename X>
__device__ inline Acc row_dot_lds(const int* __restrict__ loc, const V* __restrict__ val, const X* tile, int e0,
                                  int e1, int b) {
  constexpr int CH = 8;
  e0 = __builtin_amdgcn_readfirstlane(e0);
  e1 = __builtin_amdgcn_readfirstlane(e1);
  Acc a[4] = {0, 0, 0, 0};
  int e = e0;
  for (; e + CH <= e1; e += CH) {
    int c[CH];
    Acc v[CH], g[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      c[u] = loc[e + u];
      v[u] = static_cast<Acc>(val[e + u]);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) g[u] = static_cast<Acc>(tile[c[u] * 64 + b]);
#pragma unroll
    for (int u = 0; u < CH; ++u) a[u & 3] += v[u] * g[u];
  }
  if (e < e1) {
    const int last = e1 - 1;
    int c[CH];
    Acc v[CH], g[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int p = min(e + u, last);
      c[u] = loc[p];
      const Acc w = static_cast<Acc>(val[p]);
      v[u] = e + u <= last ? w : static_cast<Acc>(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) g[u] = static_cast<Acc>(tile[c[u] * 64 + b]);
#pragma unroll
    for (int u = 0; u < CH; ++u) a[u & 3] += v[u] * g[u];
  }
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// row_dot_lds for a row of len <= 64 entries whose tile columns and values
// were staged in LDS (el[k], ev[k] for k < len) in the tile's load round,
// with every operand read from LDS.  Entry k goes into accumulator k mod 4
// in entry order, as in row_dot_wave; chunks of 4 keep the entries' columns
// and values (now vector registers) within the occupancy budget, and the
// predicated last chunk adds weight-0 terms only, so the sums are the same.
template <typename Acc, typename X>
__device__ inline Acc row_dot_lds_staged(const int* el, const Acc* ev, const X* tile, int len, int b) {
  constexpr int CH = 4;
  len = __builtin_amdgcn_readfirstlane(len);
  Acc a[4] = {0, 0, 0, 0};
  int k = 0;
  for (; k + CH <= len; k += CH) {
    int c[CH];
    Acc v[CH], g[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      c[u] = el[k + u];
      v[u] = ev[k + u];
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) g[u] = static_cast<Acc>(tile[c[u] * 64 + b]);
#pragma unroll
    for (int u = 0; u < CH; ++u) a[u & 3] += v[u] * g[u];
  }
  if (k < len) {
    const int last = len - 1;
    int c[CH];
    Acc v[CH], g[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int p = min(k + u, last);
      c[u] = el[p];
      const Acc w = ev[p];
      v[u] = k + u <= last ? w : static_cast<Acc>(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) g[u] = static_cast<Acc>(tile[c[u] * 64 + b]);
#pragma unroll
    for (int u = 0; u < CH; ++u) a[u & 3] += v[u] * g[u];
  }
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// Staging of tile t of a B = 64 plan in LDS: tile position u of lane b
// holds the gathered row of the tile's u-th column for lane b (src may
// differ per lane).  Of a tile shape's WAVES wavefronts, wavefront w stages
// positions w, w + WAVES, ...; NS positions at most.  Its record
// (tile_descriptors) comes in one round of scalar loads with no load before
// it, so the gathers start one memory latency after launch.
template <typename T, int SHAPE>
constexpr int kStageCols = (kTileCols<T, SHAPE> + tile_waves(SHAPE) - 1) / tile_waves(SHAPE);

// The wavefront's record of tile t, read into registers (wave-uniform).
template <int NS>
struct TileRec {
  int row, e0, e1, ns;
  int col[NS];
};

template <int NS, int WAVES>
__device__ inline TileRec<NS> tile_record(const RowPlan& p, int t) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int* __restrict__ d = p.desc + (static_cast<size_t>(t) * WAVES + wave) * desc_stride(NS);
  TileRec<NS> r;
  r.row = d[0];
  r.e0 = d[1];
  r.e1 = d[2];
  r.ns = d[3];
#pragma unroll
  for (int k = 0; k < NS; ++k) r.col[k] = d[4 + k];
  return r;
}

// tile_load issues a wavefront's share of the staging loads, tile_store
// writes them to LDS; a kernel issues its other loads in between, then meets
// the others at the barrier.
// A lane whose instance is done (live false) loads nothing: once a run's live
// instances sit in consecutive lanes, the cache lines of the finished half
// of a staged row are not fetched at all.
template <bool SPLIT, typename T, int NS>
__device__ inline void tile_load(const TileRec<NS>& r, const T* src, int b, bool live, T (&v)[NS]) {
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    v[k] = static_cast<T>(0);
    if (k < r.ns && live) v[k] = vld<SPLIT, 6>(src, r.col[k], b);
  }
}

template <int WAVES, typename T, int NS>
__device__ inline void tile_store(const TileRec<NS>& r, const T (&v)[NS], T* tile, int b) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
#pragma unroll
  for (int k = 0; k < NS; ++k)
    if (k < r.ns) tile[(wave + k * WAVES) * 64 + b] = v[k];
  __syncthreads();
}

// B = 16 or 32, one row per B-lane group (G = 64 / B rows per wavefront):
// each group's column indices and values come through the scalar cache a
// chunk at a time, every lane takes its own group's entry with a select, and
// the CH gathers of a chunk are issued back to back.  rows[g] is the wave's
// row for group g (-1: none); a group whose row is done loads nothing and
// gathers column 0 with weight 0.
template <int LOGB, typename Acc, bool SPLIT, typename V, typename X>
__device__ inline Acc row_dot_groups(const int* __restrict__ ptr, const int* __restrict__ idx,
                                     const V* __restrict__ val, const X* __restrict__ x, const int (&rows)[64 >> LOGB],
                                     int b) {
  constexpr int G = 64 >> LOGB;
  constexpr int CH = G == 2 ? 8 : 4;
  static_assert(CH <= kCsrPad, "CSR padding too short for the chunk");
  const int g = (threadIdx.x & 63) >> LOGB;
  int e0[G], len[G];
  int maxLen = 0;
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int r = rows[q];
    e0[q] = r >= 0 ? __builtin_amdgcn_readfirstlane(ptr[r]) : 0;
    len[q] = r >= 0 ? __builtin_amdgcn_readfirstlane(ptr[r + 1]) - e0[q] : 0;
    maxLen = max(maxLen, len[q]);
  }
  Acc a0 = 0, a1 = 0;
  for (int k = 0; k < maxLen; k += CH) {
    int c[CH];
    Acc v[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      c[u] = 0;
      v[u] = 0;
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      if (k < len[q]) {
        int cq[CH];
        Acc vq[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          cq[u] = idx[e0[q] + k + u];
          const Acc w = static_cast<Acc>(val[e0[q] + k + u]);
          vq[u] = k + u < len[q] ? w : static_cast<Acc>(0);
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          c[u] = g == q ? cq[u] : c[u];
          v[u] = g == q ? vq[u] : v[u];
        }
      }
    }
    Acc gth[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) gth[u] = static_cast<Acc>(vld<SPLIT, LOGB>(x, c[u], b));
#pragma unroll
    for (int u = 0; u < CH; ++u) (u & 1 ? a1 : a0) += v[u] * gth[u];
  }
  return a0 + a1;
}

// B = 64, two rows per wavefront: row_dot_wave for rows [a0, a1) and
// [c0, c1) with their chunks interleaved, so each wave keeps two chunks of
// gathers in flight.  A row that has run out reads the other row's entries
// with zero weights (at least one of them has entries in every round).
template <typename Acc, bool SPLIT, typename V, typename X>
__device__ inline void row_dot_wave2(const int* __restrict__ idx, const V* __restrict__ val, const X* __restrict__ x,
                                     int a0, int a1, int c0, int c1, int b, Acc& outA, Acc& outC) {
  constexpr int CH = 8;
  a0 = __builtin_amdgcn_readfirstlane(a0);
  a1 = __builtin_amdgcn_readfirstlane(a1);
  c0 = __builtin_amdgcn_readfirstlane(c0);
  c1 = __builtin_amdgcn_readfirstlane(c1);
  Acc sa[2] = {0, 0}, sc[2] = {0, 0};
  for (int ea = a0, ec = c0; ea < a1 || ec < c1; ea += CH, ec += CH) {
    const bool ha = ea < a1, hc = ec < c1;
    const int fa = ha ? ea : ec, la = ha ? a1 - 1 : c1 - 1;  // entries read for row a
    const int fc = hc ? ec : ea, lc = hc ? c1 - 1 : a1 - 1;  // entries read for row c
    int ca[CH], cc[CH];
    Acc va[CH], vc[CH], ga[CH], gc[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int pa = min(fa + u, la), pc = min(fc + u, lc);
      ca[u] = idx[pa];
      cc[u] = idx[pc];
      const Acc wa = static_cast<Acc>(val[pa]), wc = static_cast<Acc>(val[pc]);
      va[u] = ha && ea + u < a1 ? wa : static_cast<Acc>(0);
      vc[u] = hc && ec + u < c1 ? wc : static_cast<Acc>(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      ga[u] = static_cast<Acc>(vld<SPLIT, 6>(x, ca[u], b));
      gc[u] = static_cast<Acc>(vld<SPLIT, 6>(x, cc[u], b));
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      sa[u & 1] += va[u] * ga[u];
      sc[u & 1] += vc[u] * gc[u];
    }
  }
  outA = sa[0] + sa[1];
  outC = sc[0] + sc[1];
}

// Row dot product for the lane layout of LOGB (the B = 64 case uses the
// wave-uniform path above; there the plan always has one lane per instance).
template <int LOGB, typename Acc, bool SPLIT = false, typename V, typename X>
__device__ inline Acc row_sum(const int* __restrict__ ptr, const int* __restrict__ idx, const V* __restrict__ val,
                              const X* __restrict__ x, const Slot& s, int b) {
  if constexpr (LOGB == 6) {
    return row_dot_wave<Acc, SPLIT>(idx, val, x, ptr[s.row], ptr[s.row + 1], b);
  } else {
    return row_dot<LOGB, Acc, SPLIT>(idx, val, x, ptr[s.row] + s.sub, ptr[s.row + 1], 1 << s.logS, b);
  }
}

// The row dot product of lane b under plan p (block bid): the grouped form
// when the plan asks for it, else row_sum (0 for a lane without a row).
template <int LOGB, typename Acc, bool SPLIT = false, typename V, typename X>
__device__ inline Acc row_sum_plan(const RowPlan& p, int bid, const int* __restrict__ ptr, const int* __restrict__ idx,
                                   const V* __restrict__ val, const X* __restrict__ x, const Slot& s, int b) {
  if constexpr (LOGB == 4 || LOGB == 5) {
    if (p.grp) {
      constexpr int G = 64 >> LOGB;
      const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
      const int base = ((bid - p.blkStart[0]) << (kLogWG - LOGB)) + wave * G;  // locate() with one lane per row
      int rows[G];
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const int r = base + q;
        rows[q] = r < p.rowCount[0] ? (p.perm ? p.perm[p.rowStart[0] + r] : p.rowStart[0] + r) : -1;
      }
      return row_dot_groups<LOGB, Acc, SPLIT>(ptr, idx, val, x, rows, b);
    }
  }
  return s.row >= 0 ? row_sum<LOGB, Acc, SPLIT>(ptr, idx, val, x, s, b) : static_cast<Acc>(0);
}

// The rows of lane (b, sub) under plan p and their (sub-lane reduced) dot
// products: one row (rows[1] = -1), or at B = 64 with two rows per
// wavefront the wave's two rows, their chunks interleaved.
template <int LOGB, int RPW, typename Acc, bool SPLIT = false, typename V, typename X>
__device__ inline void lane_rows(const RowPlan& p, int bid, const Slot& s, const int* __restrict__ ptr,
                                 const int* __restrict__ idx, const V* __restrict__ val, const X* __restrict__ x, int b,
                                 int (&rows)[2], Acc (&acc)[2]) {
  if constexpr (LOGB == 6 && RPW == 2) {
    {  // one bucket, rows in order (no permutation at B = 64)
      const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
      const int r0 = ((bid - p.blkStart[0]) * kWaves + wave) * 2;
      rows[0] = r0 < p.rowCount[0] ? r0 : -1;
      rows[1] = r0 + 1 < p.rowCount[0] ? r0 + 1 : -1;
      const int a0 = rows[0] >= 0 ? ptr[rows[0]] : 0, a1 = rows[0] >= 0 ? ptr[rows[0] + 1] : 0;
      const int c0 = rows[1] >= 0 ? ptr[rows[1]] : 0, c1 = rows[1] >= 0 ? ptr[rows[1] + 1] : 0;
      row_dot_wave2<Acc, SPLIT>(idx, val, x, a0, a1, c0, c1, b, acc[0], acc[1]);
      return;
    }
  }
  rows[0] = s.row;
  rows[1] = -1;
  acc[0] = reduce_sub<LOGB>(row_sum_plan<LOGB, Acc, SPLIT>(p, bid, ptr, idx, val, x, s, b), s.logS);
  acc[1] = 0;
}

// Per-workgroup FP64 partial sums of NQ quantities for each instance
// b = threadIdx % B, stored at out[((q0 + q) * B + b) * nblk + blk].
template <int LOGB, int NQ>
__device__ inline void block_partials(const double (&v)[NQ], double* __restrict__ out, int q0, int nblk, int blk) {
  constexpr int B = 1 << LOGB;
  __shared__ double sh[NQ][kWG / 64][B];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double s = v[q];
    for (int o = 32; o >= B; o >>= 1) s += __shfl_xor(s, o);
    if (lane < B) sh[q][wave][lane] = s;
  }
  __syncthreads();
  if (threadIdx.x < B) {
    const int b = threadIdx.x;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double t = sh[q][0][b];
      for (int w = 1; w < kWaves; ++w) t += sh[q][w][b];
      out[(static_cast<size_t>(q0 + q) * B + b) * nblk + blk] = t;
    }
  }
}

template <typename T>
__device__ inline T clamp_to(T v, T lo, T hi) {
  v = v > lo ? v : lo;  // max(v, lo) then min(., hi): src/pdhg/pdhg.cpp:121-122
  return v < hi ? v : hi;
}

// ---------------------------------------------------------------------------
// Per-instance state (device arrays of length B)
// ---------------------------------------------------------------------------
struct Inst {
  double *tau, *sigma, *omega, *fpR, *fpPrev;
  long long *iter, *since, *restartAt, *applyAt, *doneIter;
  int *done, *restarts;
  const double *cNorm, *rhsNorm;
  CheckRecord* rec;
};

// Stride of a per-instance (ld = B, mask = B-1) or shared (ld = 1, mask = 0) array.
struct Ld {
  int ld, mask;
  __device__ size_t at(int i, int b) const { return static_cast<size_t>(i) * ld + (b & mask); }
};

// Case data of row i for lane b: one value shared by all lanes (mask 0) or
// one per lane in the run's vector layout.
template <bool SPLIT, int LOGB, typename T>
__device__ inline T cld(const T* __restrict__ p, const Ld& l, int i, int b) {
  return l.mask ? vld<SPLIT, LOGB>(p, i, b) : p[i];
}

// Check logic of one instance from its nine fixed-order sums S (column side:
// ||x - x+||^2, ||x+ - x0||^2, dual residual^2, c'x+, column dual objective;
// row side: primal residual^2, row dual objective, ||y - y+||^2,
// ||y+ - y0||^2), all FP64 (reference pdhg.cpp:604-656): relative KKT,
// convergence, the restart rule, the primal-weight update and new steps.
// A case is checked on its own iteration count: every checkEvery iterations
// and at its iteration limit (pdhg.cpp:593).  A block that ends elsewhere for
// this case (cases of a continuous batch start at different run iterations)
// only advances its count.
__device__ inline void finalize_instance(Inst st, int b, const double (&S)[9], const RunScalars& sc, int blockIters,
                                         int checkEvery) {
  const double dx2 = S[0], ax2 = S[1], dres2 = S[2], pobjS = S[3], dobjC = S[4];
  const double pres2 = S[5], dobjR = S[6], dy2 = S[7], ay2 = S[8];
  const long long it = st.iter[b] + blockIters - 1;
  if ((it + 1) % checkEvery != 0 && it + 1 < sc.maxIter) {
    st.iter[b] += blockIters;
    return;
  }
  const double so = sc.beta * sc.gamma;
  const double pobj = pobjS * so;
  const double dobj = (dobjR + dobjC) * so;
  const double relP = sqrt(pres2) / (1.0 + st.rhsNorm[b]);
  const double relD = sqrt(dres2) / (1.0 + st.cNorm[b]);
  const double gap = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
  double omega = st.omega[b];
  CheckRecord r;
  r.it = static_cast<double>(it);
  r.relP = relP;
  r.relD = relD;
  r.gap = gap;
  r.pobj = pobj + sc.offset;
  r.dobj = dobj + sc.offset;
  r.omega = omega;
  r.restarted = 0;
  st.iter[b] += blockIters;
  if (relP <= sc.tol && relD <= sc.tol && gap <= sc.tol) {
    st.done[b] = 1;
    st.doneIter[b] = it;
    st.rec[b] = r;
    return;
  }
  long long since = st.since[b] + checkEvery;
  const double fp = sqrt(omega * dx2 + dy2 / omega);
  double fpR = st.fpR[b];
  if (isinf(fpR)) fpR = fp;
  const bool restart = fp <= 0.2 * fpR || (fp <= 0.8 * fpR && fp > st.fpPrev[b]) ||
                       since >= static_cast<long long>(0.36 * static_cast<double>(it + 1));
  double fpPrev = fp;
  if (restart) {
    const double dxn = sqrt(ax2), dyn = sqrt(ay2);
    if (dxn > 1e-10 && dyn > 1e-10) omega = exp(0.5 * log(dyn / dxn) + 0.5 * log(omega));
    fpR = sqrt(omega * dx2 + dy2 / omega);
    fpPrev = INFINITY;
    since = 0;
    st.restarts[b] += 1;
    st.omega[b] = omega;
    st.tau[b] = sc.eta / omega;
    st.sigma[b] = sc.eta * omega;
    st.restartAt[b] = it + 1;
    st.applyAt[b] = it + 1;
    r.restarted = 1;
    r.omega = omega;
  }
  st.since[b] = since;
  st.fpR[b] = fpR;
  st.fpPrev[b] = fpPrev;
  st.rec[b] = r;
}

// Host helpers
int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// Device buffer list owned by a run; freed together.
struct Buffers {
  std::vector<void*> ptrs;
  ~Buffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T* alloc(size_t count) {
    void* p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
};

// Host-side row mapping of one CSR matrix for B instances.

}  // namespace

}  // namespace pdhgamd
