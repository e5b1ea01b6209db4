// Row tiles for the width-64 iteration kernels staged in LDS (host side):
// rows in locality order, cut into tiles whose distinct columns fit the
// workgroup's LDS, and every entry's column numbered within its tile; and the
// sliced ELLPACK copy of a matrix for the width-1 SpMV.  Host-only: the
// kernels include it for the constants.
#pragma once

#include <algorithm>
#include <vector>

#include "core.h"

namespace pdhgamd {

// A workgroup of W wavefronts (one row each) per tile, and the LDS it stages
// the tile's distinct columns in.  Three shapes, each filling a CU's 32
// wavefront slots (8 per SIMD) within its 160 KB of LDS:
//   shape 1: 16 wavefronts, 64 KB of columns (128 FP64), 2 workgroups per CU;
//   shape 2:  8 wavefronts, 32 KB (64 FP64), 4 workgroups per CU;
//   shape 3:  4 wavefronts, 16 KB (32 FP64), 8 workgroups per CU.
// Smaller tiles share fewer columns but let more workgroups, at different
// points of their load / barrier / sum sequence, share a CU.
constexpr int kTileShapes = 3;
constexpr int tile_waves(int shape) { return shape == 3 ? 4 : shape == 2 ? 8 : 16; }
constexpr int tile_bytes(int shape) { return shape == 3 ? 16 * 1024 : shape == 2 ? 32 * 1024 : 64 * 1024; }
template <typename T, int SHAPE = 1>
constexpr int kTileCols = tile_bytes(SHAPE) / (64 * static_cast<int>(sizeof(T)));
// Shape 1, the largest.
constexpr int kTileWaves = tile_waves(1);
constexpr int kTileWG = 64 * kTileWaves;
constexpr int kTileRows = kTileWaves;

// Zero entries (column 0, weight 0) after the last nonzero of the CSR arrays
// the iteration kernels read: whole-chunk scalar loads may run up to a
// chunk past a row's end.
constexpr int kCsrPad = 64;

template <typename U>
std::vector<U> csr_padded(const std::vector<U>& v) {
  std::vector<U> out(v);
  out.resize(v.size() + kCsrPad, U(0));
  return out;
}

// Tiles over the rows of a CSR matrix: tile t holds rows perm[tileRow[t] ..
// tileRow[t+1]) and their distinct columns ucols[tileCol[t] .. tileCol[t+1]);
// loc[e] is entry e's position among its tile's columns (padded like the CSR
// arrays).
struct TileSet {
  std::vector<int> perm, tileRow, tileCol, ucols, loc;
  int tiles = 0;
};

// Tiles of at most maxRows rows and maxCols distinct columns over the rows
// in locality order (by the column of their middle entry), so that rows of a
// tile share columns.  False when a row alone has more than maxCols entries.
inline bool build_tiles(const Csr& M, int maxCols, TileSet& t, int maxRows = kTileRows) {
  const int rows = M.rows;
  for (int i = 0; i < rows; ++i)
    if (M.ptr[i + 1] - M.ptr[i] > maxCols) return false;
  std::vector<int> key(rows);
  for (int i = 0; i < rows; ++i) {
    const int len = M.ptr[i + 1] - M.ptr[i];
    key[i] = len ? M.idx[M.ptr[i] + len / 2] : 0;
  }
  t.perm.resize(rows);
  for (int i = 0; i < rows; ++i) t.perm[i] = i;
  std::stable_sort(t.perm.begin(), t.perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  t.loc.assign(static_cast<size_t>(M.nnz()) + kCsrPad, 0);
  t.tileRow.assign(1, 0);
  t.tileCol.assign(1, 0);
  t.ucols.clear();
  std::vector<int> mark(M.cols, -1), pos(M.cols, 0);
  int tile = 0, inTile = 0;
  for (int r = 0; r < rows; ++r) {
    const int i = t.perm[r];
    for (int pass = 0; pass < 2; ++pass) {
      const size_t keep = t.ucols.size();
      bool fits = inTile < maxRows;
      for (int e = M.ptr[i]; fits && e < M.ptr[i + 1]; ++e) {
        const int c = M.idx[e];
        if (mark[c] == tile) continue;
        mark[c] = tile;
        pos[c] = static_cast<int>(t.ucols.size()) - t.tileCol.back();
        t.ucols.push_back(c);
        fits = pos[c] < maxCols;
      }
      if (fits) break;
      // Close the tile (dropping this row's new columns) and retry in a new one.
      for (size_t u = keep; u < t.ucols.size(); ++u) mark[t.ucols[u]] = -1;
      t.ucols.resize(keep);
      t.tileRow.push_back(r);
      t.tileCol.push_back(static_cast<int>(t.ucols.size()));
      ++tile;
      inTile = 0;
    }
    for (int e = M.ptr[i]; e < M.ptr[i + 1]; ++e) t.loc[e] = pos[M.idx[e]];
    ++inTile;
  }
  if (inTile > 0 || rows == 0) {
    t.tileRow.push_back(rows);
    t.tileCol.push_back(static_cast<int>(t.ucols.size()));
  }
  t.tiles = static_cast<int>(t.tileRow.size()) - 1;
  return t.tiles > 0;
}

// One record per (tile, wavefront) for the tiled kernels, so that a
// wavefront gets everything it needs before its gathers in one round of
// scalar loads: [0] its row (-1: none), [1], [2] the row's entry range,
// [3] the number of tile columns it stages, [4 + k] the k-th of them (tile
// position wave + k * waves, the matrix column ucols[tileCol[t] +
// position]).  Records are `desc_stride(stageCols)` ints apart, a multiple of
// eight so that each starts on a 32-byte boundary.  The tiles hold at most
// `waves` rows (one per wavefront).
constexpr int desc_stride(int stageCols) { return (4 + stageCols + 7) / 8 * 8; }

inline std::vector<int> tile_descriptors(const Csr& M, const TileSet& t, int stageCols, int waves = kTileWaves) {
  const int stride = desc_stride(stageCols);
  std::vector<int> d(static_cast<size_t>(t.tiles) * waves * stride, 0);
  for (int tt = 0; tt < t.tiles; ++tt) {
    const int r0 = t.tileRow[tt], r1 = t.tileRow[tt + 1], u0 = t.tileCol[tt], u1 = t.tileCol[tt + 1];
    for (int w = 0; w < waves; ++w) {
      int* o = &d[(static_cast<size_t>(tt) * waves + w) * stride];
      const int j = r0 + w < r1 ? t.perm[r0 + w] : -1;
      o[0] = j;
      o[1] = j >= 0 ? M.ptr[j] : 0;
      o[2] = j >= 0 ? M.ptr[j + 1] : 0;
      int ns = 0;
      for (int k = 0; k < stageCols && u0 + w + k * waves < u1; ++k) o[4 + ns++] = t.ucols[u0 + w + k * waves];
      o[3] = ns;
    }
  }
  return d;
}

// Sliced ELLPACK copy of a CSR matrix for the width-1 SpMV (k_spmv_sell): slice s holds rows
// perm[64 s .. 64 s + 63], its width is its longest row, and padding repeats
// the row's last column (or column 0) with weight 0.  Rows keep their order
// (neighbours share columns, so they gather together) unless that pads the
// copy by more than a quarter; then rows are ordered by length (longest
// first) within windows of 512, so a slice's rows have similar lengths and
// still come from nearby rows.
struct SellHost {
  std::vector<int> soff, sidx, perm;
  std::vector<double> sval;
  int slices = 0;
};

inline SellHost build_sell(const Csr& M) {
  constexpr int kWindow = 512;
  SellHost h;
  const int rows = M.rows;
  h.perm.resize(rows);
  for (int i = 0; i < rows; ++i) h.perm[i] = i;
  auto len = [&](int i) { return M.ptr[i + 1] - M.ptr[i]; };
  auto padded = [&] {
    long long total = 0;
    for (int s0 = 0; s0 < rows; s0 += 64) {
      int w = 0;
      for (int i = s0; i < std::min(rows, s0 + 64); ++i) w = std::max(w, len(h.perm[i]));
      total += 64LL * w;
    }
    return total;
  };
  if (4 * padded() > 5LL * std::max(1, M.nnz()))
    for (int w0 = 0; w0 < rows; w0 += kWindow)
      std::stable_sort(h.perm.begin() + w0, h.perm.begin() + std::min(rows, w0 + kWindow),
                       [&](int a, int b) { return len(a) > len(b); });
  h.slices = (rows + 63) / 64;
  h.soff.assign(h.slices + 1, 0);
  long long total = 0;
  for (int s = 0; s < h.slices; ++s) {
    int w = 0;
    for (int i = 64 * s; i < std::min(rows, 64 * s + 64); ++i) w = std::max(w, len(h.perm[i]));
    total += 64LL * w;
    PDHG_REQUIRE(total < (1LL << 31), "sliced matrix too large for 32-bit offsets");
    h.soff[s + 1] = static_cast<int>(total);
  }
  h.sidx.assign(static_cast<size_t>(total), 0);
  h.sval.assign(static_cast<size_t>(total), 0.0);
  for (int s = 0; s < h.slices; ++s) {
    const int w = (h.soff[s + 1] - h.soff[s]) / 64;
    for (int lane = 0; lane < 64 && 64 * s + lane < rows; ++lane) {
      const int r = h.perm[64 * s + lane], e0 = M.ptr[r], n = len(r);
      for (int k = 0; k < w; ++k) {
        const size_t at = static_cast<size_t>(h.soff[s]) + 64 * static_cast<size_t>(k) + lane;
        h.sidx[at] = k < n ? M.idx[e0 + k] : (n > 0 ? M.idx[e0 + n - 1] : 0);
        h.sval[at] = k < n ? M.val[e0 + k] : 0.0;
      }
    }
  }
  return h;
}

}  // namespace pdhgamd
