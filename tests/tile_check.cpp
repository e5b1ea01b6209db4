// Invariants of the width-64 row tiles (tiles.h), checked on the host:
// every row in exactly one tile, at most `rows` rows (one per wavefront of the
// tile shape) and maxCols distinct columns per tile, loc[e] naming entry e's column among its tile's columns,
// the refusal of a matrix with a row longer than maxCols, and the
// per-(tile, wavefront) records: each wavefront's row and entry range, and
// every tile column staged by exactly one wavefront at its tile position;
// and the sliced ELLPACK copy of the width-1 SpMV.
// Usage: tile_check [matrix.bin ...]   (int32 rows, cols, nnz, ptr, idx)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>

#include "tiles.h"

using pdhgamd::Csr;
using pdhgamd::TileSet;

static bool check(const Csr& M, int maxCols, int rows, const char* name) {
  TileSet t;
  int maxLen = 0;
  for (int i = 0; i < M.rows; ++i) maxLen = std::max(maxLen, M.ptr[i + 1] - M.ptr[i]);
  const bool got = pdhgamd::build_tiles(M, maxCols, t, rows);
  if (got != (maxLen <= maxCols)) {
    std::printf("%s: build_tiles returned %d with longest row %d, budget %d\n", name, got, maxLen, maxCols);
    return false;
  }
  if (!got) return true;
  if (static_cast<int>(t.tileRow.size()) != t.tiles + 1 || t.tileRow[0] != 0 || t.tileRow[t.tiles] != M.rows ||
      t.tileCol[0] != 0 || t.tileCol[t.tiles] != static_cast<int>(t.ucols.size()) ||
      static_cast<int>(t.loc.size()) != M.nnz() + pdhgamd::kCsrPad) {
    std::printf("%s: bad tile bounds\n", name);
    return false;
  }
  std::vector<int> seen(M.rows, 0);
  for (int k = 0; k < t.tiles; ++k) {
    const int r0 = t.tileRow[k], r1 = t.tileRow[k + 1], u0 = t.tileCol[k], u1 = t.tileCol[k + 1];
    if (r1 <= r0 && M.rows > 0) {
      std::printf("%s: empty tile %d\n", name, k);
      return false;
    }
    if (r1 - r0 > rows || u1 - u0 > maxCols || u1 < u0) {
      std::printf("%s: tile %d has %d rows, %d columns\n", name, k, r1 - r0, u1 - u0);
      return false;
    }
    std::vector<int> cols(t.ucols.begin() + u0, t.ucols.begin() + u1);
    std::sort(cols.begin(), cols.end());
    if (std::adjacent_find(cols.begin(), cols.end()) != cols.end()) {
      std::printf("%s: tile %d lists a column twice\n", name, k);
      return false;
    }
    for (int r = r0; r < r1; ++r) {
      const int i = t.perm[r];
      if (i < 0 || i >= M.rows || seen[i]++) {
        std::printf("%s: row %d placed twice or out of range\n", name, i);
        return false;
      }
      for (int e = M.ptr[i]; e < M.ptr[i + 1]; ++e) {
        const int l = t.loc[e];
        if (l < 0 || l >= u1 - u0 || t.ucols[u0 + l] != M.idx[e]) {
          std::printf("%s: entry %d of row %d maps to a wrong tile column\n", name, e, i);
          return false;
        }
      }
    }
  }
  for (int i = 0; i < M.rows; ++i)
    if (seen[i] != 1) {
      std::printf("%s: row %d in no tile\n", name, i);
      return false;
    }
  const int waves = rows;
  const int stageCols = (maxCols + waves - 1) / waves;
  const int stride = pdhgamd::desc_stride(stageCols);
  const std::vector<int> d = pdhgamd::tile_descriptors(M, t, stageCols, waves);
  if (d.size() != static_cast<size_t>(t.tiles) * waves * stride || stride % 8 != 0) {
    std::printf("%s: descriptor size\n", name);
    return false;
  }
  for (int k = 0; k < t.tiles; ++k) {
    const int r0 = t.tileRow[k], r1 = t.tileRow[k + 1], u0 = t.tileCol[k], u1 = t.tileCol[k + 1];
    std::vector<int> staged(u1 - u0, 0);
    for (int w = 0; w < waves; ++w) {
      const int* o = &d[(static_cast<size_t>(k) * waves + w) * stride];
      const int j = r0 + w < r1 ? t.perm[r0 + w] : -1;
      if (o[0] != j || (j >= 0 && (o[1] != M.ptr[j] || o[2] != M.ptr[j + 1])) || o[3] < 0 || o[3] > stageCols) {
        std::printf("%s: tile %d wave %d record\n", name, k, w);
        return false;
      }
      for (int q = 0; q < o[3]; ++q) {
        const int pos = w + q * waves;
        if (pos >= u1 - u0 || o[4 + q] != t.ucols[u0 + pos]) {
          std::printf("%s: tile %d wave %d stages a wrong column\n", name, k, w);
          return false;
        }
        staged[pos]++;
      }
    }
    for (int pos = 0; pos < u1 - u0; ++pos)
      if (staged[pos] != 1) {
        std::printf("%s: tile %d column %d staged %d times\n", name, k, pos, staged[pos]);
        return false;
      }
  }
  return true;
}

// The sliced ELLPACK copy (build_sell): perm a permutation of the rows,
// slice offsets in steps of 64 x the slice's longest row, every row's entries
// in its lane in CSR order, the rest of its lane padded with weight 0 and a
// column of the row (or 0), lanes past the last row empty; rows reordered
// only when their own order pads by more than a quarter.
static bool check_sell(const Csr& M, const char* name) {
  const pdhgamd::SellHost h = pdhgamd::build_sell(M);
  const int rows = M.rows;
  auto len = [&](int i) { return M.ptr[i + 1] - M.ptr[i]; };
  if (h.slices != (rows + 63) / 64 || static_cast<int>(h.soff.size()) != h.slices + 1 || h.soff[0] != 0 ||
      static_cast<int>(h.perm.size()) != rows || h.sidx.size() != h.sval.size() ||
      static_cast<long long>(h.sidx.size()) != h.soff[h.slices]) {
    std::printf("%s: sliced copy sizes\n", name);
    return false;
  }
  std::vector<int> seen(rows, 0);
  bool identity = true;
  long long natural = 0;
  for (int s0 = 0; s0 < rows; s0 += 64) {
    int w = 0;
    for (int i = s0; i < std::min(rows, s0 + 64); ++i) w = std::max(w, len(i));
    natural += 64LL * w;
  }
  for (int s = 0; s < h.slices; ++s) {
    const int off = h.soff[s], w = (h.soff[s + 1] - off) / 64;
    if ((h.soff[s + 1] - off) % 64 != 0) {
      std::printf("%s: slice %d size\n", name, s);
      return false;
    }
    int longest = 0;
    for (int lane = 0; lane < 64; ++lane) {
      const int pos = 64 * s + lane;
      const int r = pos < rows ? h.perm[pos] : -1;
      if (pos < rows && (r < 0 || r >= rows || seen[r]++)) {
        std::printf("%s: perm is not a permutation\n", name);
        return false;
      }
      identity = identity && (pos >= rows || r == pos);
      const int n = r >= 0 ? len(r) : 0;
      longest = std::max(longest, n);
      for (int k = 0; k < w; ++k) {
        const size_t at = static_cast<size_t>(off) + 64 * static_cast<size_t>(k) + lane;
        const bool entry = k < n;
        const int wantIdx = entry ? M.idx[M.ptr[r] + k] : (n > 0 ? M.idx[M.ptr[r] + n - 1] : 0);
        const double wantVal = entry ? M.val[M.ptr[r] + k] : 0.0;
        if (h.sidx[at] != wantIdx || h.sval[at] != wantVal) {
          std::printf("%s: slice %d lane %d entry %d\n", name, s, lane, k);
          return false;
        }
      }
    }
    if (longest != w) {
      std::printf("%s: slice %d width %d, longest row %d\n", name, s, w, longest);
      return false;
    }
  }
  // The order expected: the rows' own, or by length (longest first, stable)
  // within windows of 512 when their own order pads by more than a quarter.
  std::vector<int> want(rows);
  for (int i = 0; i < rows; ++i) want[i] = i;
  if (4 * natural > 5LL * std::max(1, M.nnz()))
    for (int w0 = 0; w0 < rows; w0 += 512)
      std::stable_sort(want.begin() + w0, want.begin() + std::min(rows, w0 + 512),
                       [&](int x, int y) { return len(x) > len(y); });
  if (want != h.perm) {
    std::printf("%s: row order of the sliced copy (identity %d)\n", name, identity ? 1 : 0);
    return false;
  }
  return true;
}

static Csr load(const char* f) {
  std::ifstream in(f, std::ios::binary);
  Csr M;
  int nnz = 0;
  in.read(reinterpret_cast<char*>(&M.rows), 4);
  in.read(reinterpret_cast<char*>(&M.cols), 4);
  in.read(reinterpret_cast<char*>(&nnz), 4);
  M.ptr.resize(M.rows + 1);
  M.idx.resize(nnz);
  in.read(reinterpret_cast<char*>(M.ptr.data()), 4 * (M.rows + 1));
  in.read(reinterpret_cast<char*>(M.idx.data()), 4 * static_cast<size_t>(nnz));
  M.val.resize(nnz);
  for (int e = 0; e < nnz; ++e) M.val[e] = 1.0 + e;  // distinct, so misplaced values show
  return M;
}

int main(int argc, char** argv) {
  bool ok = true;
  int files = 0, random = 0;
  for (int a = 1; a < argc; ++a) {
    const Csr M = load(argv[a]);
    // Each tile shape's FP64 and FP32 budgets (tiles.h), then small budgets.
    for (int shape = 1; shape <= pdhgamd::kTileShapes; ++shape)
      for (int bytes : {8, 4})
        ok = check(M, pdhgamd::tile_bytes(shape) / (64 * bytes), pdhgamd::tile_waves(shape), argv[a]) && ok;
    for (int mc : {64, 45}) ok = check(M, mc, pdhgamd::kTileRows, argv[a]) && ok;
    ok = check_sell(M, argv[a]) && ok;
    ++files;
  }
  std::mt19937 rng(1);
  for (int trial = 0; trial < 300; ++trial) {
    Csr M;
    M.rows = static_cast<int>(rng() % 70);
    M.cols = 1 + static_cast<int>(rng() % 90);
    std::vector<int> c(M.cols);
    for (int i = 0; i < M.rows; ++i) {
      const int len = static_cast<int>(rng() % std::min(M.cols + 1, 24));
      for (int j = 0; j < M.cols; ++j) c[j] = j;
      for (int k = 0; k < len; ++k) std::swap(c[k], c[k + rng() % (M.cols - k)]);
      std::sort(c.begin(), c.begin() + len);
      M.idx.insert(M.idx.end(), c.begin(), c.begin() + len);
      M.ptr.push_back(static_cast<int>(M.idx.size()));
    }
    M.val.resize(M.idx.size());
    for (size_t e = 0; e < M.val.size(); ++e) M.val[e] = 1.0 + static_cast<double>(e);
    ok = check(M, 1 + static_cast<int>(rng() % 30), pdhgamd::tile_waves(1 + trial % pdhgamd::kTileShapes), "random") &&
         ok;
    ok = check_sell(M, "random") && ok;
    ++random;
  }
  {  // uneven rows (mostly 1 entry, every 37th row 200): the sliced copy reorders them
    Csr M;
    M.rows = 2000;
    M.cols = 500;
    for (int i = 0; i < M.rows; ++i) {
      const int len = i % 37 == 0 ? 200 : 1;
      for (int k = 0; k < len; ++k) M.idx.push_back((i + 3 * k) % M.cols);
      std::sort(M.idx.end() - len, M.idx.end());
      M.idx.erase(std::unique(M.idx.end() - len, M.idx.end()), M.idx.end());
      M.ptr.push_back(static_cast<int>(M.idx.size()));
    }
    M.val.resize(M.idx.size());
    for (size_t e = 0; e < M.val.size(); ++e) M.val[e] = 1.0 + static_cast<double>(e);
    ok = check_sell(M, "uneven") && ok;
  }
  std::printf("%s: %d matrices from files, %d random\n", ok ? "ok" : "FAILED", files, random);
  return ok ? 0 : 1;
}
