// Invariants of the width-64 row tiles (tiles.h), checked on the host:
// every row in exactly one tile, at most `rows` rows (one per wavefront of the
// tile shape) and maxCols distinct columns per tile, loc[e] naming entry e's column among its tile's columns,
// the refusal of a matrix with a row longer than maxCols, and the
// per-(tile, wavefront) records: each wavefront's row and entry range, and
// every tile column staged by exactly one wavefront at its tile position.
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
  M.val.assign(nnz, 1.0);
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
    M.val.assign(M.idx.size(), 1.0);
    ok = check(M, 1 + static_cast<int>(rng() % 30), pdhgamd::tile_waves(1 + trial % pdhgamd::kTileShapes), "random") &&
         ok;
    ++random;
  }
  std::printf("%s: %d matrices from files, %d random\n", ok ? "ok" : "FAILED", files, random);
  return ok ? 0 : 1;
}
