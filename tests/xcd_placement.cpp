// Host model of the L2 fills of the gathered vector in one width-64 FP64
// launch: workgroup w runs on XCD w mod 8 (the hardware's round-robin), each
// XCD has its own L2, so a column gathered by workgroups on k XCDs is fetched
// k times.  Counts the distinct (XCD, column) pairs, with and without the
// remap that gives each XCD a contiguous range of logical blocks
// (remap_xcd, csrc/kcommon.h), for the LDS tiles (one tile per workgroup,
// tiles.h) and for one row per wavefront (four rows per 256-thread block,
// natural row order).
// Usage: xcd_placement matrix.bin ...   (int32 rows, cols, nnz, ptr, idx)
#include <cstdio>
#include <fstream>
#include <vector>

#include "tiles.h"

using pdhgamd::Csr;
using pdhgamd::TileSet;

// The logical block that hardware block `bid` of `nblk` serves under the
// remap (restated from remap_xcd in csrc/kcommon.h).
static int remap(int bid, int nblk) {
  const int q = nblk >> 3, r = nblk & 7, x = bid & 7, l = bid >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + l;
}

// Distinct (XCD, column) pairs when logical block L gathers cols(L).
template <typename Cols>
static long long pairs(int nblk, int ncols, bool rm, Cols cols) {
  std::vector<int> xcdOf(nblk);
  std::vector<char> hit(nblk, 0);
  for (int b = 0; b < nblk; ++b) {
    const int L = rm ? remap(b, nblk) : b;
    xcdOf[L] = b & 7;
    hit[L] = 1;
  }
  for (int L = 0; L < nblk; ++L)
    if (!hit[L]) return -1;  // the remap is not a bijection
  std::vector<unsigned char> seen(static_cast<size_t>(ncols) * 8, 0);
  long long n = 0;
  for (int L = 0; L < nblk; ++L)
    cols(L, [&](int c) {
      unsigned char& s = seen[static_cast<size_t>(c) * 8 + xcdOf[L]];
      if (!s) {
        s = 1;
        ++n;
      }
    });
  return n;
}

int main(int argc, char** argv) {
  for (int a = 1; a < argc; ++a) {
    std::ifstream f(argv[a], std::ios::binary);
    int rows = 0, cols = 0, nnz = 0;
    f.read(reinterpret_cast<char*>(&rows), 4);
    f.read(reinterpret_cast<char*>(&cols), 4);
    f.read(reinterpret_cast<char*>(&nnz), 4);
    Csr M;
    M.rows = rows;
    M.cols = cols;
    M.ptr.resize(rows + 1);
    M.idx.resize(nnz);
    f.read(reinterpret_cast<char*>(M.ptr.data()), 4LL * (rows + 1));
    f.read(reinterpret_cast<char*>(M.idx.data()), 4LL * nnz);
    if (!f) {
      std::printf("%s: short file\n", argv[a]);
      return 1;
    }
    TileSet t;
    if (!pdhgamd::build_tiles(M, pdhgamd::kTileCols<double>, t)) {
      std::printf("%s: no tiles\n", argv[a]);
      return 1;
    }
    auto tileCols = [&](int L, auto&& use) {
      for (int u = t.tileCol[L]; u < t.tileCol[L + 1]; ++u) use(t.ucols[u]);
    };
    const int rowBlocks = (rows + 3) / 4;
    auto rowCols = [&](int L, auto&& use) {
      for (int r = 4 * L; r < std::min(rows, 4 * L + 4); ++r)
        for (int e = M.ptr[r]; e < M.ptr[r + 1]; ++e) use(M.idx[e]);
    };
    std::vector<char> used(cols, 0);
    long long distinct = 0;
    for (int c : M.idx)
      if (!used[c]) used[c] = 1, ++distinct;
    std::printf("%s distinct %lld tiled_on %lld tiled_off %lld onerow_on %lld onerow_off %lld\n", argv[a], distinct,
                pairs(t.tiles, cols, true, tileCols), pairs(t.tiles, cols, false, tileCols),
                pairs(rowBlocks, cols, true, rowCols), pairs(rowBlocks, cols, false, rowCols));
  }
  return 0;
}
