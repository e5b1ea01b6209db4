// Host-side core types of the PDHG engine: the LP model (CSC), CSR views,
// the deterministic SplitMix64 generator and the error type.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace pdhgamd {

constexpr double kInf = std::numeric_limits<double>::infinity();
// |v| >= 1e20 is an infinite bound (reference src/util/common.h:13-17).
constexpr double kInfBound = 1e20;
inline bool finite_bound(double v) { return std::fabs(v) < kInfBound; }
inline double sq(double v) { return v * v; }

class Error : public std::runtime_error {
 public:
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

#define PDHG_REQUIRE(cond, msg)                                                     \
  do {                                                                              \
    if (!(cond)) throw ::pdhgamd::Error(std::string(msg) + " [" #cond "]");         \
  } while (0)

// SplitMix64 (the reference uses the same generator, src/util/rng.h:10-24, so
// streams are identical on every platform).
class Rng {
 public:
  explicit Rng(uint64_t seed) : s_(seed) {}
  uint64_t next() {
    uint64_t z = (s_ += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
  int below(int n) { return static_cast<int>(next() % static_cast<uint64_t>(n)); }
  // Standard normal by the Box-Muller transform (one value per call).
  double normal() {
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }

 private:
  uint64_t s_;
};

// Compressed sparse column matrix (reference src/core/sparse.h:10-16).
struct Csc {
  int rows = 0, cols = 0;
  std::vector<int> start{0};
  std::vector<int> index;
  std::vector<double> value;
  int nnz() const { return start.empty() ? 0 : start.back(); }
};

// Compressed sparse row matrix (the layout both SpMV kernels stream).
struct Csr {
  int rows = 0, cols = 0;
  std::vector<int> ptr{0};
  std::vector<int> idx;
  std::vector<double> val;
  int nnz() const { return ptr.empty() ? 0 : ptr.back(); }
};

// LP in minimization form (reference src/core/model.h:24-64, LP part only).
struct Model {
  int m = 0, n = 0;
  Csc A;
  std::vector<double> c, l, u, L, U;
  double offset = 0;
  // Known optimum of generated LPs (empty otherwise).
  std::vector<double> xstar, ystar;
  // Refinery family metadata: purchase columns of the first four crudes.
  std::vector<std::vector<int>> crudeCols;
  std::vector<double> crudePrice;
};

// Transpose of a CSC matrix, i.e. the CSR of the matrix, rows sorted by
// column (reference src/core/sparse.cpp:59-75).
Csr csr_of(const Csc& a);
// The CSC of A read as the CSR of A^T (reference src/pdhg/pdhg.cpp:35-43).
Csr csr_of_transpose(const Csc& a);

}  // namespace pdhgamd
