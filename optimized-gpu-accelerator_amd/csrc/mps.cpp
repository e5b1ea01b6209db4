// MPS reader feeding the engine (the LP part of the reference's readMps,
// src/io/mps.h:15-19, src/io/mps.cpp:40-400, followed by toMinimization,
// src/core/model.cpp:92-101).  Plain or gzip files (zlib reads both).  Data
// lines are whitespace-separated; once a ROWS line has more than two fields
// (names with blanks) the file is read in fixed columns 2-3, 5-12, 15-22,
// 25-36, 40-47, 50-61, as the reference does.  Sections: NAME, OBJSENSE,
// ROWS, COLUMNS (integer MARKER lines skipped: the engine solves the LP
// relaxation), RHS, RANGES, BOUNDS (UP LO FX FR MI PL BV LI UI), ENDATA.
// Numbers may use Fortran D exponents.  Quadratic sections and
// semi-continuous bounds are rejected.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "core.h"

namespace pdhgamd {

namespace {

std::vector<std::string> tokens(const std::string& s) {
  std::vector<std::string> t;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
    size_t j = i;
    while (j < s.size() && !std::isspace(static_cast<unsigned char>(s[j]))) ++j;
    if (j > i) t.emplace_back(s.substr(i, j - i));
    i = j;
  }
  return t;
}

std::string trim(const std::string& s) {
  const size_t b = s.find_first_not_of(" \t");
  if (b == std::string::npos) return "";
  return s.substr(b, s.find_last_not_of(" \t") - b + 1);
}

// The six fixed-format fields of a data line (1-based columns 2-3, 5-12,
// 15-22, 25-36, 40-47, 50-61), trimmed.
std::vector<std::string> fixed_fields(const std::string& line) {
  static const int beg[6] = {1, 4, 14, 24, 39, 49}, len[6] = {2, 8, 8, 12, 8, 12};
  std::vector<std::string> f(6);
  const int L = static_cast<int>(line.size());
  for (int k = 0; k < 6; ++k)
    if (beg[k] < L) f[k] = trim(line.substr(beg[k], std::min(len[k], L - beg[k])));
  return f;
}

std::string upper(std::string s) {
  for (char& ch : s) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
  return s;
}

struct MpsReader {
  std::string path;
  int lineNo = 0;
  // Rows: index >= 0 for constraints, -1 the objective, -2 further N rows
  // (dropped with their entries, as the reference does).
  std::unordered_map<std::string, int> rowIndex, colIndex;
  std::vector<char> rowType;
  std::vector<double> rhs, range;
  std::vector<char> hasRange;
  std::vector<double> cost, lo, up;
  std::vector<char> hasLo;
  std::vector<std::tuple<int, int, double>> trip;  // (column, row, value)
  double offset = 0;
  bool maximize = false;

  [[noreturn]] void fail(const std::string& what) const {
    throw Error("MPS " + path + " line " + std::to_string(lineNo) + ": " + what);
  }
  double num(const std::string& s) const {
    std::string t = s;
    for (char& ch : t)  // Fortran exponents (1.0D+03) occur in legacy files
      if (ch == 'D' || ch == 'd') ch = 'E';
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0') fail("bad number '" + s + "'");
    return v;
  }

  // A fixed-format data line of section `sec` as the whitespace-token form
  // the handlers take (set names always present, possibly empty).
  std::vector<std::string> from_fixed(const std::string& sec, const std::string& line) const {
    const auto f = fixed_fields(line);
    std::vector<std::string> t;
    if (sec == "ROWS") return {f[0], f[1]};
    if (sec == "COLUMNS") {
      if (f[2] == "'MARKER'" || f[2] == "MARKER") return {f[1], f[2], f[3]};
      t = {f[1], f[2], f[3]};
    } else if (sec == "RHS" || sec == "RANGES") {
      t = {f[1].empty() ? std::string("SET") : f[1], f[2], f[3]};
    } else if (sec == "BOUNDS") {
      const std::string ty = upper(f[0]);
      const bool noValue = ty == "FR" || ty == "MI" || ty == "PL" || ty == "BV";
      t = {f[0], f[1].empty() ? std::string("SET") : f[1], f[2]};
      if (!noValue) t.push_back(f[3]);
      return t;
    } else {
      return {f[0]};
    }
    if (!f[4].empty()) {
      t.push_back(f[4]);
      t.push_back(f[5]);
    }
    return t;
  }
  int row(const std::string& name) const {
    auto it = rowIndex.find(name);
    if (it == rowIndex.end()) fail("unknown row " + name);
    return it->second;
  }
  int col(const std::string& name, bool create) {
    auto it = colIndex.find(name);
    if (it != colIndex.end()) return it->second;
    if (!create) fail("unknown column " + name);
    const int j = static_cast<int>(cost.size());
    colIndex.emplace(name, j);
    cost.push_back(0);
    lo.push_back(0);
    up.push_back(kInf);
    hasLo.push_back(0);
    return j;
  }

  void rows_line(const std::vector<std::string>& t) {
    if (t.size() != 2 || t[1].empty()) fail("ROWS line needs a type and a name");
    const std::string ty = upper(t[0]);
    if (rowIndex.count(t[1])) fail("duplicate row " + t[1]);
    if (ty == "N") {
      const bool first = std::none_of(rowIndex.begin(), rowIndex.end(), [](const auto& kv) { return kv.second == -1; });
      rowIndex.emplace(t[1], first ? -1 : -2);
      return;
    }
    if (ty != "E" && ty != "L" && ty != "G") fail("row type " + t[0]);
    rowIndex.emplace(t[1], static_cast<int>(rowType.size()));
    rowType.push_back(ty[0]);
    rhs.push_back(0);
    range.push_back(0);
    hasRange.push_back(0);
  }

  void columns_line(const std::vector<std::string>& t) {
    if (t.size() >= 3 && (t[1] == "'MARKER'" || t[1] == "MARKER")) return;
    if (t.size() != 3 && t.size() != 5) fail("COLUMNS line needs a column and one or two (row, value) pairs");
    const int j = col(t[0], true);
    for (size_t k = 1; k + 1 < t.size(); k += 2) {
      const int r = row(t[k]);
      const double v = num(t[k + 1]);
      if (r == -1) cost[j] += v;
      else if (r >= 0) trip.emplace_back(j, r, v);
    }
  }

  // RHS and RANGES: an optional set name, then (row, value) pairs.
  template <typename F>
  void pairs_line(const std::vector<std::string>& t, F&& put) {
    const size_t k0 = t.size() % 2 == 1 ? 1 : 0;
    if (t.size() < k0 + 2) fail("expected (row, value) pairs");
    for (size_t k = k0; k + 1 < t.size(); k += 2) put(row(t[k]), num(t[k + 1]));
  }

  void bounds_line(const std::vector<std::string>& t) {
    const std::string ty = upper(t[0]);
    const bool noValue = ty == "FR" || ty == "MI" || ty == "PL" || ty == "BV";
    // [type, set, column(, value)] or, without a set name, [type, column(, value)].
    const size_t full = noValue ? 3 : 4;
    if (t.size() != full && t.size() != full - 1) fail("BOUNDS line has " + std::to_string(t.size()) + " fields");
    const int j = col(t[t.size() == full ? 2 : 1], false);
    const double v = noValue ? 0.0 : num(t.back());
    if (ty == "UP" || ty == "UI") {
      up[j] = v;
      // Classic rule: a negative upper bound without an explicit lower bound
      // makes the lower bound -inf.
      if (v < 0 && !hasLo[j]) lo[j] = -kInf;
    } else if (ty == "LO" || ty == "LI") {
      lo[j] = v;
      hasLo[j] = 1;
    } else if (ty == "FX") {
      lo[j] = up[j] = v;
      hasLo[j] = 1;
    } else if (ty == "FR") {
      lo[j] = -kInf;
      up[j] = kInf;
      hasLo[j] = 1;
    } else if (ty == "MI") {
      lo[j] = -kInf;
      hasLo[j] = 1;
    } else if (ty == "PL") {
      up[j] = kInf;
    } else if (ty == "BV") {
      lo[j] = 0;
      up[j] = 1;
      hasLo[j] = 1;
    } else {
      fail("bound type " + t[0] + " is not supported by the LP engine");
    }
  }

  Model finish() const {
    const int m = static_cast<int>(rowType.size()), n = static_cast<int>(cost.size());
    Model md;
    md.m = m;
    md.n = n;
    md.L.resize(m);
    md.U.resize(m);
    for (int i = 0; i < m; ++i) {
      const double b = rhs[i];
      double L = rowType[i] == 'L' ? -kInf : b, U = rowType[i] == 'G' ? kInf : b;
      if (hasRange[i]) {  // sign-dependent on E rows
        const double R = range[i];
        if (rowType[i] == 'E') {
          if (R >= 0) U = b + std::fabs(R);
          else L = b - std::fabs(R);
        } else if (rowType[i] == 'L') {
          L = b - std::fabs(R);
        } else {
          U = b + std::fabs(R);
        }
      }
      md.L[i] = L;
      md.U[i] = U;
    }
    // CSC in (column, row) order; duplicate entries summed, zeros dropped.
    auto t = trip;
    std::stable_sort(t.begin(), t.end(), [](const auto& a, const auto& b) {
      return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b) : std::get<1>(a) < std::get<1>(b);
    });
    md.A.rows = m;
    md.A.cols = n;
    md.A.start.assign(n + 1, 0);
    for (size_t e = 0; e < t.size();) {
      const int j = std::get<0>(t[e]), r = std::get<1>(t[e]);
      double v = 0;
      for (; e < t.size() && std::get<0>(t[e]) == j && std::get<1>(t[e]) == r; ++e) v += std::get<2>(t[e]);
      if (v == 0.0) continue;
      md.A.index.push_back(r);
      md.A.value.push_back(v);
      md.A.start[j + 1]++;
    }
    for (int j = 0; j < n; ++j) md.A.start[j + 1] += md.A.start[j];
    const double s = maximize ? -1.0 : 1.0;  // minimisation form
    md.c.resize(n);
    for (int j = 0; j < n; ++j) md.c[j] = s * cost[j];
    md.offset = s * offset;
    md.l = lo;
    md.u = up;
    return md;
  }
};

}  // namespace

Model read_mps(const std::string& path, int* sense) {
  gzFile f = gzopen(path.c_str(), "rb");
  if (!f) throw Error("cannot open MPS file " + path);
  MpsReader rd;
  rd.path = path;
  std::string sec, line;
  std::vector<char> buf(1 << 16);
  bool ended = false, fixed = false;
  try {
    while (!ended && gzgets(f, buf.data(), static_cast<int>(buf.size()))) {
      line.assign(buf.data());
      ++rd.lineNo;
      while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
      if (line.empty() || line[0] == '*') continue;
      auto t = tokens(line);
      if (t.empty()) continue;
      if (!std::isspace(static_cast<unsigned char>(line[0]))) {  // section header
        sec = upper(t[0]);
        if (sec == "ENDATA") ended = true;
        else if (sec == "OBJSENSE" || sec == "OBJSENCE") {
          if (t.size() >= 2) rd.maximize = upper(t[1]).rfind("MAX", 0) == 0;
        } else if (sec != "NAME" && sec != "ROWS" && sec != "COLUMNS" && sec != "RHS" && sec != "RANGES" &&
                   sec != "BOUNDS") {
          rd.fail("section " + t[0] + " is not supported by the LP engine");
        }
        continue;
      }
      if (sec == "ROWS" && t.size() > 2) fixed = true;  // names with blanks: fixed columns from here on
      if (fixed && sec != "OBJSENSE" && sec != "OBJSENCE") t = rd.from_fixed(sec, line);
      if (sec == "OBJSENSE" || sec == "OBJSENCE") rd.maximize = upper(t[0]).rfind("MAX", 0) == 0;
      else if (sec == "ROWS") rd.rows_line(t);
      else if (sec == "COLUMNS") rd.columns_line(t);
      else if (sec == "RHS")
        rd.pairs_line(t, [&](int r, double v) {
          if (r == -1) rd.offset = -v;  // RHS on the objective row: constant term -rhs
          else if (r >= 0) rd.rhs[r] = v;
        });
      else if (sec == "RANGES")
        rd.pairs_line(t, [&](int r, double v) {
          if (r >= 0) {
            rd.range[r] = v;
            rd.hasRange[r] = 1;
          }
        });
      else if (sec == "BOUNDS") rd.bounds_line(t);
      else rd.fail("data line outside a section");
    }
  } catch (...) {
    gzclose(f);
    throw;
  }
  gzclose(f);
  if (rd.rowIndex.empty()) throw Error("MPS " + path + ": no ROWS section");
  if (sense) *sense = rd.maximize ? -1 : 1;
  return rd.finish();
}

}  // namespace pdhgamd
