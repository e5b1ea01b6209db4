// Synthetic LP generators of the benchmark workloads (all SplitMix64 seeded).
//
//  * random LP with a known optimal pair (x*, y*): m x n, equality rows b = A x*,
//    c = A'y* + s*, x >= 0.  Uniform rows per column, or log-normal row lengths.
//  * refinery-style multi-period blending LP: block staircase over periods,
//    crude purchase / distillation / inventories, stream blending into products
//    under quality specs, every column bounded, rows and columns scaled by
//    10^U(-d, d) so that preconditioning has to recover the scale.
//  * the price family of a refinery model (purchase price of four crudes).
//  * the reference's calibration LP (src/parametric/parametric.cpp:543-563).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "core.h"

namespace pdhgamd {

namespace {

// Builds a CSC matrix from triplets: entries of a column sorted by row,
// duplicates summed, explicit zeros dropped (the convention of the
// reference's SparseMatrix::fromTriplets, src/core/sparse.cpp:11-57).
Csc from_triplets(int rows, int cols, const std::vector<int>& ri, const std::vector<int>& ci,
                  const std::vector<double>& v) {
  Csc a;
  a.rows = rows;
  a.cols = cols;
  std::vector<int> cnt(cols + 1, 0);
  for (int c : ci) cnt[c + 1]++;
  for (int j = 0; j < cols; ++j) cnt[j + 1] += cnt[j];
  std::vector<int> order(v.size());
  {
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (size_t e = 0; e < v.size(); ++e) order[pos[ci[e]]++] = static_cast<int>(e);
  }
  a.start.assign(cols + 1, 0);
  a.index.reserve(v.size());
  a.value.reserve(v.size());
  for (int j = 0; j < cols; ++j) {
    auto b = order.begin() + cnt[j], e = order.begin() + cnt[j + 1];
    std::stable_sort(b, e, [&](int p, int q) { return ri[p] < ri[q]; });
    int last = -1;
    for (auto it = b; it != e; ++it) {
      if (ri[*it] == last) {
        a.value.back() += v[*it];
      } else {
        a.index.push_back(ri[*it]);
        a.value.push_back(v[*it]);
        last = ri[*it];
      }
    }
    int w = a.start[j];
    for (int k = a.start[j]; k < static_cast<int>(a.index.size()); ++k)
      if (a.value[k] != 0.0) {
        a.index[w] = a.index[k];
        a.value[w] = a.value[k];
        ++w;
      }
    a.index.resize(w);
    a.value.resize(w);
    a.start[j + 1] = w;
  }
  return a;
}

}  // namespace

Csr csr_of(const Csc& a) {
  Csr t;
  t.rows = a.rows;
  t.cols = a.cols;
  const int nnz = a.nnz();
  std::vector<int> cnt(a.rows + 1, 0);
  for (int e = 0; e < nnz; ++e) cnt[a.index[e] + 1]++;
  for (int i = 0; i < a.rows; ++i) cnt[i + 1] += cnt[i];
  t.ptr = cnt;
  t.idx.resize(nnz);
  t.val.resize(nnz);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (int j = 0; j < a.cols; ++j)
    for (int e = a.start[j]; e < a.start[j + 1]; ++e) {
      int p = pos[a.index[e]]++;
      t.idx[p] = j;
      t.val[p] = a.value[e];
    }
  return t;
}

Csr csr_of_transpose(const Csc& a) {
  Csr t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.ptr = a.start;
  t.idx = a.index;
  t.val = a.value;
  return t;
}

Model gen_random_lp(uint64_t seed, int m, int n, double perCol, double rowSigma) {
  PDHG_REQUIRE(m > 0 && n > 0 && perCol > 0, "random LP: bad size");
  Rng rng(seed);
  std::vector<int> ri, ci;
  std::vector<double> v;
  const size_t expect = static_cast<size_t>(perCol * n) + 16;
  ri.reserve(expect);
  ci.reserve(expect);
  v.reserve(expect);
  if (rowSigma <= 0) {
    const int per = static_cast<int>(std::lround(perCol));
    for (int j = 0; j < n; ++j)
      for (int e = 0; e < per; ++e) {
        ri.push_back(rng.below(m));
        ci.push_back(j);
        v.push_back(rng.normal());
      }
  } else {
    // Log-normal row lengths with mean perCol * n / m: exp(sigma z - sigma^2/2) has mean 1.
    const double mean = perCol * n / m;
    for (int i = 0; i < m; ++i) {
      double len = mean * std::exp(rowSigma * rng.normal() - 0.5 * rowSigma * rowSigma);
      int L = static_cast<int>(std::lround(len));
      L = std::max(0, std::min(n, L));
      for (int e = 0; e < L; ++e) {
        ri.push_back(i);
        ci.push_back(rng.below(n));
        v.push_back(rng.normal());
      }
    }
  }
  Model md;
  md.m = m;
  md.n = n;
  md.A = from_triplets(m, n, ri, ci, v);
  // Known optimum: a random half of x* is zero, the rest U(0,1); s* is U(0,1)
  // on the zero set and 0 elsewhere; y* ~ N(0,1).
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int j = n - 1; j > 0; --j) std::swap(perm[j], perm[rng.below(j + 1)]);
  md.xstar.assign(n, 0.0);
  std::vector<double> s(n, 0.0);
  for (int r = 0; r < n; ++r) {
    int j = perm[r];
    if (r < n / 2) s[j] = rng.uniform();
    else md.xstar[j] = rng.uniform();
  }
  md.ystar.resize(m);
  for (int i = 0; i < m; ++i) md.ystar[i] = rng.normal();
  std::vector<double> b(m, 0.0);
  md.c.assign(n, 0.0);
  for (int j = 0; j < n; ++j) {
    double acc = 0;
    for (int e = md.A.start[j]; e < md.A.start[j + 1]; ++e) {
      b[md.A.index[e]] += md.A.value[e] * md.xstar[j];
      acc += md.A.value[e] * md.ystar[md.A.index[e]];
    }
    md.c[j] = acc + s[j];
  }
  md.L = b;
  md.U = b;
  md.l.assign(n, 0.0);
  md.u.assign(n, kInf);
  return md;
}

Model gen_refinery(uint64_t seed, int T, int S, int K, int P, int Q, double decades) {
  PDHG_REQUIRE(T > 0 && S > 0 && K >= 4 && P > 0 && Q > 0, "refinery: bad size");
  Rng rng(seed);
  // Plant data.
  std::vector<double> yield(static_cast<size_t>(K) * S, 0.0);
  for (int k = 0; k < K; ++k) {
    int cnt = 0;
    for (int s = 0; s < S; ++s)
      if (rng.uniform() < 0.25) {
        yield[static_cast<size_t>(k) * S + s] = 0.01 + 0.59 * rng.uniform();
        ++cnt;
      }
    if (cnt == 0) yield[static_cast<size_t>(k) * S + rng.below(S)] = 0.01 + 0.59 * rng.uniform();
  }
  for (int s = 0; s < S; ++s) {  // every stream comes from at least one crude
    bool any = false;
    for (int k = 0; k < K; ++k) any = any || yield[static_cast<size_t>(k) * S + s] > 0;
    if (!any) yield[static_cast<size_t>(rng.below(K)) * S + s] = 0.01 + 0.59 * rng.uniform();
  }
  std::vector<double> qual(static_cast<size_t>(S) * Q), spec(static_cast<size_t>(P) * Q);
  std::vector<char> isMax(static_cast<size_t>(P) * Q);
  for (auto& q : qual) q = 100.0 * rng.uniform();
  for (size_t t = 0; t < spec.size(); ++t) {
    isMax[t] = rng.uniform() < 0.5;
    double slack = 5.0 + 25.0 * rng.uniform();
    spec[t] = isMax[t] ? 50.0 + slack : 50.0 - slack;
  }
  std::vector<double> crudePrice(K), prodPrice(P), avail(K), demand(P);
  for (int k = 0; k < K; ++k) {
    crudePrice[k] = 20.0 + 100.0 * rng.uniform();
    avail[k] = 50.0 + 100.0 * rng.uniform();
  }
  for (int p = 0; p < P; ++p) {
    prodPrice[p] = 20.0 + 100.0 * rng.uniform();
    demand[p] = 20.0 + 60.0 * rng.uniform();
  }
  double cduCap = 0;
  for (int k = 0; k < K; ++k) cduCap += 0.6 * avail[k];

  // Column layout per period: buy[K] dist[K] invC[K] invS[S] x[S*P] sell[P].
  const int perT = 3 * K + S + S * P + P;
  const int n = T * perT;
  auto buy = [&](int t, int k) { return t * perT + k; };
  auto dist = [&](int t, int k) { return t * perT + K + k; };
  auto invC = [&](int t, int k) { return t * perT + 2 * K + k; };
  auto invS = [&](int t, int s) { return t * perT + 3 * K + s; };
  auto flow = [&](int t, int s, int p) { return t * perT + 3 * K + S + s * P + p; };
  auto sell = [&](int t, int p) { return t * perT + 3 * K + S + S * P + p; };
  Model md;
  md.n = n;
  md.c.assign(n, 0.0);
  md.l.assign(n, 0.0);
  md.u.assign(n, 0.0);
  md.crudeCols.assign(4, {});
  for (int t = 0; t < T; ++t) {
    for (int k = 0; k < K; ++k) {
      md.c[buy(t, k)] = crudePrice[k] * (0.85 + 0.3 * rng.uniform());
      md.u[buy(t, k)] = avail[k];
      md.c[dist(t, k)] = 1.0;
      md.u[dist(t, k)] = 2.0 * avail[k];
      md.c[invC(t, k)] = 0.2;
      md.u[invC(t, k)] = 100.0;
      if (k < 4) md.crudeCols[k].push_back(buy(t, k));
    }
    for (int s = 0; s < S; ++s) {
      md.c[invS(t, s)] = 0.1;
      md.u[invS(t, s)] = 50.0;
      for (int p = 0; p < P; ++p) {
        md.c[flow(t, s, p)] = 0.01;
        md.u[flow(t, s, p)] = demand[p];
      }
    }
    for (int p = 0; p < P; ++p) {
      md.c[sell(t, p)] = -prodPrice[p] * (0.85 + 0.3 * rng.uniform());
      md.u[sell(t, p)] = demand[p];
    }
  }
  std::vector<int> ri, ci;
  std::vector<double> v;
  int row = 0;
  auto put = [&](int r, int j, double a) {
    ri.push_back(r);
    ci.push_back(j);
    v.push_back(a);
  };
  for (int t = 0; t < T; ++t) {
    for (int k = 0; k < K; ++k, ++row) {  // crude balance
      if (t > 0) put(row, invC(t - 1, k), 1.0);
      put(row, buy(t, k), 1.0);
      put(row, dist(t, k), -1.0);
      put(row, invC(t, k), -1.0);
      md.L.push_back(0.0);
      md.U.push_back(0.0);
    }
    for (int k = 0; k < K; ++k) put(row, dist(t, k), 1.0);  // distillation capacity
    md.L.push_back(-kInf);
    md.U.push_back(cduCap);
    ++row;
    for (int s = 0; s < S; ++s, ++row) {  // stream balance
      if (t > 0) put(row, invS(t - 1, s), 1.0);
      for (int k = 0; k < K; ++k) {
        double y = yield[static_cast<size_t>(k) * S + s];
        if (y > 0) put(row, dist(t, k), y);
      }
      for (int p = 0; p < P; ++p) put(row, flow(t, s, p), -1.0);
      put(row, invS(t, s), -1.0);
      md.L.push_back(0.0);
      md.U.push_back(0.0);
    }
    for (int p = 0; p < P; ++p, ++row) {  // product volume
      for (int s = 0; s < S; ++s) put(row, flow(t, s, p), 1.0);
      put(row, sell(t, p), -1.0);
      md.L.push_back(0.0);
      md.U.push_back(0.0);
    }
    for (int p = 0; p < P; ++p)  // blend quality specs
      for (int q = 0; q < Q; ++q, ++row) {
        const size_t pq = static_cast<size_t>(p) * Q + q;
        for (int s = 0; s < S; ++s) put(row, flow(t, s, p), qual[static_cast<size_t>(s) * Q + q] - spec[pq]);
        md.L.push_back(isMax[pq] ? -kInf : 0.0);
        md.U.push_back(isMax[pq] ? 0.0 : kInf);
      }
  }
  md.m = row;
  md.A = from_triplets(md.m, n, ri, ci, v);
  // Deliberate bad scaling: A' = Dr A Dc, x = Dc x', c' = Dc c, l' = l / Dc,
  // row bounds scaled by Dr.  The optimum value is unchanged.
  if (decades > 0) {
    std::vector<double> dr(md.m), dc(n);
    for (auto& s : dr) s = std::pow(10.0, decades * (2.0 * rng.uniform() - 1.0));
    for (auto& s : dc) s = std::pow(10.0, decades * (2.0 * rng.uniform() - 1.0));
    for (int j = 0; j < n; ++j) {
      for (int e = md.A.start[j]; e < md.A.start[j + 1]; ++e) md.A.value[e] *= dr[md.A.index[e]] * dc[j];
      md.c[j] *= dc[j];
      md.l[j] /= dc[j];
      md.u[j] /= dc[j];
    }
    for (int i = 0; i < md.m; ++i) {
      md.L[i] *= dr[i];
      md.U[i] *= dr[i];
    }
  }
  md.crudePrice.assign(crudePrice.begin(), crudePrice.begin() + 4);
  return md;
}

// cost = c0 + sum_q theta_q * d_q, d_q = c0 on crude q's purchase columns.
void family_costs(const Model& base, int k, const double* thetas, double* costs) {
  PDHG_REQUIRE(base.crudeCols.size() == 4, "family: base model has no crude purchase columns");
  for (int b = 0; b < k; ++b) {
    double* cb = costs + static_cast<size_t>(b) * base.n;
    std::copy(base.c.begin(), base.c.end(), cb);
    for (int q = 0; q < 4; ++q)
      for (int j : base.crudeCols[q]) cb[j] = base.c[j] + thetas[b * 4 + q] * base.c[j];
  }
}

void gen_family(const Model& base, uint64_t seed, int k, double range, double* thetas, double* costs) {
  Rng rng(seed);
  for (int b = 0; b < k; ++b)
    for (int q = 0; q < 4; ++q) thetas[b * 4 + q] = range * (2.0 * rng.uniform() - 1.0);
  family_costs(base, k, thetas, costs);
}

// src/parametric/parametric.cpp:543-563 (Rng(7) there; the seed is a parameter here).
Model gen_calibration(uint64_t seed, int s) {
  Rng rng(seed);
  const int m = s, n = 2 * s;
  std::vector<int> ri, ci;
  std::vector<double> v;
  for (int j = 0; j < n; ++j)
    for (int e = 0; e < 4; ++e) {
      ri.push_back(rng.below(m));
      ci.push_back(j);
      v.push_back(0.5 + rng.uniform());
    }
  Model md;
  md.m = m;
  md.n = n;
  md.A = from_triplets(m, n, ri, ci, v);
  md.c.resize(n);
  md.l.assign(n, 0.0);
  md.u.assign(n, 10.0);
  for (int j = 0; j < n; ++j) md.c[j] = rng.uniform() - 0.3;
  md.L.assign(m, 1.0);
  md.U.assign(m, kInf);
  return md;
}

}  // namespace pdhgamd
