#include "certify.h"

#include <algorithm>
#include <cmath>

namespace pdhgamd {

namespace {

// ---- outward-rounded arithmetic (src/cert/interval.h:14-91) -------------
// A round-to-nearest result is tested for exactness with an error-free
// transformation (TwoSum for +, FMA remainder for *); only an inexact result
// is moved one ulp in the safe direction, so no rounding-mode switch is needed.
inline double next_down(double v) { return (std::isinf(v) || std::isnan(v)) ? v : std::nextafter(v, -kInf); }
inline double next_up(double v) { return (std::isinf(v) || std::isnan(v)) ? v : std::nextafter(v, kInf); }

inline double sum_error(double a, double b, double s) {  // a + b = s + err exactly
  const double bv = s - a;
  return (a - (s - bv)) + (b - bv);
}
inline double add_lo(double a, double b) {
  const double s = a + b;
  if (std::isinf(s) || std::isnan(s)) return s;
  return sum_error(a, b, s) < 0 ? next_down(s) : s;
}
inline double add_hi(double a, double b) {
  const double s = a + b;
  if (std::isinf(s) || std::isnan(s)) return s;
  return sum_error(a, b, s) > 0 ? next_up(s) : s;
}
inline double mul_lo(double a, double b) {
  if (a == 0 || b == 0) return 0.0;
  const double p = a * b;
  if (std::isinf(p)) return p;
  return std::fma(a, b, -p) < 0 ? next_down(p) : p;
}
inline double mul_hi(double a, double b) {
  if (a == 0 || b == 0) return 0.0;
  const double p = a * b;
  if (std::isinf(p)) return p;
  return std::fma(a, b, -p) > 0 ? next_up(p) : p;
}

struct Ival {
  double lo, hi;
};

// Rigorous lower bound of min { z x : z in [z.lo, z.hi], x in [l, u] }.
double min_product(Ival z, double l, double u) {
  if (z.lo == 0 && z.hi == 0) return 0.0;
  if (std::isinf(l) && z.hi > 0) return -kInf;
  if (std::isinf(u) && z.lo < 0) return -kInf;
  double best = kInf;
  for (double zz : {z.lo, z.hi})
    for (double xx : {l, u}) {
      double p;
      if (zz == 0 || xx == 0) p = 0.0;
      else if (std::isinf(xx)) p = zz * xx;
      else p = mul_lo(zz, xx);
      best = std::min(best, p);
    }
  return best;
}

inline double as_inf(double v) { return finite_bound(v) ? v : (v > 0 ? kInf : -kInf); }

// Neumaier / Kahan-Babuska compensated sum (src/cert/interval.h:93-103).
struct CompSum {
  double s = 0, comp = 0;
  void add(double v) {
    const double t = s + v;
    if (std::fabs(s) >= std::fabs(v)) comp += (s - t) + v;
    else comp += (v - t) + s;
    s = t;
  }
  double value() const { return s + comp; }
};

}  // namespace

SafeBound safe_dual_bound(const Csc& A, const double* c, const double* l, const double* u, const double* L,
                          const double* U, const double* yIn) {
  SafeBound r;
  const int m = A.rows, n = A.cols;
  // Row multipliers projected onto their sign cone (free: the bound holds for any y).
  std::vector<double> y(yIn, yIn + m);
  for (int i = 0; i < m; ++i) {
    const bool hasL = finite_bound(L[i]), hasU = finite_bound(U[i]);
    double v = y[i];
    if (!hasL && !hasU) v = 0;
    else if (!hasL) v = std::min(v, 0.0);
    else if (!hasU) v = std::max(v, 0.0);
    r.rowProjection = std::max(r.rowProjection, std::fabs(v - y[i]));
    y[i] = v;
  }
  double lb = 0;
  bool minusInf = false;
  for (int i = 0; i < m; ++i) lb = add_lo(lb, min_product(Ival{y[i], y[i]}, as_inf(L[i]), as_inf(U[i])));
  const double maxPert = 1e-9;
  for (int j = 0; j < n; ++j) {
    // z_j = c_j - (A^T y)_j as an interval.
    Ival z{c[j], c[j]};
    for (int e = A.start[j]; e < A.start[j + 1]; ++e) {
      const double a = A.value[e], yv = y[A.index[e]];
      const double plo = mul_lo(a, yv), phi = mul_hi(a, yv);
      z = Ival{add_lo(z.lo, -phi), add_hi(z.hi, -plo)};
    }
    // Wrong-sign reduced cost on an infinite bound: clip and record the cost
    // perturbation under which the bound stays rigorous.
    double shift = 0;
    if (!finite_bound(l[j]) && z.hi > 0) {
      shift = std::max(shift, z.hi);
      z.hi = 0;
      z.lo = std::min(z.lo, 0.0);
    }
    if (!finite_bound(u[j]) && z.lo < 0) {
      shift = std::max(shift, -z.lo);
      z.lo = 0;
      z.hi = std::max(z.hi, 0.0);
    }
    if (shift > 0) {
      const double rel = shift / (1.0 + std::fabs(c[j]));
      r.costPerturbation = std::max(r.costPerturbation, rel);
      if (rel > maxPert) {
        r.infiniteTerms++;
        minusInf = true;
        continue;
      }
    }
    const double t = min_product(z, as_inf(l[j]), as_inf(u[j]));
    if (std::isinf(t)) {
      r.infiniteTerms++;
      minusInf = true;
      continue;
    }
    lb = add_lo(lb, t);
  }
  r.bound = minusInf ? -kInf : lb;
  return r;
}

void implied_bounds(const Csc& A, const double* L, const double* U, std::vector<double>& lo,
                    std::vector<double>& up) {
  const int m = A.rows, n = A.cols;
  const Csr rows = csr_of(A);
  std::vector<double> l2 = lo, u2 = up;
  std::vector<char> queued(m, 1);
  std::vector<int> queue(m), next;
  for (int i = 0; i < m; ++i) queue[i] = i;
  bool infeasible = false;
  for (int round = 0; round < 30 && !queue.empty() && !infeasible; ++round) {
    next.clear();
    for (int i : queue) {
      queued[i] = 0;
      // Finite parts and infinite-term counts of the min / max row activity.
      double minF = 0, maxF = 0;
      int minI = 0, maxI = 0;
      for (int e = rows.ptr[i]; e < rows.ptr[i + 1]; ++e) {
        const int j = rows.idx[e];
        const double a = rows.val[e];
        const double bmin = a > 0 ? l2[j] : u2[j], bmax = a > 0 ? u2[j] : l2[j];
        if (finite_bound(bmin)) minF += a * bmin; else ++minI;
        if (finite_bound(bmax)) maxF += a * bmax; else ++maxI;
      }
      const double Li = L[i], Ui = U[i];
      const double scale = 1.0 + std::max(finite_bound(Li) ? std::fabs(Li) : 0.0, finite_bound(Ui) ? std::fabs(Ui) : 0.0);
      if ((minI == 0 && finite_bound(Ui) && minF > Ui + 1e-6 * scale) ||
          (maxI == 0 && finite_bound(Li) && maxF < Li - 1e-6 * scale)) {
        infeasible = true;
        break;
      }
      for (int e = rows.ptr[i]; e < rows.ptr[i + 1]; ++e) {
        const int j = rows.idx[e];
        const double a = rows.val[e];
        const double bmin = a > 0 ? l2[j] : u2[j], bmax = a > 0 ? u2[j] : l2[j];
        // Activity of the row without column j.
        const bool minOk = finite_bound(bmin) ? minI == 0 : minI == 1;
        const bool maxOk = finite_bound(bmax) ? maxI == 0 : maxI == 1;
        const double resMin = finite_bound(bmin) ? minF - a * bmin : minF;
        const double resMax = finite_bound(bmax) ? maxF - a * bmax : maxF;
        double newLo = -kInf, newUp = kInf;
        if (finite_bound(Ui) && minOk) {
          const double lim = (Ui - resMin) / a;
          if (a > 0) newUp = lim;
          else newLo = lim;
        }
        if (finite_bound(Li) && maxOk) {
          const double lim = (Li - resMax) / a;
          if (a > 0) newLo = std::max(newLo, lim);
          else newUp = std::min(newUp, lim);
        }
        bool changed = false;
        // Only significant tightenings, each relaxed by a safety margin.
        const double range = (finite_bound(l2[j]) && finite_bound(u2[j])) ? u2[j] - l2[j] : kInf;
        if (finite_bound(newLo)) {
          const double cand = newLo - 1e-7 * (1 + std::fabs(newLo));
          if (cand > l2[j] + 1e-3 * std::min(range, 1.0 + std::fabs(cand))) {
            l2[j] = cand;
            changed = true;
          }
        }
        if (finite_bound(newUp)) {
          const double cand = newUp + 1e-7 * (1 + std::fabs(newUp));
          const double r2 = (finite_bound(l2[j]) && finite_bound(u2[j])) ? u2[j] - l2[j] : kInf;
          if (cand < u2[j] - 1e-3 * std::min(r2, 1.0 + std::fabs(cand))) {
            u2[j] = cand;
            changed = true;
          }
        }
        if (l2[j] > u2[j] + 1e-6 * (1 + std::fabs(l2[j]))) {
          infeasible = true;
          break;
        }
        if (l2[j] > u2[j]) l2[j] = u2[j];
        if (changed)
          for (int k = A.start[j]; k < A.start[j + 1]; ++k) {
            const int r = A.index[k];
            if (r != i && !queued[r]) {
              queued[r] = 1;
              next.push_back(r);
            }
          }
      }
      if (infeasible) break;
    }
    queue.swap(next);
  }
  // Tightened bounds (from whatever rounds completed) replace the originals,
  // relaxed once more by 1e-6 (1 + |v|) (src/parametric/parametric.cpp:69-72).
  for (int j = 0; j < n; ++j) {
    if (finite_bound(l2[j]) && (!finite_bound(lo[j]) || l2[j] > lo[j])) lo[j] = l2[j] - 1e-6 * (1 + std::fabs(l2[j]));
    if (finite_bound(u2[j]) && (!finite_bound(up[j]) || u2[j] < up[j])) up[j] = u2[j] + 1e-6 * (1 + std::fabs(u2[j]));
  }
}

double max_primal_violation(const Csc& A, const double* l, const double* u, const double* L, const double* U,
                            const double* x) {
  const int m = A.rows, n = A.cols;
  double viol = 0;
  for (int j = 0; j < n; ++j) {
    if (!std::isfinite(x[j])) return kInf;
    if (finite_bound(l[j]) && x[j] < l[j]) viol = std::max(viol, (l[j] - x[j]) / (1 + std::fabs(l[j])));
    if (finite_bound(u[j]) && x[j] > u[j]) viol = std::max(viol, (x[j] - u[j]) / (1 + std::fabs(u[j])));
  }
  std::vector<CompSum> act(m);
  for (int j = 0; j < n; ++j)
    for (int e = A.start[j]; e < A.start[j + 1]; ++e) act[A.index[e]].add(A.value[e] * x[j]);
  for (int i = 0; i < m; ++i) {
    const double r = act[i].value();
    if (!std::isfinite(r)) return kInf;
    if (finite_bound(L[i]) && r < L[i]) viol = std::max(viol, (L[i] - r) / (1 + std::fabs(L[i])));
    if (finite_bound(U[i]) && r > U[i]) viol = std::max(viol, (r - U[i]) / (1 + std::fabs(U[i])));
  }
  return viol;
}

FamilyCert certify_case(const Csc& A, const double* c, const double* l, const double* u, const double* L,
                        const double* U, double offset, const double* x, const double* y, double objective) {
  FamilyCert fc;
  const int n = A.cols;
  std::vector<double> lo(l, l + n), up(u, u + n);
  implied_bounds(A, L, U, lo, up);
  SafeBound sb = safe_dual_bound(A, c, lo.data(), up.data(), L, U, y);
  fc.bound = sb.bound + offset;
  if (x) fc.primalViolation = max_primal_violation(A, l, u, L, U, x);
  if (std::isfinite(sb.bound)) fc.relGap = std::fabs(objective - fc.bound) / (1 + std::fabs(objective));
  fc.accepted = std::isfinite(sb.bound) && fc.relGap <= 1e-4;
  return fc;
}

}  // namespace pdhgamd
