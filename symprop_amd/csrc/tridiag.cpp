// Host part of the eigen-step: the R largest eigenpairs of a symmetric
// tridiagonal matrix T (diagonal d[0..P), off-diagonal e[0..P-1)), the
// reduced form of the P x P Gram (sytrd.hip, or rocSOLVER's sytrd).  The same
// arithmetic as LAPACK's dstebz (Sturm-sequence bisection) and dstein
// (inverse iteration with partial pivoting, reorthogonalisation inside
// clusters of close eigenvalues), for the few wanted pairs only.  Every loop
// is bounded.
#include "tridiag.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace spamd {

namespace {

// Numbers of eigenvalues of T below each of the K shifts x[k]: the Sturm
// sequences (LDL^T pivots of T - x I) of all shifts advance together, so the
// K division chains overlap.
template <int K>
void sturm_counts(const double *d, const double *e2, int P, const double *x, double pivmin,
                  int *cnt) {
  double q[K];
  int c[K];
  for (int k = 0; k < K; ++k) {
    q[k] = d[0] - x[k];
    if (std::fabs(q[k]) < pivmin) q[k] = -pivmin;
    c[k] = q[k] < 0.0;
  }
  for (int i = 1; i < P; ++i) {
    const double di = d[i], ei = e2[i - 1];
    for (int k = 0; k < K; ++k) {
      double t = (di - x[k]) - ei / q[k];
      if (std::fabs(t) < pivmin) t = -pivmin;
      q[k] = t;
      c[k] += t < 0.0;
    }
  }
  for (int k = 0; k < K; ++k) cnt[k] = c[k];
}

double unit_random(unsigned long long &s) {
  s += 0x9E3779B97F4A7C15ull;
  unsigned long long z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

}  // namespace

void tridiag_top(const double *d, const double *e, int P, int R, double *w, double *Z) {
  if (P <= 0 || R <= 0 || R > P) throw std::invalid_argument("tridiag_top: size out of range");
  // Far outside [2^-400, 2^400] the pivot floor 10 eps ||T|| and its inverse
  // leave the normal range; solve T scaled by a power of two (exact), as
  // LAPACK's drivers scale such matrices, and scale the eigenvalues back.
  {
    double tn = 0.0;
    for (int i = 0; i < P; ++i) {
      const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < P ? std::fabs(e[i]) : 0.0);
      tn = std::max(tn, std::fabs(d[i]) + r);
    }
    if (tn > 0.0 && std::isfinite(tn) && (tn < 0x1p-400 || tn > 0x1p400)) {
      const int ex = -std::ilogb(tn);
      std::vector<double> ds(P), es(P > 1 ? P - 1 : 0);
      for (int i = 0; i < P; ++i) ds[i] = std::ldexp(d[i], ex);
      for (int i = 0; i + 1 < P; ++i) es[i] = std::ldexp(e[i], ex);
      tridiag_top(ds.data(), es.data(), P, R, w, Z);
      for (int j = 0; j < R; ++j) w[j] = std::ldexp(w[j], -ex);
      return;
    }
  }
  std::vector<double> e2(P, 0.0);
  double lo = d[0], hi = d[0], tn = 0.0, emax2 = 0.0;
  for (int i = 0; i < P; ++i) {
    if (i + 1 < P) {
      e2[i] = e[i] * e[i];
      emax2 = std::max(emax2, e2[i]);
    }
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < P ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
    tn = std::max(tn, std::fabs(d[i]) + r);
  }
  const double pad = 4.0 * DBL_EPSILON * std::max(tn, DBL_MIN) + 4.0 * DBL_MIN;
  lo -= pad;
  hi += pad;
  const double pivmin = DBL_MIN * std::max(1.0, emax2);

  if (tn == 0.0) {  // T = 0: every unit vector is an eigenvector
    for (int j = 0; j < R; ++j) {
      w[j] = 0.0;
      for (int i = 0; i < P; ++i) Z[(size_t)j * P + i] = i == P - 1 - j ? 1.0 : 0.0;
    }
    return;
  }

  // eigenvalues: w[j] = (j+1)-th largest, by bisection to full precision.
  // Up to 16 open intervals: one Sturm pass of 8 or 16 shifts spread over
  // the distinct open intervals (several eigenvalues share one until the
  // shifts separate them), and every shift's count narrows every open
  // interval (multisection).  More than 16: groups of 16 midpoints.
  std::vector<double> a(R, lo), b(R, hi);
  std::vector<int> open(R);
  auto narrow = [&](int j, double x, int cnt) {
    if (!(x > a[j] && x < b[j])) return;
    if (cnt > P - 1 - j) b[j] = x;  // the wanted one is below x
    else a[j] = x;
  };
  auto multisect = [&](auto kc, int m) {
    constexpr int K = decltype(kc)::value;
    int rep[K], nd = 0;  // one open index per distinct interval, in order
    for (int q = 0; q < m; ++q) {
      const int j = open[q];
      if (nd == 0 || a[j] != a[rep[nd - 1]] || b[j] != b[rep[nd - 1]]) rep[nd++] = j;
    }
    double x[K];
    int cnt[K], ns = 0;
    for (int i = 0; i < nd; ++i) {
      const int sl = K / nd + (i < K % nd ? 1 : 0);
      const double lo_i = a[rep[i]], w_i = b[rep[i]] - a[rep[i]];
      for (int q = 1; q <= sl; ++q) x[ns++] = lo_i + w_i * ((double)q / (sl + 1));
    }
    sturm_counts<K>(d, e2.data(), P, x, pivmin, cnt);
    for (int q = 0; q < m; ++q)
      for (int k = 0; k < K; ++k) narrow(open[q], x[k], cnt[k]);
  };
  for (int it = 0; it < 256; ++it) {
    int m = 0;
    for (int j = 0; j < R; ++j) {
      const double tol = 2.0 * DBL_EPSILON * std::max(std::fabs(a[j]), std::fabs(b[j])) + pivmin;
      const double x = 0.5 * (a[j] + b[j]);
      if (b[j] - a[j] > tol && x > a[j] && x < b[j]) open[m++] = j;
    }
    if (m == 0) break;
    if (m <= 8) {
      multisect(std::integral_constant<int, 8>(), m);
    } else if (m <= 16) {
      multisect(std::integral_constant<int, 16>(), m);
    } else {
      for (int g = 0; g < m; g += 16) {
        constexpr int K = 16;
        double x[K];
        int cnt[K];
        for (int k = 0; k < K; ++k) {
          const int j = open[std::min(g + k, m - 1)];
          x[k] = 0.5 * (a[j] + b[j]);
        }
        sturm_counts<K>(d, e2.data(), P, x, pivmin, cnt);
        for (int k = 0; k < K && g + k < m; ++k) narrow(open[g + k], x[k], cnt[k]);
      }
    }
  }
  for (int j = 0; j < R; ++j) w[j] = 0.5 * (a[j] + b[j]);

  // eigenvectors: clusters are runs closer than 1e-3 ||T|| (dstein); shifts
  // inside a cluster are kept 10 eps ||T|| apart; each iterate is
  // orthogonalised against the cluster's earlier vectors
  const double ortol = 1e-3 * tn;
  const double pertol = 10.0 * DBL_EPSILON * std::max(tn, DBL_MIN);
  std::vector<double> lw(P), rd(P), u1(P), u2(P);
  std::vector<char> swap(P);
  int j0 = 0;
  while (j0 < R) {
    int j1 = j0 + 1;
    while (j1 < R && w[j1 - 1] - w[j1] <= ortol) ++j1;
    double prev_shift = 0.0;
    unsigned long long seed = 0x5EED0000ull + (unsigned long long)j0;
    for (int j = j0; j < j1; ++j) {
      double x = w[j];
      if (j > j0 && prev_shift - x < pertol) x = prev_shift - pertol;
      prev_shift = x;
      // T - x I = P L U with partial pivoting (LAPACK dgttrf, streamed)
      double a = d[0] - x, c = P > 1 ? e[0] : 0.0;
      for (int i = 0; i + 1 < P; ++i) {
        const double b = e[i];
        const double dn = d[i + 1] - x;
        const double cn = i + 2 < P ? e[i + 1] : 0.0;
        if (std::fabs(a) >= std::fabs(b)) {
          const double piv = a == 0.0 ? pertol : a;
          const double l = b / piv;
          rd[i] = 1.0 / piv;
          u1[i] = c;
          u2[i] = 0.0;
          lw[i] = l;
          swap[i] = 0;
          a = dn - l * c;
          c = cn;
        } else {
          const double l = a / b;
          rd[i] = 1.0 / b;
          u1[i] = dn;
          u2[i] = cn;
          lw[i] = l;
          swap[i] = 1;
          a = c - l * dn;
          c = -l * cn;
        }
      }
      rd[P - 1] = 1.0 / (a == 0.0 ? pertol : a);
      u1[P - 1] = 0.0;
      u2[P - 1] = 0.0;
      double *y = Z + (size_t)j * P;
      for (int i = 0; i < P; ++i) y[i] = unit_random(seed);
      for (int it = 0; it < 4; ++it) {
        for (int i = 0; i + 1 < P; ++i) {
          if (swap[i]) {
            const double tmp = y[i];
            y[i] = y[i + 1];
            y[i + 1] = tmp - lw[i] * y[i];
          } else {
            y[i + 1] -= lw[i] * y[i];
          }
        }
        for (int i = P - 1; i >= 0; --i) {
          double s = y[i];
          if (i + 1 < P) s -= u1[i] * y[i + 1];
          if (i + 2 < P) s -= u2[i] * y[i + 2];
          y[i] = s * rd[i];
        }
        for (int q = j0; q < j; ++q) {
          const double *yq = Z + (size_t)q * P;
          double dot = 0.0;
          for (int i = 0; i < P; ++i) dot += yq[i] * y[i];
          for (int i = 0; i < P; ++i) y[i] -= dot * yq[i];
        }
        double nrm = 0.0;
        for (int i = 0; i < P; ++i) nrm += y[i] * y[i];
        const double inv = nrm > 0.0 ? 1.0 / std::sqrt(nrm) : 0.0;
        for (int i = 0; i < P; ++i) y[i] *= inv;
      }
    }
    j0 = j1;
  }
}

}  // namespace spamd
