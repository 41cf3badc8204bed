// Runs the library's host tridiagonal eigensolver (symprop_amd/csrc/
// tridiag.cpp) on one matrix: reads "P R" then d (P values) and e (P-1
// values) from stdin, prints w (R values) then Z (R*P values).
#include <cstdio>
#include <vector>

#include "tridiag.hpp"

int main() {
  int P, R;
  if (std::scanf("%d %d", &P, &R) != 2) return 2;
  std::vector<double> d(P), e(P > 1 ? P - 1 : 1), w(R), Z((size_t)R * P);
  for (auto &x : d)
    if (std::scanf("%lf", &x) != 1) return 2;
  for (int i = 0; i + 1 < P; ++i)
    if (std::scanf("%lf", &e[i]) != 1) return 2;
  spamd::tridiag_top(d.data(), e.data(), P, R, w.data(), Z.data());
  for (double x : w) std::printf("%.17g\n", x);
  for (double x : Z) std::printf("%.17g\n", x);
  return 0;
}
