// Test driver: reads m and an m x m column-major matrix (text) from stdin, writes the
// eigenvalues and the column-major eigenvector matrix of fame6d::sym_eig to stdout.
#include <cstdio>
#include <vector>

#include "dense_eig.hpp"

int main() {
  int m = 0;
  if (std::scanf("%d", &m) != 1 || m < 0) return 2;
  std::vector<double> A((size_t)m * m);
  for (auto& a : A)
    if (std::scanf("%lf", &a) != 1) return 2;
  std::vector<double> w, V;
  fame6d::sym_eig(m, A, w, V);
  for (double x : w) std::printf("%.17g\n", x);
  for (double x : V) std::printf("%.17g\n", x);
  return 0;
}
