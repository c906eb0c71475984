// Eigendecomposition of a small real symmetric matrix on the host: the projected problem of
// the restarted Lanczos (an arrowhead block plus a tridiagonal, order m <= 256).
//
// Method: Householder reduction to tridiagonal form with the transformations accumulated,
// then the QL iteration with implicit Wilkinson-type shifts on the tridiagonal, applying
// every plane rotation to the accumulated basis (Bowdler, Martin, Reinsch and Wilkinson,
// "The QR and QL algorithms for symmetric matrices", Numer. Math. 11 (1968), with the
// Householder tridiagonalisation of Martin, Reinsch and Wilkinson, Numer. Math. 11 (1968)).
// Cost ~ (4/3 + 3) m^3 flops, against ~ 6 m^3 per sweep for cyclic Jacobi.
#pragma once

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace fame6d {

// A: m x m column-major (symmetric; the full matrix is read).  On return w holds the
// eigenvalues in ascending order and V (m x m column-major) the orthonormal eigenvectors in
// the same order.  Throws std::runtime_error if the QL iteration does not converge.
inline void sym_eig(int m, const std::vector<double>& A, std::vector<double>& w, std::vector<double>& V) {
  // z(r, c) = working matrix (row-major here for the reduction), d/e = tridiagonal.
  std::vector<double> z((size_t)m * m), d(m, 0.0), e(m, 0.0);
  auto Z = [&](int r, int c) -> double& { return z[(size_t)r * m + c]; };
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) Z(r, c) = 0.5 * (A[(size_t)c * m + r] + A[(size_t)r * m + c]);

  // Householder: for i = m-1 .. 1 annihilate row i left of the subdiagonal.  The vector u of
  // each reflector is kept in row i (columns 0..i-1) and u/h in column i (rows 0..i-1).
  for (int i = m - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0;
    if (l > 0) {
      double scale = 0.0;
      for (int k = 0; k <= l; ++k) scale += std::fabs(Z(i, k));
      if (scale == 0.0) {
        e[i] = Z(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          Z(i, k) /= scale;
          h += Z(i, k) * Z(i, k);
        }
        double f = Z(i, l);
        const double g0 = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g0;
        h -= f * g0;
        Z(i, l) = f - g0;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          Z(j, i) = Z(i, j) / h;
          double g = 0.0;
          for (int k = 0; k <= j; ++k) g += Z(j, k) * Z(i, k);
          for (int k = j + 1; k <= l; ++k) g += Z(k, j) * Z(i, k);
          e[j] = g / h;
          f += e[j] * Z(i, j);
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          const double fj = Z(i, j);
          const double gj = e[j] - hh * fj;
          e[j] = gj;
          for (int k = 0; k <= j; ++k) Z(j, k) -= fj * e[k] + gj * Z(i, k);
        }
      }
    } else {
      e[i] = Z(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  // accumulate the reflectors into Z (Z becomes the orthogonal matrix of the reduction)
  for (int i = 0; i < m; ++i) {
    if (d[i] != 0.0) {
      for (int j = 0; j < i; ++j) {
        double g = 0.0;
        for (int k = 0; k < i; ++k) g += Z(i, k) * Z(k, j);
        for (int k = 0; k < i; ++k) Z(k, j) -= g * Z(k, i);
      }
    }
    d[i] = Z(i, i);
    Z(i, i) = 1.0;
    for (int j = 0; j < i; ++j) Z(j, i) = Z(i, j) = 0.0;
  }

  // QL with implicit shifts on (d, e); e[i] couples i-1 and i, moved to e[i-1].
  for (int i = 1; i < m; ++i) e[i - 1] = e[i];
  if (m > 0) e[m - 1] = 0.0;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < m; ++l) {
    int iter = 0;
    int mm;
    do {
      for (mm = l; mm < m - 1; ++mm) {
        const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
        if (std::fabs(e[mm]) <= eps * dd) break;
      }
      if (mm != l) {
        if (iter++ == 60) throw std::runtime_error("symmetric eigensolver: QL iteration did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[mm] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        bool underflow = false;
        for (i = mm - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {  // recover from underflow
            d[i + 1] -= p;
            e[mm] = 0.0;
            underflow = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < m; ++k) {  // rotate columns i, i+1 of the basis
            f = Z(k, i + 1);
            Z(k, i + 1) = s * Z(k, i) + c * f;
            Z(k, i) = c * Z(k, i) - s * f;
          }
        }
        if (underflow && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[mm] = 0.0;
      }
    } while (mm != l);
  }

  // ascending order; Z holds the eigenvectors in its columns
  std::vector<int> ord(m);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  w.resize(m);
  V.assign((size_t)m * m, 0.0);
  for (int j = 0; j < m; ++j) {
    w[j] = d[ord[j]];
    for (int r = 0; r < m; ++r) V[(size_t)j * m + r] = Z(r, ord[j]);
  }
}

}  // namespace fame6d
