// Eigen-decomposition of a dense symmetric matrix on the host (no HIP):
// Householder reduction to tridiagonal form, then implicit symmetric QR
// steps with Wilkinson shifts (Golub & Van Loan, Matrix Computations,
// algorithms 8.3.1 and 8.3.2), the orthogonal transformations accumulated.
// Used to compress the FMM's M2L operators (the Gram matrix of a level's
// stacked operators, a few hundred rows).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace abh {

// a: n x n symmetric, row-major.  On return w holds the eigenvalues in
// descending order and v (n x n, row-major) the matching orthonormal
// eigenvectors as columns: a = v diag(w) v^T.  Returns false if the QR
// iteration did not converge (30 n steps).
inline bool sym_eig(int n, const std::vector<double> &a, std::vector<double> &w,
                    std::vector<double> &v) {
  std::vector<double> t(a);  // working copy, reduced in place
  std::vector<double> q(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) q[size_t(i) * n + i] = 1.0;
  auto T = [&](int i, int j) -> double & { return t[size_t(i) * n + j]; };
  auto Q = [&](int i, int j) -> double & { return q[size_t(i) * n + j]; };
  std::vector<double> hv(n), p(n);
  // Householder: T <- H T H with H = I - 2 h h^T zeroing column k below k+1
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;  // length of the part below the diagonal
    double norm = 0.0;
    for (int i = 0; i < m; ++i) norm += T(k + 1 + i, k) * T(k + 1 + i, k);
    norm = std::sqrt(norm);
    if (norm == 0.0) continue;
    const double x0 = T(k + 1, k);
    const double alpha = x0 > 0.0 ? -norm : norm;
    for (int i = 0; i < m; ++i) hv[i] = T(k + 1 + i, k);
    hv[0] -= alpha;
    double hn = 0.0;
    for (int i = 0; i < m; ++i) hn += hv[i] * hv[i];
    hn = std::sqrt(hn);
    if (hn == 0.0) continue;
    for (int i = 0; i < m; ++i) hv[i] /= hn;
    // p = S h (S the trailing block), kk = h^T p, p <- p - kk h,
    // S <- S - 2 (h p^T + p h^T)
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int j = 0; j < m; ++j) s += T(k + 1 + i, k + 1 + j) * hv[j];
      p[i] = s;
    }
    double kk = 0.0;
    for (int i = 0; i < m; ++i) kk += hv[i] * p[i];
    for (int i = 0; i < m; ++i) p[i] -= kk * hv[i];
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        T(k + 1 + i, k + 1 + j) -= 2.0 * (hv[i] * p[j] + p[i] * hv[j]);
    T(k + 1, k) = alpha;
    T(k, k + 1) = alpha;
    for (int i = 1; i < m; ++i) {
      T(k + 1 + i, k) = 0.0;
      T(k, k + 1 + i) = 0.0;
    }
    // Q <- Q H on columns k+1..n-1
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int j = 0; j < m; ++j) s += Q(r, k + 1 + j) * hv[j];
      for (int j = 0; j < m; ++j) Q(r, k + 1 + j) -= 2.0 * s * hv[j];
    }
  }
  std::vector<double> d(n), e(std::max(n - 1, 0));
  for (int i = 0; i < n; ++i) d[i] = T(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = T(i + 1, i);
  // implicit QR on the tridiagonal (d, e); rotations R_k on rows k, k+1:
  // T <- R T R^T, Q <- Q R^T
  const double eps = 2.220446049250313e-16;
  int hi = n - 1;
  int64_t steps = 0;
  while (hi > 0) {
    for (int i = 0; i < hi; ++i)
      if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
    if (e[hi - 1] == 0.0) {
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && e[lo - 1] != 0.0) --lo;
    if (++steps > int64_t(30) * n) return false;
    // Wilkinson shift from the trailing 2 x 2 block
    const double dl = 0.5 * (d[hi - 1] - d[hi]);
    const double eb = e[hi - 1];
    const double den = dl + (dl >= 0.0 ? 1.0 : -1.0) * std::hypot(dl, eb);
    const double mu = d[hi] - eb * eb / den;
    double x = d[lo] - mu, z = e[lo], bulge = 0.0;
    for (int k = lo; k < hi; ++k) {
      const double r = std::hypot(x, z);
      const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : z / r;
      if (k > lo) e[k - 1] = r;
      const double a = d[k], b = e[k], cc = d[k + 1];
      d[k] = c * c * a + 2.0 * c * s * b + s * s * cc;
      d[k + 1] = s * s * a - 2.0 * c * s * b + c * c * cc;
      e[k] = c * s * (cc - a) + (c * c - s * s) * b;
      if (k + 1 < hi) {
        bulge = s * e[k + 1];
        e[k + 1] = c * e[k + 1];
      }
      for (int row = 0; row < n; ++row) {
        const double q0 = Q(row, k), q1 = Q(row, k + 1);
        Q(row, k) = c * q0 + s * q1;
        Q(row, k + 1) = -s * q0 + c * q1;
      }
      x = e[k];
      z = bulge;
    }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int i, int j) { return d[i] > d[j]; });
  w.resize(n);
  v.assign(size_t(n) * n, 0.0);
  for (int c = 0; c < n; ++c) {
    w[c] = d[ord[c]];
    for (int r = 0; r < n; ++r) v[size_t(r) * n + c] = Q(r, ord[c]);
  }
  return true;
}

}  // namespace abh
