// CPU oracle: radial kernel profiles shared by the operator restatements.
// TEST INFRASTRUCTURE ONLY.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

// Mirrors abh_kernel in include/aboria_hip.h.
struct Kernel {
  int32_t type;
  int32_t reserved;
  double param;
  double w0, w_row, w_col;
};

enum { K_CONSTANT = 0, K_GAUSSIAN = 1, K_WENDLAND_C2 = 2, K_MULTIQUADRIC = 3, K_SMOOTH_INV_R = 4 };

// phi as a function of the squared distance s = sum_d dx_d^2.
inline double phi(const Kernel &k, double s) {
  switch (k.type) {
    case K_CONSTANT:
      return 1.0;
    case K_GAUSSIAN:
      return std::exp(-(s / (k.param * k.param)));
    case K_WENDLAND_C2: {
      const double q = std::sqrt(s) / k.param;
      if (!(q < 1.0)) return 0.0;
      const double t = 1.0 - q;
      const double t2 = t * t;
      return (t2 * t2) * (4.0 * q + 1.0);
    }
    case K_MULTIQUADRIC:
      return std::sqrt(s + k.param * k.param);
    case K_SMOOTH_INV_R:
      return 1.0 / (std::sqrt(s) + k.param);
    default:
      return 0.0;
  }
}

// f(dx, a_i, b_j) = phi * (w0 + w_row*a_i + w_col*b_j) when scalars exist.
inline double weight(const Kernel &k, const double *row_s, const double *col_s, size_t i, size_t j) {
  if (!row_s && !col_s) return 1.0;
  double w = k.w0;
  if (row_s) w = w + k.w_row * row_s[i];
  if (col_s) w = w + k.w_col * col_s[j];
  return w;
}

}  // namespace orc
