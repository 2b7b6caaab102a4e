// Kernel profiles of the operators (sparse, dense, FMM): a radial phi of the
// squared distance times an optional affine per-particle weight, and the
// compensated FP32 accumulator used by the mixed-precision paths.
#pragma once

#include "common.hpp"

namespace abh {

struct KernelDesc {
  int32_t type;
  double param, param2;  // param2 = param^2
  double w0, wr, wc;
  const double *row_s;
  const double *col_s;
};

__device__ __forceinline__ double phi64(const KernelDesc &k, double s) {
  switch (k.type) {
    case ABH_KERNEL_CONSTANT: return 1.0;
    case ABH_KERNEL_GAUSSIAN: return exp(-(s / k.param2));
    case ABH_KERNEL_WENDLAND_C2: {
      const double q = sqrt(s) / k.param;
      if (!(q < 1.0)) return 0.0;
      const double t = 1.0 - q;
      const double t2 = t * t;
      return (t2 * t2) * (4.0 * q + 1.0);
    }
    case ABH_KERNEL_MULTIQUADRIC: return sqrt(s + k.param2);
    case ABH_KERNEL_SMOOTH_INV_R: return 1.0 / (sqrt(s) + k.param);
  }
  return 0.0;
}

__device__ __forceinline__ float phi32(const KernelDesc &k, float s) {
  const float p = float(k.param);
  switch (k.type) {
    case ABH_KERNEL_CONSTANT: return 1.0f;
    case ABH_KERNEL_GAUSSIAN: return expf(-(s / float(k.param2)));
    case ABH_KERNEL_WENDLAND_C2: {
      const float q = sqrtf(s) / p;
      if (!(q < 1.0f)) return 0.0f;
      const float t = 1.0f - q;
      const float t2 = t * t;
      return (t2 * t2) * (4.0f * q + 1.0f);
    }
    case ABH_KERNEL_MULTIQUADRIC: return sqrtf(s + float(k.param2));
    case ABH_KERNEL_SMOOTH_INV_R: return 1.0f / (sqrtf(s) + p);
  }
  return 0.0f;
}

// Neumaier's variant of Kahan summation: sum + c carries the FP32 sum.
__device__ __forceinline__ void kb_add(float &sum, float &c, float v) {
  const float t = sum + v;
  if (fabsf(sum) >= fabsf(v)) c += (sum - t) + v;
  else c += (v - t) + sum;
  sum = t;
}


}  // namespace abh
