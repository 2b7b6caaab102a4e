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
  double inv, inv2;      // 1 / param, 1 / param^2 (set_inverses)
};

inline void set_inverses(KernelDesc &k) {
  k.inv = k.param != 0.0 ? 1.0 / k.param : 0.0;
  k.inv2 = k.param2 != 0.0 ? 1.0 / k.param2 : 0.0;
}

// sqrt(s) for s >= 1e-280: the hardware reciprocal square root refined by
// one Goldschmidt step and one Newton correction (about one ulp; the libm
// square root is correctly rounded and costs twice the instructions with its
// range scaling)
__device__ __forceinline__ double sqrt_pos(double s) {
  const double y = __builtin_amdgcn_rsq(s);
  double g = s * y, h = 0.5 * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, s);
  return fma(d, h, g);
}

// sqrt(s) for s >= 0 (s below 1e-280 gives 0)
__device__ __forceinline__ double sqrt_fast(double s) {
  const double g = sqrt_pos(s);
  return s >= 1e-280 ? g : 0.0;
}

// 1 / x for a normal x: hardware reciprocal and two Newton steps (~1 ulp)
__device__ __forceinline__ double rcp_fast(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}

// The radial profiles in FP64 (operators, local matrices, FMM transfer
// operators): within a few ulp of the libm forms, which is inside every
// tolerance they are held to (the radius tests never use them).
__device__ __forceinline__ double phi64(const KernelDesc &k, double s) {
  switch (k.type) {
    case ABH_KERNEL_CONSTANT: return 1.0;
    case ABH_KERNEL_GAUSSIAN: return exp(-(s * k.inv2));
    case ABH_KERNEL_WENDLAND_C2: {
      const double q = sqrt_fast(s) * k.inv;
      if (!(q < 1.0)) return 0.0;
      const double t = 1.0 - q;
      const double t2 = t * t;
      return (t2 * t2) * fma(4.0, q, 1.0);
    }
    case ABH_KERNEL_MULTIQUADRIC: return sqrt_fast(s + k.param2);
    case ABH_KERNEL_SMOOTH_INV_R: return rcp_fast(sqrt_fast(s) + k.param);
  }
  return 0.0;
}

// FP32 profiles with the hardware square root and reciprocal (1 ulp)
__device__ __forceinline__ float phi32(const KernelDesc &k, float s) {
  const float p = float(k.param);
  switch (k.type) {
    case ABH_KERNEL_CONSTANT: return 1.0f;
    case ABH_KERNEL_GAUSSIAN: return expf(-(s * float(k.inv2)));
    case ABH_KERNEL_WENDLAND_C2: {
      const float q = __builtin_amdgcn_sqrtf(s) * float(k.inv);
      if (!(q < 1.0f)) return 0.0f;
      const float t = 1.0f - q;
      const float t2 = t * t;
      return (t2 * t2) * fmaf(4.0f, q, 1.0f);
    }
    case ABH_KERNEL_MULTIQUADRIC: return __builtin_amdgcn_sqrtf(s + float(k.param2));
    case ABH_KERNEL_SMOOTH_INV_R: return __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(s) + p);
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
