// Learned approximate-inverse preconditioner.
//
//   k_knn        k nearest neighbours of every particle (one wavefront per
//                particle, the k-1 best kept sorted across the lanes)
//   k_lsq        least-squares approximate-inverse row of every particle on
//                its neighbour pattern: normal equations accumulated in FP64
//                over every particle within reach, solved in LDS
//   k_patch      the local solve of every particle's patch (its k nearest
//                neighbours): the row of (K + shift I)^{-1} of the patch
//                matrix that belongs to the particle, Cholesky in LDS
//   k_resid      the patch-residual loss || M y - e_1 ||^2 of the network's
//                training and its gradient, one wavefront per row
//   k_gemm       FP32 tiled GEMM with fused bias / SiLU, the forward and
//                backward passes of the network's training
//   k_mlp_bf16   whole-network inference on MFMA (BF16 operands, FP32
//                accumulation): one wavefront per 32 rows, activations in
//                LDS, weights streamed from L2
//   k_ainv_apply x_i = sum_k w_ik b_nbr_ik
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "cells.hpp"
#include "kernel_fn.hpp"
#include "solvers.hpp"

using abh::DevBuf;
using abh::GridParams;
using abh::KernelDesc;

namespace {

constexpr int kMaxK = 64;     // neighbour pattern size limit (one per lane)
constexpr int kLsqMaxK = 32;  // least-squares rows: k x k normal equations

__device__ __forceinline__ bool pair_less(double ad, int ai, double bd, int bi) {
  return ad < bd || (ad == bd && ai < bi);
}

// ---------------------------------------------------------------------------
// k nearest neighbours
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_knn(const double *__restrict__ pos,
                                             const uint32_t *__restrict__ keys,
                                             const uint32_t *__restrict__ beg,
                                             const uint32_t *__restrict__ end, GridParams g,
                                             int64_t n, int k, int32_t *__restrict__ nbr,
                                             double *__restrict__ dist2) {
  const int lane = threadIdx.x & 63;
  const int64_t i = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int K1 = k - 1;
  double xi[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = pos[i * D + d];
  int iv[3] = {0, 0, 0};
  {
    int64_t r = keys[i];
    for (int d = D - 1; d >= 0; --d) {
      iv[d] = int(r % g.size[d]);
      r /= g.size[d];
    }
  }
  double ld = INFINITY;
  int li = INT32_MAX;
  for (int w = 1; K1 > 0; ++w) {
    ld = INFINITY;
    li = INT32_MAX;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool whole = true;
    for (int d = 0; d < D; ++d) {
      lo[d] = max(iv[d] - w, 0);
      hi[d] = min(iv[d] + w, g.size[d] - 1);
      whole = whole && lo[d] == 0 && hi[d] == g.size[d] - 1;
    }
    for (int x0 = lo[0]; x0 <= hi[0]; ++x0)
      for (int x1 = lo[1]; x1 <= hi[1]; ++x1)
        for (int x2 = lo[2]; x2 <= hi[2]; ++x2) {
          const int64_t b = (int64_t(x0) * g.size[1] + x1) * g.size[2] + x2;
          const int b0 = int(beg[b]), b1 = int(end[b]);
          for (int j0 = b0; j0 < b1; j0 += 64) {
            const int j = j0 + lane;
            double cd = INFINITY;
            if (j < b1 && j != i) {
              double s = 0.0;
#pragma unroll
              for (int d = 0; d < D; ++d) {
                const double dx = pos[int64_t(j) * D + d] - xi[d];
                s = s + dx * dx;
              }
              cd = s;
            }
            const double wd = __shfl(ld, K1 - 1, 64);
            const int wi = __shfl(li, K1 - 1, 64);
            uint64_t mask = __ballot(cd != INFINITY && pair_less(cd, j, wd, wi));
            while (mask) {
              const int src = __builtin_ctzll(mask);
              mask &= mask - 1;
              const double nd = __shfl(cd, src, 64);
              const int ni = __shfl(j, src, 64);
              const bool lt = lane < K1 && pair_less(ld, li, nd, ni);
              const int p = __popcll(__ballot(lt));
              const double ud = __shfl_up(ld, 1, 64);
              const int ui = __shfl_up(li, 1, 64);
              if (p < K1) {
                if (lane > p && lane < K1) {
                  ld = ud;
                  li = ui;
                }
                if (lane == p) {
                  ld = nd;
                  li = ni;
                }
              }
            }
          }
        }
    if (whole || K1 <= 0) break;
    const double far = __shfl(ld, K1 - 1, 64);
    if (far == INFINITY) continue;  // fewer than k-1 particles in the block
    const double rk = sqrt(far);
    bool inside = true;
    for (int d = 0; d < D; ++d) {
      const double blo = double(lo[d]) * g.side[d] + g.lo[d];
      const double bhi = double(hi[d] + 1) * g.side[d] + g.lo[d];
      if ((lo[d] > 0 && xi[d] - rk < blo) || (hi[d] < g.size[d] - 1 && xi[d] + rk >= bhi))
        inside = false;
    }
    if (inside) break;
  }
  int32_t *out = nbr + i * k;
  if (lane == 0) {
    out[0] = int32_t(i);
    if (dist2) dist2[i * k] = 0.0;
  }
  if (lane < K1) {
    const bool ok = ld != INFINITY;
    out[1 + lane] = ok ? li : -1;
    if (dist2) dist2[i * k + 1 + lane] = ok ? ld : -1.0;
  }
}

// ---------------------------------------------------------------------------
// least-squares approximate-inverse rows
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ double coeff(const double *a, const double *b, bool same,
                                        const KernelDesc &k, double r2, double shift) {
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double dx = b[d] - a[d];
    s = s + dx * dx;
  }
  double v = s <= r2 ? abh::phi64(k, s) : 0.0;
  if (same) v = v + shift;
  return v;
}

template <int D>
__global__ __launch_bounds__(256) void k_lsq(const double *__restrict__ pos,
                                             const uint32_t *__restrict__ beg,
                                             const uint32_t *__restrict__ end, GridParams g,
                                             int64_t n, int k, const int32_t *__restrict__ nbr,
                                             KernelDesc kern, double r2, double shift,
                                             double *__restrict__ w, uint8_t *__restrict__ fail) {
  __shared__ double ps[kLsqMaxK][D];
  __shared__ int32_t sid[kLsqMaxK];
  __shared__ double C[64][kLsqMaxK + 1];
  __shared__ double G[kLsqMaxK][kLsqMaxK + 1];
  __shared__ double rhs[kLsqMaxK];
  __shared__ double reach;
  const int64_t i = blockIdx.x;
  const int t = threadIdx.x;
  if (t < k) {
    const int32_t s = nbr[i * k + t];
    sid[t] = s;
#pragma unroll
    for (int d = 0; d < D; ++d) ps[t][d] = s >= 0 ? pos[int64_t(s) * D + d] : 0.0;
    rhs[t] = 0.0;
  }
  __syncthreads();
  if (t == 0) {
    double m = 0.0;
    for (int q = 0; q < k; ++q) {
      if (sid[q] < 0) continue;
      double s = 0.0;
      for (int d = 0; d < D; ++d) {
        const double dx = ps[q][d] - pos[i * D + d];
        s += dx * dx;
      }
      m = fmax(m, sqrt(s));
    }
    reach = sqrt(r2) + m;
  }
  __syncthreads();
  const double R = reach;
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < D; ++d) {
    const double x = pos[i * D + d];
    lo[d] = max(int(floor((x - R - g.lo[d]) * g.inv[d])), 0);
    hi[d] = min(int(floor((x + R - g.lo[d]) * g.inv[d])), g.size[d] - 1);
    lo[d] = min(lo[d], g.size[d] - 1);
    hi[d] = max(hi[d], 0);
  }
  const int kk = k * k;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int x0 = lo[0]; x0 <= hi[0]; ++x0)
    for (int x1 = lo[1]; x1 <= hi[1]; ++x1)
      for (int x2 = lo[2]; x2 <= hi[2]; ++x2) {
        const int64_t b = (int64_t(x0) * g.size[1] + x1) * g.size[2] + x2;
        const int b0 = int(beg[b]), b1 = int(end[b]);
        for (int j0 = b0; j0 < b1; j0 += 64) {
          const int m = min(64, b1 - j0);
          __syncthreads();
          for (int e = t; e < 64 * k; e += 256) {
            const int c = e / k, a = e - c * k;
            double v = 0.0;
            if (c < m && sid[a] >= 0) {
              const int p = j0 + c;
              double xp[3] = {0, 0, 0};
#pragma unroll
              for (int d = 0; d < D; ++d) xp[d] = pos[int64_t(p) * D + d];
              v = coeff<D>(xp, ps[a], p == sid[a], kern, r2, shift);
            }
            C[c][a] = v;
          }
          __syncthreads();
          for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            if (e < kk) {
              const int a = e / k, bb = e - a * k;
              double s = acc[q];
              for (int c = 0; c < m; ++c) s = s + C[c][a] * C[c][bb];
              acc[q] = s;
            }
          }
          if (i >= j0 && i < j0 + m && t < k) rhs[t] = C[i - j0][t];
        }
      }
  for (int q = 0; q < 4; ++q) {
    const int e = t + 256 * q;
    if (e < kk) G[e / k][e % k] = acc[q];
  }
  __syncthreads();
  // Cholesky of the k x k normal equations, then the two triangular solves
  __shared__ int bad;
  if (t == 0) bad = 0;
  for (int j = 0; j < k; ++j) {
    __syncthreads();
    if (t == 0) {
      const double d = G[j][j];
      if (!(d > 0.0)) bad = 1;
      G[j][j] = sqrt(d > 0.0 ? d : 1.0);
    }
    __syncthreads();
    if (t > j && t < k) G[t][j] = G[t][j] / G[j][j];
    __syncthreads();
    if (t > j && t < k)
      for (int c = j + 1; c <= t; ++c) G[t][c] = G[t][c] - G[t][j] * G[c][j];
  }
  __syncthreads();
  if (t == 0) {
    double y[kLsqMaxK];
    for (int a = 0; a < k; ++a) {
      double v = rhs[a];
      for (int q = 0; q < a; ++q) v -= G[a][q] * y[q];
      y[a] = v / G[a][a];
    }
    for (int a = k - 1; a >= 0; --a) {
      double v = y[a];
      for (int q = a + 1; q < k; ++q) v -= G[q][a] * y[q];
      y[a] = v / G[a][a];
    }
    for (int a = 0; a < k; ++a) w[i * k + a] = bad ? 0.0 : y[a];
    if (bad) fail[0] = 1;
  }
}

// ---------------------------------------------------------------------------
// patch local solves: A_P = K(x_a, x_b) (cut at the radius) + shift I over a
// particle's k nearest neighbours P (itself first); w = A_P^{-1} e_1, the
// particle's row of the patch's local inverse.  Missing neighbours (fewer
// than k particles) pad A_P with identity rows and get weight 0.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ double patch_entry(const double *__restrict__ pos,
                                              const int32_t *__restrict__ nb, int a, int b,
                                              const KernelDesc &kern, double r2, double shift) {
  const int32_t ia = nb[a], ib = nb[b];
  if (ia < 0 || ib < 0) return a == b ? 1.0 : 0.0;
  return coeff<D>(pos + int64_t(ia) * D, pos + int64_t(ib) * D, a == b, kern, r2, shift);
}

template <int D>
__global__ __launch_bounds__(64) void k_patch(const double *__restrict__ pos, int64_t n, int k,
                                              const int32_t *__restrict__ nbr, KernelDesc kern,
                                              double r2, double shift, double *__restrict__ w,
                                              uint8_t *__restrict__ fail) {
  __shared__ double G[kMaxK][kMaxK + 1];
  __shared__ int bad;
  const int64_t i = blockIdx.x;
  const int t = threadIdx.x;
  const int32_t *nb = nbr + i * k;
  for (int e = t; e < k * k; e += 64) {
    const int a = e / k, b = e - a * k;
    G[a][b] = patch_entry<D>(pos, nb, a, b, kern, r2, shift);
  }
  if (t == 0) bad = 0;
  // Cholesky, column by column (row t of the lower factor kept by lane t)
  for (int j = 0; j < k; ++j) {
    __syncthreads();
    if (t == 0) {
      const double d = G[j][j];
      if (!(d > 0.0)) bad = 1;
      G[j][j] = sqrt(d > 0.0 ? d : 1.0);
    }
    __syncthreads();
    if (t > j && t < k) G[t][j] = G[t][j] / G[j][j];
    __syncthreads();
    if (t > j && t < k)
      for (int c = j + 1; c <= t; ++c) G[t][c] = G[t][c] - G[t][j] * G[c][j];
  }
  __syncthreads();
  if (t == 0) {
    double y[kMaxK];
    for (int a = 0; a < k; ++a) {  // L y = e_1
      double v = a == 0 ? 1.0 : 0.0;
      for (int q = 0; q < a; ++q) v -= G[a][q] * y[q];
      y[a] = v / G[a][a];
    }
    for (int a = k - 1; a >= 0; --a) {  // L^T w = y
      double v = y[a];
      for (int q = a + 1; q < k; ++q) v -= G[q][a] * y[q];
      y[a] = v / G[a][a];
    }
    for (int a = 0; a < k; ++a) w[i * k + a] = (bad || nb[a] < 0) ? 0.0 : y[a];
    if (bad) fail[0] = 1;
  }
}

// the patch matrices in FP32, times scale: mats[(i k + a) k + b]
template <int D>
__global__ __launch_bounds__(256) void k_patch_mats(const double *__restrict__ pos, int64_t n,
                                                    int k, const int32_t *__restrict__ nbr,
                                                    KernelDesc kern, double r2, double shift,
                                                    double scale, float *__restrict__ mats) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int64_t kk = int64_t(k) * k;
  if (e >= n * kk) return;
  const int64_t i = e / kk;
  const int r = int(e - i * kk), a = r / k, b = r - a * k;
  mats[e] = float(scale * patch_entry<D>(pos, nbr + i * k, a, b, kern, r2, shift));
}

__global__ void k_features(const double *__restrict__ pos, int64_t n, int k, int D,
                           const int32_t *__restrict__ nbr, double inv_scale,
                           float *__restrict__ f) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  const int32_t j = nbr[e];
  for (int d = 0; d < D; ++d)
    f[e * D + d] = j >= 0 ? float((pos[int64_t(j) * D + d] - pos[i * D + d]) * inv_scale) : 0.0f;
}

__global__ void k_ainv_apply(const int32_t *__restrict__ nbr, const double *__restrict__ w,
                             int64_t n, int k, const double *__restrict__ b,
                             double *__restrict__ x) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int q = 0; q < k; ++q) {
    const int32_t j = nbr[i * k + q];
    if (j >= 0) s = s + w[i * k + q] * b[j];
  }
  x[i] = s;
}

__global__ void k_scale_to_f64(const float *__restrict__ y, int64_t m, double s,
                               double *__restrict__ w) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < m) w[e] = double(y[e]) * s;
}

// ---------------------------------------------------------------------------
// MLP: FP32 GEMM with fused epilogue
// ---------------------------------------------------------------------------
enum Act { kNone = 0, kSilu = 1 };

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// C[M x N] = op(A) op(B) (+ C if beta) (+ bias[n]); Z = pre-activation copy;
// C = act(pre-activation).  Row-major operands; op(A)[m][k] =
// TA ? A[k*lda + m] : A[m*lda + k], op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n].
// With part set, block z takes k in [z * kchunk, (z + 1) * kchunk) and writes
// its raw partial product to part[z][m][n] (no epilogue); k_split_sum adds
// the partials in z order.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float *__restrict__ A,
                                              int lda, const float *__restrict__ B, int ldb,
                                              float *__restrict__ C, int ldc, int beta,
                                              const float *__restrict__ bias, int act,
                                              float *__restrict__ Z, int kchunk = 0,
                                              float *__restrict__ part = nullptr) {
  __shared__ float As[16][64 + 4];
  __shared__ float Bs[16][64 + 4];
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  float acc[4][4] = {};
  const int kb = part ? int(blockIdx.z) * kchunk : 0;
  const int ke = part ? min(K, kb + kchunk) : K;
  for (int k0 = kb; k0 < ke; k0 += 16) {
    for (int e = t; e < 1024; e += 256) {
      int kk, mm;
      if (TA) {
        mm = e & 63;
        kk = e >> 6;
      } else {
        kk = e & 15;
        mm = e >> 4;
      }
      const int m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < M && k < ke) ? (TA ? A[int64_t(k) * lda + m] : A[int64_t(m) * lda + k]) : 0.0f;
    }
    for (int e = t; e < 1024; e += 256) {
      int kk, nn;
      if (TB) {
        kk = e & 15;
        nn = e >> 4;
      } else {
        nn = e & 63;
        kk = e >> 6;
      }
      const int nx = n0 + nn, k = k0 + kk;
      Bs[kk][nn] = (nx < N && k < ke) ? (TB ? B[int64_t(nx) * ldb + k] : B[int64_t(k) * ldb + nx]) : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ty * 4 + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tx * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = acc[r][c] + a[r] * b[c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = m0 + ty * 4 + r;
    if (m >= M) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int nx = n0 + tx * 4 + c;
      if (nx >= N) continue;
      float v = acc[r][c];
      if (part) {
        part[(int64_t(blockIdx.z) * M + m) * N + nx] = v;
        continue;
      }
      float *cp = &C[int64_t(m) * ldc + nx];
      if (beta) v = v + *cp;
      if (bias) v = v + bias[nx];
      if (Z) Z[int64_t(m) * N + nx] = v;
      *cp = act == kSilu ? silu(v) : v;
    }
  }
}

// dA (in place) = dA * silu'(Z)
__global__ void k_silu_back(float *__restrict__ dA, const float *__restrict__ Z, int64_t m) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < m) dA[e] = dA[e] * silu_grad(Z[e]);
}

// C[m][n] = sum over z (in order) of part[z][m][n]
__global__ void k_split_sum(int M, int N, int nz, const float *__restrict__ part,
                            float *__restrict__ C, int ldc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(M) * N) return;
  float v = 0.0f;
  for (int z = 0; z < nz; ++z) v = v + part[int64_t(z) * M * N + e];
  C[(e / N) * ldc + e % N] = v;
}

// part[z][c] = sum of dZ[r][c] over rows r of chunk z (64 columns x 4 row
// phases per block, the phases combined in order); k_split_sum then adds
// the chunks in order (fixed order throughout)
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ dZ, int rows, int cols,
                                                int rchunk, float *__restrict__ part) {
  __shared__ float red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + tx;
  const int r0 = blockIdx.y * rchunk, r1 = min(rows, r0 + rchunk);
  float s = 0.0f;
  if (c < cols)
    for (int r = r0 + ty; r < r1; r += 4) s = s + dZ[int64_t(r) * cols + c];
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && c < cols)
    part[int64_t(blockIdx.y) * cols + c] = ((red[0][tx] + red[1][tx]) + red[2][tx]) + red[3][tx];
}

// dY = 2 (Y - T) / (rows * cols); partial[block] = sum (Y - T)^2 in FP64
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             int64_t m, float scale, float *__restrict__ dY,
                                             double *__restrict__ partial) {
  __shared__ double red[4];
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double s = 0.0;
  if (e < m) {
    const float d = Y[e] - T[e];
    dY[e] = 2.0f * d * scale;
    s = double(d) * double(d);
  }
  s = abh::wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Patch-residual loss: per row r (one wavefront, k <= 64 outputs)
// res = M_r y_r - e_1, loss_r = |res|^2, dY_r = 2 M_r^T res / rows;
// partial[block] = the block's sum of loss_r in FP64 (4 rows per block).
__global__ __launch_bounds__(256) void k_resid(const float *__restrict__ Y,
                                               const float *__restrict__ M, int64_t rows, int k,
                                               float inv_rows, float *__restrict__ dY,
                                               double *__restrict__ partial) {
  __shared__ float ys[4][kMaxK], rs[4][kMaxK];
  __shared__ double red[4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r = int64_t(blockIdx.x) * 4 + wv;
  const bool live = r < rows;
  if (live && lane < k) ys[wv][lane] = Y[r * k + lane];
  __syncthreads();
  double l2 = 0.0;
  if (live && lane < k) {
    const float *Mr = M + r * k * k;
    float v = lane == 0 ? -1.0f : 0.0f;
    for (int b = 0; b < k; ++b) v += Mr[int64_t(lane) * k + b] * ys[wv][b];
    rs[wv][lane] = v;
    l2 = double(v) * double(v);
  }
  __syncthreads();
  if (live && lane < k) {
    const float *Mr = M + r * k * k;
    float g = 0.0f;
    for (int a = 0; a < k; ++a) g += Mr[int64_t(a) * k + lane] * rs[wv][a];
    dY[r * k + lane] = 2.0f * g * inv_rows;
  }
  l2 = abh::wave_sum(l2);
  if (lane == 0) red[wv] = l2;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// out[step ? *step : 0] = scale * sum of the partials
__global__ void k_sum_partials(const double *__restrict__ partial, int nb, double scale,
                               double *__restrict__ out, const int32_t *__restrict__ step) {
  __shared__ double red[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += partial[b];
  s = abh::wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[step ? *step : 0] = ((red[0] + red[1]) + (red[2] + red[3])) * scale;
}

// Adam with the step size of training step *step (lr_steps, per epoch)
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float b1, float b2,
                       const float *__restrict__ lr_steps, const int32_t *__restrict__ step,
                       float eps) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float lr_t = lr_steps[*step];
  const float ge = g[e];
  const float me = b1 * m[e] + (1.0f - b1) * ge;
  const float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  p[e] = p[e] - lr_t * me / (sqrtf(ve) + eps);
}

// the rows of training step *step: perm[*step * nrows + r]
__global__ void k_gather_rows(const float *__restrict__ src, const int32_t *__restrict__ perm,
                              const int32_t *__restrict__ step, int nrows, int cols,
                              float *__restrict__ dst) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(nrows) * cols) return;
  const int64_t r = e / cols;
  const int c = int(e - r * cols);
  const int32_t *rows = perm + int64_t(*step) * nrows;
  dst[e] = src[int64_t(rows[r]) * cols + c];
}

// ---------------------------------------------------------------------------
// MLP: one training step in two kernels.  k_train_chunk takes kFR rows of
// the step's batch per block through the whole network -- forward with the
// activations in LDS, the loss, backward -- and writes the chunk's weight
// and bias gradients to its own slice of a partial buffer; k_train_update
// sums the slices in chunk order and applies Adam.  For networks of at most
// kFMaxL layers no wider than kFMaxW (the learned preconditioner's).
// ---------------------------------------------------------------------------

constexpr int kFR = 8;       // rows per block
constexpr int kFT = 1024;    // threads per block
constexpr int kFG = kFT / 128;  // thread groups of 128 (rows kFR / kFG each)
constexpr int kFMaxW = 128;  // widest layer
constexpr int kFMaxL = 4;    // layers

struct FusedNet {
  int L;
  int dims[kFMaxL + 1];
  int64_t w_off[kFMaxL], b_off[kFMaxL];
  int64_t nparams;
};

// floats of the block's LDS: input, per hidden layer Z and activation,
// output, two gradient buffers, the staged transposed weights
inline size_t fused_lds_floats(const FusedNet &f) {
  size_t n = size_t(kFR) * f.dims[0];
  for (int l = 0; l + 1 < f.L; ++l) n += 2 * size_t(kFR) * f.dims[l + 1];
  n += size_t(kFR) * f.dims[f.L];
  n += 2 * size_t(kFR) * kFMaxW;
  n += size_t(kFMaxW) * (kFMaxW + 1);
  return n;
}

// Ws[k * (dout + 1) + n] = W[n][k]; eight loads per thread in flight
__device__ __forceinline__ void stage_wt(const float *__restrict__ W, int din, int dout,
                                         float *__restrict__ Ws) {
  const int total = din * dout;
  for (int base = 0; base < total; base += kFT * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = base + u * kFT + int(threadIdx.x);
      v[u] = e < total ? W[e] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = base + u * kFT + int(threadIdx.x);
      if (e < total) {
        const int n = e / din, k = e - n * din;
        Ws[k * (dout + 1) + n] = v[u];
      }
    }
  }
}

// mats != nullptr: patch-residual loss (M_r: k x k per row, k = n_out);
// otherwise the mean squared error against t.  rows: the batch size.
__global__ __launch_bounds__(kFT) void k_train_chunk(FusedNet f, const float *__restrict__ P,
                                                     const float *__restrict__ x,
                                                     const float *__restrict__ t,
                                                     const float *__restrict__ mats,
                                                     const int32_t *__restrict__ perm,
                                                     const int32_t *__restrict__ step, int rows,
                                                     float *__restrict__ part,
                                                     double *__restrict__ lossp) {
  extern __shared__ __align__(16) float sm[];
  const int L = f.L, nin = f.dims[0], nout = f.dims[L];
  const int chunk = blockIdx.x;
  const int r0 = chunk * kFR;
  const int nr = min(kFR, rows - r0);
  const int32_t *rid = perm + int64_t(*step) * rows + r0;
  // LDS carve-up
  float *act[kFMaxL + 1], *Z[kFMaxL];
  float *q = sm;
  act[0] = q;
  q += kFR * nin;
  for (int l = 0; l + 1 < L; ++l) {
    Z[l] = q;
    q += kFR * f.dims[l + 1];
    act[l + 1] = q;
    q += kFR * f.dims[l + 1];
  }
  float *Y = q;
  q += kFR * nout;
  float *dA = q;
  q += kFR * kFMaxW;
  float *dB = q;
  q += kFR * kFMaxW;
  float *Ws = q;
  const int tid = threadIdx.x;
  for (int e = tid; e < kFR * nin; e += kFT) {
    const int r = e / nin, c = e - r * nin;
    act[0][e] = r < nr ? x[int64_t(rid[r]) * nin + c] : 0.0f;
  }
  // forward
  for (int l = 0; l < L; ++l) {
    const int din = f.dims[l], dout = f.dims[l + 1];
    __syncthreads();
    stage_wt(P + f.w_off[l], din, dout, Ws);
    __syncthreads();
    const int n = tid & 127, g = tid >> 7;
    if (n < dout) {
      float acc[kFR / kFG];
#pragma unroll
      for (int i = 0; i < kFR / kFG; ++i) acc[i] = 0.0f;
      const float *in = act[l] + (g * (kFR / kFG)) * din;
      for (int k = 0; k < din; ++k) {
        const float w = Ws[k * (dout + 1) + n];
#pragma unroll
        for (int i = 0; i < kFR / kFG; ++i) acc[i] = acc[i] + in[i * din + k] * w;
      }
      const float b = P[f.b_off[l] + n];
#pragma unroll
      for (int i = 0; i < kFR / kFG; ++i) {
        const int r = g * (kFR / kFG) + i;
        const float z = acc[i] + b;
        if (l + 1 < L) {
          Z[l][r * dout + n] = z;
          act[l + 1][r * dout + n] = silu(z);
        } else {
          Y[r * dout + n] = z;
        }
      }
    }
  }
  __syncthreads();
  // loss and dY (into dA), rows past nr contribute nothing
  __shared__ double red[kFT / 64];
  double l2 = 0.0;
  if (mats) {
    // the rows' k x k matrices staged in the weight buffer (row stride k+1,
    // as many rows at a time as fit), then one wavefront per row
    const int wv = tid >> 6, lane = tid & 63;
    const int ms = nout + 1, per = nout * ms;
    const int grp = max(1, min(kFR, (kFMaxW * (kFMaxW + 1)) / per));
    for (int rg = 0; rg < kFR; rg += grp) {
      const int re = min(kFR, rg + grp);
      __syncthreads();
      const int tot = (re - rg) * nout * nout;
      for (int base = 0; base < tot; base += kFT * 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = base + u * kFT + tid;
          const int rr = rg + e / (nout * nout);
          v[u] = (e < tot && rr < nr)
                     ? mats[int64_t(rid[rr]) * nout * nout + (e % (nout * nout))]
                     : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = base + u * kFT + tid;
          if (e < tot) {
            const int rr = rg + e / (nout * nout), j = (e / nout) % nout, b = e % nout;
            Ws[(rr - rg) * per + j * ms + b] = v[u];
          }
        }
      }
      __syncthreads();
      for (int r = rg + wv; r < re; r += kFT / 64) {
        const float *Mr = Ws + (r - rg) * per;
        float v = 0.0f;
        if (r < nr && lane < nout) {
          v = lane == 0 ? -1.0f : 0.0f;
          for (int b = 0; b < nout; ++b) v += Mr[lane * ms + b] * Y[r * nout + b];
          l2 += double(v) * double(v);
        }
        dB[r * kFMaxW + lane] = v;  // the residual, staged for M^T res
      }
      __syncthreads();
      for (int r = rg + wv; r < re; r += kFT / 64) {
        if (lane >= nout) continue;
        const float *Mr = Ws + (r - rg) * per;
        float gsum = 0.0f;
        if (r < nr)
          for (int a = 0; a < nout; ++a) gsum += Mr[a * ms + lane] * dB[r * kFMaxW + a];
        dA[r * nout + lane] = 2.0f * gsum / float(rows);
      }
    }
  } else {
    const float scale = 1.0f / float(int64_t(rows) * nout);
    for (int e = tid; e < kFR * nout; e += kFT) {
      const int r = e / nout, c = e - r * nout;
      float d = 0.0f;
      if (r < nr) d = Y[e] - t[int64_t(rid[r]) * nout + c];
      dA[e] = 2.0f * d * scale;
      l2 += double(d) * double(d);
    }
  }
  l2 = abh::wave_sum(l2);
  if ((tid & 63) == 0) red[tid >> 6] = l2;
  __syncthreads();
  if (tid == 0) {
    double t8 = 0.0;
    for (int w = 0; w < kFT / 64; ++w) t8 += red[w];
    lossp[chunk] = t8;
  }
  // backward: dA holds dZ of layer l (kFR x dout)
  float *part_c = part + int64_t(chunk) * f.nparams;
  for (int l = L - 1; l >= 0; --l) {
    const int din = f.dims[l], dout = f.dims[l + 1];
    const float *ain = act[l];
    // dW[n][k] = sum_r dZ[r][n] act[r][k]: lane k, 16 outputs n = nb + kFG j
    {
      const int k = tid & 127, g = tid >> 7;
      if (k < din) {
        for (int nb = g; nb < dout; nb += 16 * kFG) {
          float acc[16];
#pragma unroll
          for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
          for (int r = 0; r < kFR; ++r) {
            const float a = ain[r * din + k];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
              const int n = nb + kFG * j;
              if (n < dout) acc[j] = acc[j] + dA[r * dout + n] * a;
            }
          }
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            const int n = nb + kFG * j;
            if (n < dout) part_c[f.w_off[l] + int64_t(n) * din + k] = acc[j];
          }
        }
      }
      if (tid < dout) {
        float sb = 0.0f;
        for (int r = 0; r < kFR; ++r) sb = sb + dA[r * dout + tid];
        part_c[f.b_off[l] + tid] = sb;
      }
    }
    if (l == 0) break;
    // dact[r][k] = sum_n dZ[r][n] W[n][k], then through the SiLU of layer l-1
    __syncthreads();
    stage_wt(P + f.w_off[l], din, dout, Ws);
    __syncthreads();
    {
      const int k = tid & 127, g = tid >> 7;
      if (k < din) {
        float acc[kFR / kFG];
#pragma unroll
        for (int i = 0; i < kFR / kFG; ++i) acc[i] = 0.0f;
        for (int n = 0; n < dout; ++n) {
          const float w = Ws[k * (dout + 1) + n];
#pragma unroll
          for (int i = 0; i < kFR / kFG; ++i)
            acc[i] = acc[i] + dA[(g * (kFR / kFG) + i) * dout + n] * w;
        }
#pragma unroll
        for (int i = 0; i < kFR / kFG; ++i) {
          const int r = g * (kFR / kFG) + i;
          dB[r * din + k] = acc[i] * silu_grad(Z[l - 1][r * din + k]);
        }
      }
    }
    __syncthreads();
    float *tmp = dA;
    dA = dB;
    dB = tmp;
  }
}

// grad[p] = sum over chunks (in order) of part[c][p], then Adam with the
// step size of step *step; block 0 also writes the step's loss
__global__ __launch_bounds__(256) void k_train_update(
    float *__restrict__ p, float *__restrict__ m, float *__restrict__ v, int64_t n,
    const float *__restrict__ part, int nchunks, float b1, float b2,
    const float *__restrict__ lr_steps, const int32_t *__restrict__ step, float eps,
    const double *__restrict__ lossp, double loss_scale, double *__restrict__ losses) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (blockIdx.x == 0) {
    __shared__ double red[4];
    double s = 0.0;
    for (int c = threadIdx.x; c < nchunks; c += 256) s += lossp[c];
    s = abh::wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) losses[*step] = ((red[0] + red[1]) + (red[2] + red[3])) * loss_scale;
  }
  if (e >= n) return;
  float g = 0.0f;
  int c = 0;
  for (; c + 8 <= nchunks; c += 8) {  // eight loads in flight, added in order
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[int64_t(c + u) * n + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) g = g + v[u];
  }
  for (; c < nchunks; ++c) g = g + part[int64_t(c) * n + e];
  const float lr_t = lr_steps[*step];
  const float me = b1 * m[e] + (1.0f - b1) * g;
  const float ve = b2 * v[e] + (1.0f - b2) * g * g;
  m[e] = me;
  v[e] = ve;
  p[e] = p[e] - lr_t * me / (sqrtf(ve) + eps);
}

__global__ void k_set_step(int32_t *__restrict__ step, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = v;
}

__global__ void k_next_step(int32_t *__restrict__ step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// ---------------------------------------------------------------------------
// MLP: BF16 MFMA inference
// ---------------------------------------------------------------------------
typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kBfMaxDim = 128;
constexpr int kBfLd = kBfMaxDim + 8;  // row pitch (bf16) of the activation tile
constexpr int kBfMaxLayers = 8;

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return uint16_t(u >> 16);
}

__global__ void k_to_bf16(const float *__restrict__ src, int64_t n, uint16_t *__restrict__ dst) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) dst[e] = to_bf16(src[e]);
}

struct BfNet {
  int nl;                      // layers
  int dims[kBfMaxLayers + 1];  // widths
  int64_t w_off[kBfMaxLayers]; // bf16 weights (out x in, row-major)
  int64_t b_off[kBfMaxLayers]; // FP32 biases in the parameter array
};

// One wavefront = 32 rows.  MFMA 32x32x16 (BF16): lane l supplies
// A[l % 32][8 (l / 32) + j] and B[8 (l / 32) + j][l % 32], j < 8, and holds
// C[8 (i / 4) + 4 (l / 32) + i % 4][l % 32] in accumulator register i.
__global__ __launch_bounds__(256) void k_mlp_bf16(BfNet net, const uint16_t *__restrict__ wb,
                                                  const float *__restrict__ params,
                                                  const float *__restrict__ x, int64_t n,
                                                  float *__restrict__ y) {
  __shared__ uint16_t act[4][32][kBfLd];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row0 = (int64_t(blockIdx.x) * 4 + wv) * 32;
  if (row0 >= n) return;
  uint16_t(*A)[kBfLd] = act[wv];
  const int nin = net.dims[0];
  for (int e = lane; e < 32 * nin; e += 64) {
    const int r = e / nin, c = e - r * nin;
    const int64_t row = row0 + r;
    A[r][c] = row < n ? to_bf16(x[row * nin + c]) : uint16_t(0);
  }
  const int half = lane >> 5, l32 = lane & 31;
  for (int L = 0; L < net.nl; ++L) {
    const int din = net.dims[L], dout = net.dims[L + 1];
    const int nblk = dout / 32;
    f32x16 acc[kBfMaxDim / 32];
#pragma unroll
    for (int b = 0; b < kBfMaxDim / 32; ++b)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[b][q] = 0.0f;
    const uint16_t *W = wb + net.w_off[L];
    for (int ks = 0; ks < din; ks += 16) {
      bf16x8 a;
      const uint16_t *ap = &A[l32][ks + 8 * half];
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = short(ap[j]);
#pragma unroll
      for (int b = 0; b < kBfMaxDim / 32; ++b) {
        if (b < nblk) {
          const uint16_t *bp = W + int64_t(b * 32 + l32) * din + ks + 8 * half;
          bf16x8 bv;
#pragma unroll
          for (int j = 0; j < 8; ++j) bv[j] = short(bp[j]);
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc[b], 0, 0, 0);
        }
      }
    }
    const float *bias = params + net.b_off[L];
    const bool last = L == net.nl - 1;
#pragma unroll
    for (int b = 0; b < kBfMaxDim / 32; ++b) {
      if (b >= nblk) continue;
      const int col = b * 32 + l32;
      const float bc = bias[col];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int r = 8 * (q / 4) + 4 * half + (q % 4);
        const float v = acc[b][q] + bc;
        if (last) {
          const int64_t row = row0 + r;
          if (row < n) y[row * dout + col] = v;
        } else {
          A[r][col] = to_bf16(silu(v));
        }
      }
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct abh_mlp {
  abh_context *ctx = nullptr;
  std::vector<int> dims;  // n_in, hidden..., n_out
  std::vector<int64_t> w_off, b_off;
  int64_t nparams = 0;
  DevBuf params, grads, am, av, wbf, loss;
  DevBuf ws;      // activations of the training / FP32 forward passes
  DevBuf gsplit;  // split-K partial weight / bias gradients
  int64_t adam_t = 0;
  int layers() const { return int(dims.size()) - 1; }
};

struct abh_ainv {
  abh_context *ctx = nullptr;
  int64_t n = 0;
  int k = 0;
  int dim = 2;
  const double *pos = nullptr;  // the cell list's positions (gather order)
  DevBuf nbr, w;
  // the operator A = K + shift I the rows were built from (lsq, patch)
  bool has_op = false;
  KernelDesc kern{};
  double r2 = 0.0, shift = 0.0;
};

namespace {

void knn_run(abh_context *ctx, const abh_cells *c, int k, int32_t *nbr, double *d2) {
  const int64_t n = c->n_alive;
  if (n == 0) return;
  const GridParams g = c->params();
  const unsigned blocks = unsigned(abh::ceil_div(n, 4));
  const double *pos = c->pos.as<double>();
  const uint32_t *keys = c->keys.as<uint32_t>(), *b = c->begin.as<uint32_t>(),
                 *e = c->end.as<uint32_t>();
  if (c->dim == 1) k_knn<1><<<blocks, 256, 0, ctx->stream>>>(pos, keys, b, e, g, n, k, nbr, d2);
  else if (c->dim == 2) k_knn<2><<<blocks, 256, 0, ctx->stream>>>(pos, keys, b, e, g, n, k, nbr, d2);
  else k_knn<3><<<blocks, 256, 0, ctx->stream>>>(pos, keys, b, e, g, n, k, nbr, d2);
  abh::check_launch();
}

void check_cells(const abh_cells *c) {
  ABH_REQUIRE(c, ABH_ERR_INVALID_ARG, "cells is null");
  for (int d = 0; d < c->dim; ++d)
    ABH_REQUIRE(!c->periodic[d], ABH_ERR_UNSUPPORTED, "needs a non-periodic cell list");
}

// FP32 forward over rows, keeping pre-activations Z[l] and activations
// act[l] (act[0] = x) in the workspace; returns the output pointer.
struct Fwd {
  std::vector<const float *> act;
  std::vector<float *> z;
  float *out = nullptr;
};

Fwd forward_f32(abh_mlp *m, int64_t rows, const float *x, float *y) {
  hipStream_t s = m->ctx->stream;
  const int nl = m->layers();
  int64_t need = 0;
  for (int l = 1; l < nl; ++l) need += 2 * rows * m->dims[l];
  m->ws.ensure(size_t(need) * 4 + 256);
  float *p = m->ws.as<float>();
  Fwd f;
  f.act.push_back(x);
  const float *P = m->params.as<float>();
  for (int l = 0; l < nl; ++l) {
    const int din = m->dims[l], dout = m->dims[l + 1];
    float *z = nullptr, *out;
    if (l < nl - 1) {
      z = p;
      p += rows * dout;
      out = p;
      p += rows * dout;
    } else {
      out = y;
    }
    dim3 grid(unsigned(abh::ceil_div(dout, 64)), unsigned(abh::ceil_div(rows, 64)));
    k_gemm<false, true><<<grid, 256, 0, s>>>(int(rows), dout, din, f.act.back(), din,
                                             P + m->w_off[l], din, out, dout, 0,
                                             P + m->b_off[l], l < nl - 1 ? kSilu : kNone, z);
    abh::check_launch();
    if (l < nl - 1) {
      f.z.push_back(z);
      f.act.push_back(out);
    }
  }
  f.out = y;
  return f;
}

// loss (device double) and gradient (device, nparams) over rows: the mean
// squared error against t, or with mats (k x k per row, k = n_out) the mean
// patch residual |M_r y_r - e_1|^2; scratch holds dY and the backward
// activations
void loss_grad(abh_mlp *m, int64_t rows, const float *x, const float *t, const float *mats,
               float *grad, double *loss_dev, DevBuf &scratch, const int32_t *step = nullptr) {
  hipStream_t s = m->ctx->stream;
  const int nl = m->layers();
  const int nout = m->dims.back();
  int64_t maxw = 0;
  for (int d : m->dims) maxw = std::max<int64_t>(maxw, d);
  const int64_t nmse = abh::ceil_div(rows * nout, 256), nres = abh::ceil_div(rows, 4);
  const int64_t nblk = std::max(nmse, nres);  // partial sums of either loss
  scratch.ensure(size_t(rows * nout + 2 * rows * maxw + 1) * 4 + size_t(nblk) * 8 + 256);
  float *Y = scratch.as<float>();
  float *dA = Y + rows * nout;          // gradient w.r.t. a layer output
  float *dB = dA + rows * maxw;         // gradient w.r.t. a layer input
  int64_t poff = rows * nout + 2 * rows * maxw;
  poff += poff & 1;  // 8-byte alignment of the FP64 partial sums
  double *partial = reinterpret_cast<double *>(scratch.as<float>() + poff);
  Fwd f = forward_f32(m, rows, x, Y);
  if (mats) {
    k_resid<<<unsigned(nres), 256, 0, s>>>(Y, mats, rows, nout, 1.0f / float(rows), dA, partial);
    abh::check_launch();
    k_sum_partials<<<1, 256, 0, s>>>(partial, int(nres), 1.0 / double(rows), loss_dev, step);
  } else {
    k_mse<<<unsigned(nmse), 256, 0, s>>>(Y, t, rows * nout, 1.0f / float(rows * nout), dA,
                                         partial);
    abh::check_launch();
    k_sum_partials<<<1, 256, 0, s>>>(partial, int(nmse), 1.0 / double(rows * nout), loss_dev,
                                     step);
  }
  abh::check_launch();
  const float *P = m->params.as<float>();
  for (int l = nl - 1; l >= 0; --l) {
    const int din = m->dims[l], dout = m->dims[l + 1];
    if (l < nl - 1) {  // through the SiLU of this hidden layer
      k_silu_back<<<unsigned(abh::ceil_div(rows * dout, 256)), 256, 0, s>>>(dA, f.z[l],
                                                                            rows * dout);
      abh::check_launch();
    }
    // dW = dZ^T act_in  (dout x din): the rows (the k dimension) are split
    // into chunks of 256 over the grid's z, the partial products added in
    // chunk order; db = column sums of dZ, split the same way
    {
      const int kz = int(std::min<int64_t>(abh::ceil_div(rows, 256), 512));
      const int kchunk = int(abh::ceil_div(abh::ceil_div(rows, kz), 16) * 16);
      const int nz = int(abh::ceil_div(rows, kchunk));
      m->gsplit.ensure(size_t(nz) * (size_t(dout) * din + dout) * sizeof(float));
      float *pw = m->gsplit.as<float>();
      float *pb = pw + int64_t(nz) * dout * din;
      dim3 grid(unsigned(abh::ceil_div(din, 64)), unsigned(abh::ceil_div(dout, 64)), unsigned(nz));
      k_gemm<true, false><<<grid, 256, 0, s>>>(dout, din, int(rows), dA, dout, f.act[l], din,
                                               nullptr, din, 0, nullptr, kNone, nullptr, kchunk,
                                               pw);
      abh::check_launch();
      k_split_sum<<<unsigned(abh::ceil_div(int64_t(dout) * din, 256)), 256, 0, s>>>(
          dout, din, nz, pw, grad + m->w_off[l], din);
      abh::check_launch();
      dim3 gb(unsigned(abh::ceil_div(dout, 64)), unsigned(nz));
      k_colsum<<<gb, 256, 0, s>>>(dA, int(rows), dout, kchunk, pb);
      abh::check_launch();
      k_split_sum<<<unsigned(abh::ceil_div(dout, 256)), 256, 0, s>>>(1, dout, nz, pb,
                                                                      grad + m->b_off[l], dout);
      abh::check_launch();
    }
    if (l > 0) {  // dact_in = dZ W  (rows x din)
      dim3 grid(unsigned(abh::ceil_div(din, 64)), unsigned(abh::ceil_div(rows, 64)));
      k_gemm<false, false><<<grid, 256, 0, s>>>(int(rows), din, dout, dA, dout,
                                                P + m->w_off[l], din, dB, din, 0, nullptr, kNone,
                                                nullptr);
      abh::check_launch();
      std::swap(dA, dB);
    }
  }
}

// Adam over shuffled full batches: targets t (n x n_out) for the mean
// squared error, or mats (n x n_out x n_out) for the patch residual.  The
// kernels of a step read the step's rows, loss slot and step size through a
// device step counter, so every step launches the same kernels with the
// same arguments; with ABH_TRAIN_GRAPH the second step is captured into a
// HIP graph (on a stream of the context's own) and every later step replays
// it.  ABH_TRAIN_FUSED runs a step as k_train_chunk + k_train_update.
void train_impl(abh_mlp *m, int64_t n, const float *x, const float *t, const float *mats,
                const abh_mlp_train_options *opts, float *loss_per_epoch) {
  ABH_REQUIRE(opts->epochs >= 0 && opts->batch >= 1 && opts->batch <= (1 << 20),
              ABH_ERR_INVALID_ARG, "invalid training options");
  ABH_REQUIRE((opts->flags & ~(ABH_TRAIN_GRAPH | ABH_TRAIN_FUSED)) == 0, ABH_ERR_INVALID_ARG,
              "unknown training flags");
  const bool graph = (opts->flags & ABH_TRAIN_GRAPH) != 0;
  const bool fused = (opts->flags & ABH_TRAIN_FUSED) != 0;
  FusedNet fn{};
  if (fused) {
    ABH_REQUIRE(m->layers() <= kFMaxL, ABH_ERR_UNSUPPORTED,
                "the fused training step takes at most 4 layers");
    for (int d : m->dims)
      ABH_REQUIRE(d <= kFMaxW, ABH_ERR_UNSUPPORTED, "the fused training step takes widths <= 128");
    ABH_REQUIRE(!mats || m->dims.back() <= kMaxK, ABH_ERR_UNSUPPORTED, "outputs above 64");
    fn.L = m->layers();
    for (int l = 0; l <= fn.L; ++l) fn.dims[l] = m->dims[size_t(l)];
    for (int l = 0; l < fn.L; ++l) {
      fn.w_off[l] = m->w_off[size_t(l)];
      fn.b_off[l] = m->b_off[size_t(l)];
    }
    fn.nparams = m->nparams;
  }
  abh::CaptureStream cs;
  if (graph) cs.begin(m->ctx);
  hipStream_t s = m->ctx->stream;
  const int nin = m->dims.front(), nout = m->dims.back();
  const int64_t tw = mats ? int64_t(nout) * nout : nout;  // floats per target row
  const int64_t B = std::min<int64_t>(opts->batch, n);
  const int64_t steps = n / B;  // full batches per epoch
  const int nchunks = int(abh::ceil_div(B, kFR));
  const size_t flds = fused ? fused_lds_floats(fn) * 4 : 0;
  DevBuf perm_d, xb, tb, scratch, losses, lr_d, step_d, part, lossp;
  if (fused) {
    part.ensure(size_t(nchunks) * size_t(m->nparams) * 4 + 64);
    lossp.ensure(size_t(nchunks) * 8 + 64);
    ABH_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_train_chunk),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(flds)));
  }
  perm_d.ensure(size_t(n) * 4 + 64);
  xb.ensure(size_t(B) * nin * 4 + 64);
  tb.ensure(size_t(B) * tw * 4 + 64);
  losses.ensure(size_t(std::max<int64_t>(steps, 1)) * 8 + 64);
  lr_d.ensure(size_t(std::max<int64_t>(steps, 1)) * 4 + 64);
  step_d.ensure(64);
  std::vector<int32_t> perm(static_cast<size_t>(n));
  std::vector<float> lrs(size_t(std::max<int64_t>(steps, 1)));
  std::mt19937_64 rng(opts->seed);
  float lr = opts->lr;
  std::vector<double> lh(size_t(std::max<int64_t>(steps, 1)));
  const int32_t *stp = step_d.as<int32_t>();
  const double loss_scale = mats ? 1.0 / double(B) : 1.0 / double(B * nout);
  auto fused_step = [&]() {
    k_train_chunk<<<unsigned(nchunks), kFT, flds, s>>>(
        fn, m->params.as<float>(), x, mats ? nullptr : t, mats, perm_d.as<int32_t>(), stp, int(B),
        part.as<float>(), lossp.as<double>());
    k_train_update<<<unsigned(abh::ceil_div(m->nparams, 256)), 256, 0, s>>>(
        m->params.as<float>(), m->am.as<float>(), m->av.as<float>(), m->nparams, part.as<float>(),
        nchunks, opts->beta1, opts->beta2, lr_d.as<float>(), stp, opts->eps, lossp.as<double>(),
        loss_scale, losses.as<double>());
    k_next_step<<<1, 64, 0, s>>>(step_d.as<int32_t>());
    abh::check_launch();
  };
  auto one_step = [&]() {
    if (fused) {
      fused_step();
      return;
    }
    k_gather_rows<<<unsigned(abh::ceil_div(B * nin, 256)), 256, 0, s>>>(
        x, perm_d.as<int32_t>(), stp, int(B), nin, xb.as<float>());
    k_gather_rows<<<unsigned(abh::ceil_div(B * tw, 256)), 256, 0, s>>>(
        mats ? mats : t, perm_d.as<int32_t>(), stp, int(B), int(tw), tb.as<float>());
    abh::check_launch();
    loss_grad(m, B, xb.as<float>(), mats ? nullptr : tb.as<float>(),
              mats ? tb.as<float>() : nullptr, m->grads.as<float>(), losses.as<double>(), scratch,
              stp);
    k_adam<<<unsigned(abh::ceil_div(m->nparams, 256)), 256, 0, s>>>(
        m->params.as<float>(), m->grads.as<float>(), m->am.as<float>(), m->av.as<float>(),
        m->nparams, opts->beta1, opts->beta2, lr_d.as<float>(), stp, opts->eps);
    k_next_step<<<1, 64, 0, s>>>(step_d.as<int32_t>());
    abh::check_launch();
  };
  hipGraph_t gr = nullptr;
  hipGraphExec_t ge = nullptr;
  struct GraphFree {
    hipGraph_t &g;
    hipGraphExec_t &e;
    ~GraphFree() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
    }
  } graph_free{gr, ge};
  bool sized = false;  // one eager step has sized every work buffer
  for (int ep = 0; ep < opts->epochs; ++ep) {
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), rng);
    for (int64_t st = 0; st < steps; ++st) {
      ++m->adam_t;
      const double b1 = opts->beta1, b2 = opts->beta2;
      lrs[size_t(st)] = float(lr * std::sqrt(1.0 - std::pow(b2, double(m->adam_t))) /
                              (1.0 - std::pow(b1, double(m->adam_t))));
    }
    ABH_HIP(hipMemcpyAsync(perm_d.p, perm.data(), size_t(n) * 4, hipMemcpyHostToDevice, s));
    if (steps > 0)
      ABH_HIP(hipMemcpyAsync(lr_d.p, lrs.data(), size_t(steps) * 4, hipMemcpyHostToDevice, s));
    k_set_step<<<1, 64, 0, s>>>(step_d.as<int32_t>(), 0);
    abh::check_launch();
    for (int64_t st = 0; st < steps; ++st) {
      if (!graph || !sized) {
        one_step();
        sized = true;
        continue;
      }
      if (!ge) {
        abh::CaptureGuard cg{s};
        ABH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        cg.open = true;
        one_step();
        cg.open = false;
        ABH_HIP(hipStreamEndCapture(s, &gr));
        ABH_HIP(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
      }
      ABH_HIP(hipGraphLaunch(ge, s));
    }
    // the perm and step-size uploads of the next epoch reuse the host
    // vectors: wait here
    ABH_HIP(hipMemcpyAsync(lh.data(), losses.p, size_t(std::max<int64_t>(steps, 1)) * 8,
                           hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    if (loss_per_epoch) {
      double sum = 0.0;
      for (int64_t st = 0; st < steps; ++st) sum += lh[size_t(st)];
      loss_per_epoch[ep] = steps > 0 ? float(sum / double(steps)) : 0.0f;
    }
    lr *= opts->lr_decay > 0.0f ? opts->lr_decay : 1.0f;
  }
}

bool bf16_ok(const abh_mlp *m) {
  if (m->layers() > kBfMaxLayers) return false;
  for (size_t l = 0; l < m->dims.size(); ++l) {
    const int d = m->dims[l];
    if (d > kBfMaxDim) return false;
    if (l == 0 ? d % 16 != 0 : d % 32 != 0) return false;
  }
  return true;
}

void forward_bf16(abh_mlp *m, int64_t n, const float *x, float *y) {
  hipStream_t s = m->ctx->stream;
  m->wbf.ensure(size_t(m->nparams) * 2 + 64);
  k_to_bf16<<<unsigned(abh::ceil_div(m->nparams, 256)), 256, 0, s>>>(
      m->params.as<float>(), m->nparams, m->wbf.as<uint16_t>());
  abh::check_launch();
  BfNet net{};
  net.nl = m->layers();
  for (int l = 0; l <= net.nl; ++l) net.dims[l] = m->dims[l];
  for (int l = 0; l < net.nl; ++l) {
    net.w_off[l] = m->w_off[l];
    net.b_off[l] = m->b_off[l];
  }
  if (n == 0) return;
  k_mlp_bf16<<<unsigned(abh::ceil_div(n, 128)), 256, 0, s>>>(net, m->wbf.as<uint16_t>(),
                                                             m->params.as<float>(), x, n, y);
  abh::check_launch();
}

void mlp_forward(abh_mlp *m, int64_t n, const float *x, float *y, bool bf16) {
  if (bf16) {
    ABH_REQUIRE(bf16_ok(m), ABH_ERR_UNSUPPORTED,
                "BF16 inference needs n_in % 16 == 0, widths % 32 == 0, widths <= 128");
    forward_bf16(m, n, x, y);
    return;
  }
  // FP32 in slabs that keep the workspace bounded
  const int64_t slab = 65536;
  for (int64_t r0 = 0; r0 < n; r0 += slab) {
    const int64_t rows = std::min(slab, n - r0);
    forward_f32(m, rows, x + r0 * m->dims[0], y + r0 * m->dims.back());
  }
}

}  // namespace

namespace abh {

void ainv_apply(const abh_ainv *a, const double *b, double *x) {
  if (a->n == 0) return;
  k_ainv_apply<<<unsigned(ceil_div(a->n, 256)), 256, 0, a->ctx->stream>>>(
      a->nbr.as<int32_t>(), a->w.as<double>(), a->n, a->k, b, x);
  check_launch();
}

int64_t ainv_size(const abh_ainv *a) { return a->n; }

}  // namespace abh

extern "C" {

abh_status abh_knn(abh_context *ctx, const abh_cells *cells, int k, int32_t *nbr, double *dist2) {
  if (!ctx) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] {
    check_cells(cells);
    ABH_REQUIRE(k >= 1 && k <= kMaxK, ABH_ERR_INVALID_ARG, "k must be in 1..64");
    ABH_REQUIRE(cells->n_alive == 0 || nbr, ABH_ERR_INVALID_ARG, "nbr is null");
    knn_run(ctx, cells, k, nbr, dist2);
  });
}

abh_status abh_mlp_create(abh_context *ctx, int n_in, int hidden, int n_hidden, int n_out,
                          uint64_t seed, abh_mlp **out) {
  if (!ctx || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_mlp *m = new (std::nothrow) abh_mlp();
  if (!m) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    ABH_REQUIRE(n_in >= 1 && hidden >= 1 && n_hidden >= 0 && n_out >= 1 && n_hidden <= 16,
                ABH_ERR_INVALID_ARG, "invalid network shape");
    m->ctx = ctx;
    m->dims.push_back(n_in);
    for (int l = 0; l < n_hidden; ++l) m->dims.push_back(hidden);
    m->dims.push_back(n_out);
    int64_t off = 0;
    for (int l = 0; l < m->layers(); ++l) {
      m->w_off.push_back(off);
      off += int64_t(m->dims[l]) * m->dims[l + 1];
      m->b_off.push_back(off);
      off += m->dims[l + 1];
    }
    m->nparams = off;
    // Glorot-uniform weights, zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> h(size_t(off), 0.0f);
    for (int l = 0; l < m->layers(); ++l) {
      const double a = std::sqrt(6.0 / double(m->dims[l] + m->dims[l + 1]));
      std::uniform_real_distribution<double> u(-a, a);
      for (int64_t e = 0; e < int64_t(m->dims[l]) * m->dims[l + 1]; ++e)
        h[size_t(m->w_off[l] + e)] = float(u(rng));
    }
    const size_t bytes = size_t(off) * 4 + 64;
    m->params.ensure(bytes);
    m->grads.ensure(bytes);
    m->am.ensure(bytes);
    m->av.ensure(bytes);
    hipStream_t s = ctx->stream;
    ABH_HIP(hipMemcpyAsync(m->params.p, h.data(), size_t(off) * 4, hipMemcpyHostToDevice, s));
    ABH_HIP(hipMemsetAsync(m->am.p, 0, bytes, s));
    ABH_HIP(hipMemsetAsync(m->av.p, 0, bytes, s));
    ABH_HIP(hipStreamSynchronize(s));
  });
  if (st != ABH_OK) {
    delete m;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = m;
  return ABH_OK;
}

abh_status abh_mlp_destroy(abh_mlp *m) {
  if (!m) return ABH_OK;
  abh_context *ctx = m->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete m;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_mlp_num_params(const abh_mlp *m, int64_t *n) {
  if (!m || !n) return ABH_ERR_INVALID_ARG;
  *n = m->nparams;
  return ABH_OK;
}

abh_status abh_mlp_get_params(const abh_mlp *m, float *host) {
  if (!m || !host) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] {
    ABH_HIP(hipMemcpyAsync(host, m->params.p, size_t(m->nparams) * 4, hipMemcpyDeviceToHost,
                           m->ctx->stream));
    ABH_HIP(hipStreamSynchronize(m->ctx->stream));
  });
}

abh_status abh_mlp_set_params(abh_mlp *m, const float *host) {
  if (!m || !host) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] {
    ABH_HIP(hipMemcpyAsync(m->params.p, host, size_t(m->nparams) * 4, hipMemcpyHostToDevice,
                           m->ctx->stream));
    ABH_HIP(hipStreamSynchronize(m->ctx->stream));
  });
}

abh_status abh_mlp_forward(abh_mlp *m, int64_t n, const float *x, float *y, int bf16) {
  if (!m) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] {
    ABH_REQUIRE(n >= 0, ABH_ERR_INVALID_ARG, "n must be >= 0");
    if (n == 0) return;
    ABH_REQUIRE(x && y, ABH_ERR_INVALID_ARG, "x or y is null");
    mlp_forward(m, n, x, y, bf16 != 0);
  });
}

abh_status abh_mlp_loss_grad(abh_mlp *m, int64_t n, const float *x, const float *t, double *loss,
                             float *grad) {
  if (!m) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] {
    ABH_REQUIRE(n >= 1 && n <= (int64_t(1) << 24), ABH_ERR_INVALID_ARG, "n out of range");
    ABH_REQUIRE(x && t, ABH_ERR_INVALID_ARG, "x or t is null");
    DevBuf scratch;
    m->loss.ensure(64);
    float *g = grad ? grad : m->grads.as<float>();
    loss_grad(m, n, x, t, nullptr, g, m->loss.as<double>(), scratch);
    double h = 0.0;
    ABH_HIP(hipMemcpyAsync(&h, m->loss.p, 8, hipMemcpyDeviceToHost, m->ctx->stream));
    ABH_HIP(hipStreamSynchronize(m->ctx->stream));
    if (loss) *loss = h;
  });
}

abh_status abh_mlp_train(abh_mlp *m, int64_t n, const float *x, const float *t,
                         const abh_mlp_train_options *opts, float *loss_per_epoch) {
  if (!m || !opts) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] {
    ABH_REQUIRE(n >= 1 && x && t, ABH_ERR_INVALID_ARG, "empty training set");
    train_impl(m, n, x, t, nullptr, opts, loss_per_epoch);
  });
}

abh_status abh_mlp_loss_grad_residual(abh_mlp *m, int64_t n, const float *x, const float *mats,
                                      double *loss, float *grad) {
  if (!m) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] {
    ABH_REQUIRE(n >= 1 && n <= (int64_t(1) << 24), ABH_ERR_INVALID_ARG, "n out of range");
    ABH_REQUIRE(x && mats, ABH_ERR_INVALID_ARG, "x or mats is null");
    ABH_REQUIRE(m->dims.back() <= kMaxK, ABH_ERR_UNSUPPORTED,
                "the patch-residual loss needs n_out <= 64");
    DevBuf scratch;
    m->loss.ensure(64);
    float *g = grad ? grad : m->grads.as<float>();
    loss_grad(m, n, x, nullptr, mats, g, m->loss.as<double>(), scratch);
    double h = 0.0;
    ABH_HIP(hipMemcpyAsync(&h, m->loss.p, 8, hipMemcpyDeviceToHost, m->ctx->stream));
    ABH_HIP(hipStreamSynchronize(m->ctx->stream));
    if (loss) *loss = h;
  });
}

abh_status abh_mlp_train_residual(abh_mlp *m, int64_t n, const float *x, const float *mats,
                                  const abh_mlp_train_options *opts, float *loss_per_epoch) {
  if (!m || !opts) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] {
    ABH_REQUIRE(n >= 1 && x && mats, ABH_ERR_INVALID_ARG, "empty training set");
    ABH_REQUIRE(m->dims.back() <= kMaxK, ABH_ERR_UNSUPPORTED,
                "the patch-residual loss needs n_out <= 64");
    train_impl(m, n, x, nullptr, mats, opts, loss_per_epoch);
  });
}

abh_status abh_ainv_create_lsq(abh_context *ctx, const abh_cells *cells, const abh_kernel *kernel,
                               double radius, double diag_shift, int k, abh_ainv **out) {
  if (!ctx || !kernel || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_ainv *a = new (std::nothrow) abh_ainv();
  if (!a) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    check_cells(cells);
    ABH_REQUIRE(k >= 1 && k <= kLsqMaxK, ABH_ERR_INVALID_ARG, "k must be in 1..32");
    ABH_REQUIRE(radius > 0.0, ABH_ERR_UNSUPPORTED, "least-squares rows need a finite radius");
    ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
                ABH_ERR_UNSUPPORTED, "unknown kernel type");
    a->ctx = ctx;
    a->n = cells->n_alive;
    a->k = k;
    a->dim = cells->dim;
    a->pos = cells->pos.as<double>();
    a->nbr.ensure(size_t(a->n) * k * 4 + 64);
    a->w.ensure(size_t(a->n) * k * 8 + 64);
    if (a->n == 0) return;
    knn_run(ctx, cells, k, a->nbr.as<int32_t>(), nullptr);
    KernelDesc kd{};
    kd.type = kernel->type;
    kd.param = kernel->param;
    kd.param2 = kernel->param * kernel->param;
    abh::set_inverses(kd);
    a->kern = kd;
    a->r2 = radius * radius;
    a->shift = diag_shift;
    a->has_op = true;
    DevBuf fail;
    fail.ensure(64);
    hipStream_t s = ctx->stream;
    ABH_HIP(hipMemsetAsync(fail.p, 0, 1, s));
    const GridParams g = cells->params();
    const double *pos = cells->pos.as<double>();
    const uint32_t *b = cells->begin.as<uint32_t>(), *e = cells->end.as<uint32_t>();
    const double r2 = radius * radius;
    const unsigned grid = unsigned(a->n);
#define ABH_LSQ(DD)                                                                        \
  k_lsq<DD><<<grid, 256, 0, s>>>(pos, b, e, g, a->n, k, a->nbr.as<int32_t>(), kd, r2, diag_shift, \
                                 a->w.as<double>(), fail.as<uint8_t>())
    if (cells->dim == 1) ABH_LSQ(1);
    else if (cells->dim == 2) ABH_LSQ(2);
    else ABH_LSQ(3);
#undef ABH_LSQ
    abh::check_launch();
    uint8_t f = 0;
    ABH_HIP(hipMemcpyAsync(&f, fail.p, 1, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    ABH_REQUIRE(!f, ABH_ERR_NUMERICAL, "least-squares normal equations not positive definite");
  });
  if (st != ABH_OK) {
    delete a;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = a;
  return ABH_OK;
}

abh_status abh_ainv_create_patch(abh_context *ctx, const abh_cells *cells,
                                 const abh_kernel *kernel, double radius, double diag_shift,
                                 int k, abh_ainv **out) {
  if (!ctx || !kernel || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_ainv *a = new (std::nothrow) abh_ainv();
  if (!a) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    check_cells(cells);
    ABH_REQUIRE(k >= 1 && k <= kMaxK, ABH_ERR_INVALID_ARG, "k must be in 1..64");
    ABH_REQUIRE(radius > 0.0, ABH_ERR_UNSUPPORTED, "patch rows need a finite radius");
    ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
                ABH_ERR_UNSUPPORTED, "unknown kernel type");
    a->ctx = ctx;
    a->n = cells->n_alive;
    a->k = k;
    a->dim = cells->dim;
    a->pos = cells->pos.as<double>();
    a->kern.type = kernel->type;
    a->kern.param = kernel->param;
    a->kern.param2 = kernel->param * kernel->param;
    abh::set_inverses(a->kern);
    a->r2 = radius * radius;
    a->shift = diag_shift;
    a->has_op = true;
    a->nbr.ensure(size_t(a->n) * k * 4 + 64);
    a->w.ensure(size_t(a->n) * k * 8 + 64);
    if (a->n == 0) return;
    hipStream_t s = ctx->stream;
    knn_run(ctx, cells, k, a->nbr.as<int32_t>(), nullptr);
    DevBuf fail;
    fail.ensure(64);
    ABH_HIP(hipMemsetAsync(fail.p, 0, 1, s));
    const unsigned grid = unsigned(a->n);
#define ABH_PATCH(DD)                                                                     \
  k_patch<DD><<<grid, 64, 0, s>>>(a->pos, a->n, k, a->nbr.as<int32_t>(), a->kern, a->r2, \
                                  a->shift, a->w.as<double>(), fail.as<uint8_t>())
    if (a->dim == 1) ABH_PATCH(1);
    else if (a->dim == 2) ABH_PATCH(2);
    else ABH_PATCH(3);
#undef ABH_PATCH
    abh::check_launch();
    uint8_t f = 0;
    ABH_HIP(hipMemcpyAsync(&f, fail.p, 1, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    ABH_REQUIRE(!f, ABH_ERR_NUMERICAL, "a patch matrix is not positive definite");
  });
  if (st != ABH_OK) {
    delete a;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = a;
  return ABH_OK;
}

abh_status abh_ainv_patch_matrices(const abh_ainv *a, double scale, float *mats) {
  if (!a) return ABH_ERR_INVALID_ARG;
  return abh::guard(a->ctx, [&] {
    ABH_REQUIRE(a->has_op, ABH_ERR_INVALID_ARG,
                "patch matrices need rows built from the operator (lsq or patch)");
    if (a->n == 0) return;
    ABH_REQUIRE(mats, ABH_ERR_INVALID_ARG, "mats is null");
    const int64_t total = a->n * int64_t(a->k) * a->k;
    const unsigned g = unsigned(abh::ceil_div(total, 256));
    hipStream_t s = a->ctx->stream;
#define ABH_PM(DD)                                                                           \
  k_patch_mats<DD><<<g, 256, 0, s>>>(a->pos, a->n, a->k, a->nbr.as<int32_t>(), a->kern, a->r2, \
                                     a->shift, scale, mats)
    if (a->dim == 1) ABH_PM(1);
    else if (a->dim == 2) ABH_PM(2);
    else ABH_PM(3);
#undef ABH_PM
    abh::check_launch();
  });
}

abh_status abh_ainv_create_learned(abh_context *ctx, const abh_cells *cells, abh_mlp *net, int k,
                                   double feature_scale, double out_scale, int bf16,
                                   abh_ainv **out) {
  if (!ctx || !net || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_ainv *a = new (std::nothrow) abh_ainv();
  if (!a) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    check_cells(cells);
    ABH_REQUIRE(k >= 1 && k <= kMaxK, ABH_ERR_INVALID_ARG, "k must be in 1..64");
    ABH_REQUIRE(net->ctx == ctx, ABH_ERR_INVALID_ARG, "network belongs to another context");
    ABH_REQUIRE(net->dims.front() == k * cells->dim && net->dims.back() == k,
                ABH_ERR_INVALID_ARG, "network must map k*dim features to k weights");
    ABH_REQUIRE(feature_scale > 0.0, ABH_ERR_INVALID_ARG, "feature_scale must be > 0");
    a->ctx = ctx;
    a->n = cells->n_alive;
    a->k = k;
    a->dim = cells->dim;
    a->pos = cells->pos.as<double>();
    a->nbr.ensure(size_t(a->n) * k * 4 + 64);
    a->w.ensure(size_t(a->n) * k * 8 + 64);
    if (a->n == 0) return;
    hipStream_t s = ctx->stream;
    knn_run(ctx, cells, k, a->nbr.as<int32_t>(), nullptr);
    DevBuf feat, y;
    feat.ensure(size_t(a->n) * k * a->dim * 4 + 64);
    y.ensure(size_t(a->n) * k * 4 + 64);
    k_features<<<unsigned(abh::ceil_div(a->n * k, 256)), 256, 0, s>>>(
        a->pos, a->n, k, a->dim, a->nbr.as<int32_t>(), 1.0 / feature_scale, feat.as<float>());
    abh::check_launch();
    mlp_forward(net, a->n, feat.as<float>(), y.as<float>(), bf16 != 0);
    k_scale_to_f64<<<unsigned(abh::ceil_div(a->n * k, 256)), 256, 0, s>>>(
        y.as<float>(), a->n * k, out_scale, a->w.as<double>());
    abh::check_launch();
    ABH_HIP(hipStreamSynchronize(s));
  });
  if (st != ABH_OK) {
    delete a;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = a;
  return ABH_OK;
}

abh_status abh_ainv_destroy(abh_ainv *a) {
  if (!a) return ABH_OK;
  abh_context *ctx = a->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete a;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_ainv_get(const abh_ainv *a, int32_t *nbr, double *w) {
  if (!a) return ABH_ERR_INVALID_ARG;
  return abh::guard(a->ctx, [&] {
    hipStream_t s = a->ctx->stream;
    if (nbr && a->n)
      ABH_HIP(hipMemcpyAsync(nbr, a->nbr.p, size_t(a->n) * a->k * 4, hipMemcpyDeviceToDevice, s));
    if (w && a->n)
      ABH_HIP(hipMemcpyAsync(w, a->w.p, size_t(a->n) * a->k * 8, hipMemcpyDeviceToDevice, s));
  });
}

abh_status abh_ainv_features(const abh_ainv *a, double feature_scale, float *features) {
  if (!a) return ABH_ERR_INVALID_ARG;
  return abh::guard(a->ctx, [&] {
    ABH_REQUIRE(feature_scale > 0.0, ABH_ERR_INVALID_ARG, "feature_scale must be > 0");
    if (a->n == 0) return;
    ABH_REQUIRE(features, ABH_ERR_INVALID_ARG, "features is null");
    k_features<<<unsigned(abh::ceil_div(a->n * a->k, 256)), 256, 0, a->ctx->stream>>>(
        a->pos, a->n, a->k, a->dim, a->nbr.as<int32_t>(), 1.0 / feature_scale, features);
    abh::check_launch();
  });
}

abh_status abh_ainv_apply(const abh_ainv *a, const double *b, double *x) {
  if (!a) return ABH_ERR_INVALID_ARG;
  return abh::guard(a->ctx, [&] {
    ABH_REQUIRE(a->n == 0 || (b && x && b != x), ABH_ERR_INVALID_ARG,
                "b and x must be distinct device vectors");
    abh::ainv_apply(a, b, x);
  });
}

}  // extern "C"
