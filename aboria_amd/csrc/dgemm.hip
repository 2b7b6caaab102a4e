// Strided-batched FP64 GEMM on the matrix cores (v_mfma_f64_16x16x4f64),
// column-major like BLAS:  C[g] = alpha op(A[g]) op(B[g]) + beta C[g].
// Used by the Schwarz local factorisations (panel / trailing updates and the
// blocked triangular solves), whose operands are 64-wide panels of matrices
// up to a few thousand rows.
//
// One workgroup (4 wavefronts) computes a 64 x 64 tile of C; each wavefront a
// 32 x 32 quarter as 2 x 2 blocks of 16 x 16.  The k dimension is staged in
// LDS 16 at a time (op(A) as [k][m], op(B) as [k][n]), the next stage's
// global loads issued before the current stage's products.  The products are
// formed transposed (C^T = op(B)^T op(A)^T): lane l supplies op(B)[k][n] as
// the A fragment and op(A)[m][k] as the B fragment (A fragment: row l % 16,
// k l / 16; B fragment: k l / 16, column l % 16), so accumulator register v
// of lane l holds C[m = l % 16][n = l / 16 + 4 v] and the 16 lanes of a
// group store 16 consecutive rows of a column (128 contiguous bytes).
//
// lower_only skips the tiles that lie strictly above the diagonal (a
// symmetric trailing update whose upper triangle is never read).
#include "common.hpp"
#include "solvers.hpp"

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kT = 64;   // tile side
constexpr int kK = 16;   // k per LDS stage
constexpr int kLd = kT + 1;

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_dgemm(int M, int N, int K, double alpha,
                                               const double *__restrict__ A, int lda,
                                               int64_t sA, const double *__restrict__ B, int ldb,
                                               int64_t sB, double beta, double *__restrict__ C,
                                               int ldc, int64_t sC, int tiles_m, int lower_only) {
  __shared__ double As[kK][kLd];  // op(A)[m][k] stored [k][m]
  __shared__ double Bs[kK][kLd];  // op(B)[k][n] stored [k][n]
  const int g = blockIdx.y;
  const int tm = int(blockIdx.x) % tiles_m, tn = int(blockIdx.x) / tiles_m;
  const int m0 = tm * kT, n0 = tn * kT;
  if (lower_only && m0 + kT - 1 < n0) return;  // uniform over the workgroup
  A += int64_t(g) * sA;
  B += int64_t(g) * sB;
  C += int64_t(g) * sC;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

  // global -> register staging of one k stage: 4 elements of each operand
  // per thread, coalesced along the operand's contiguous dimension
  double ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int m, k;
      if (!TA) { m = t & 63; k = (t >> 6) + 4 * i; }   // A[m + k lda], m fastest
      else { k = t & 15; m = (t >> 4) + 16 * i; }       // A[k + m lda], k fastest
      const int gm = m0 + m, gk = k0 + k;
      double v = 0.0;
      if (gm < M && gk < K) v = TA ? A[int64_t(gm) * lda + gk] : A[int64_t(gk) * lda + gm];
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int n, k;
      if (!TB) { k = t & 15; n = (t >> 4) + 16 * i; }   // B[k + n ldb], k fastest
      else { n = t & 63; k = (t >> 6) + 4 * i; }        // B[n + k ldb], n fastest
      const int gn = n0 + n, gk = k0 + k;
      double v = 0.0;
      if (gn < N && gk < K) v = TB ? B[int64_t(gk) * ldb + gn] : B[int64_t(gn) * ldb + gk];
      rb[i] = v;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (!TA) As[(t >> 6) + 4 * i][t & 63] = ra[i];
      else As[t & 15][(t >> 4) + 16 * i] = ra[i];
      if (!TB) Bs[t & 15][(t >> 4) + 16 * i] = rb[i];
      else Bs[(t >> 6) + 4 * i][t & 63] = rb[i];
    }
  };

  f64x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  const int kl = lane >> 4, rl = lane & 15;
  load(0);
  for (int k0 = 0; k0 < K; k0 += kK) {
    __syncthreads();
    store();
    __syncthreads();
    if (k0 + kK < K) load(k0 + kK);
#pragma unroll
    for (int ks = 0; ks < kK / 4; ++ks) {
      const int k = ks * 4 + kl;
      double fa[2], fb[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) fa[j] = Bs[k][wn + j * 16 + rl];  // op(B)^T rows = n
#pragma unroll
      for (int i = 0; i < 2; ++i) fb[i] = As[k][wm + i * 16 + rl];  // op(A)^T cols = m
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[j], fb[i], acc[i][j], 0, 0, 0);
    }
  }
  // acc[i][j][v] = C[m = m0 + wm + 16 i + rl][n = n0 + wn + 16 j + kl + 4 v]
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int gm = m0 + wm + 16 * i + rl;
    if (gm >= M) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gn = n0 + wn + 16 * j + kl + 4 * v;
        if (gn >= N) continue;
        double *c = C + int64_t(gn) * ldc + gm;
        const double r = alpha * acc[i][j][v];
        *c = beta == 0.0 ? r : r + beta * *c;
      }
  }
}

}  // namespace

namespace abh {

void dgemm_strided_batched(hipStream_t st, bool ta, bool tb, int m, int n, int k, double alpha,
                           const double *A, int lda, int64_t sA, const double *B, int ldb,
                           int64_t sB, double beta, double *C, int ldc, int64_t sC, int batch,
                           bool lower_only) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  ABH_REQUIRE(k >= 0 && lda >= 1 && ldb >= 1 && ldc >= m, ABH_ERR_INVALID_ARG,
              "dgemm: bad leading dimension");
  ABH_REQUIRE(batch <= 65535, ABH_ERR_INVALID_ARG, "dgemm: batch too large");
  const int tm = int(ceil_div(m, kT)), tn = int(ceil_div(n, kT));
  ABH_REQUIRE(int64_t(tm) * tn < (int64_t(1) << 31), ABH_ERR_INVALID_ARG, "dgemm: too many tiles");
  if (k == 0) {
    // C = beta C: the product is empty
    ABH_REQUIRE(beta == 1.0, ABH_ERR_INVALID_ARG, "dgemm: k == 0 needs beta == 1");
    return;
  }
  dim3 grid(unsigned(tm * tn), unsigned(batch));
  const int lo = lower_only ? 1 : 0;
  if (!ta && !tb)
    k_dgemm<false, false><<<grid, 256, 0, st>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                                                 ldc, sC, tm, lo);
  else if (!ta && tb)
    k_dgemm<false, true><<<grid, 256, 0, st>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                                                ldc, sC, tm, lo);
  else if (ta && !tb)
    k_dgemm<true, false><<<grid, 256, 0, st>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                                                ldc, sC, tm, lo);
  else
    k_dgemm<true, true><<<grid, 256, 0, st>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C,
                                               ldc, sC, tm, lo);
  check_launch();
}

}  // namespace abh
