// Batched dense kernel operators: many independent particle sets (problems
// of uneven size) evaluated in one launch, and their GMRES solve.
//
// Work decomposition: a tile is 256 consecutive rows of one problem; one
// 256-thread workgroup per tile streams the problem's sources through LDS
// in chunks of 256 (positions + x), each thread accumulating one row
// (the P2P tile of the dense operator).  Tiles are ordered largest problem
// first so the long rows of the big problems start early.
#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "batch.hpp"
#include "kernel_fn.hpp"
#include "solvers.hpp"
#include "solvers.hpp"

using abh::DevBuf;
using abh::KernelDesc;

namespace {

constexpr int kRows = 256;

// Tunable variants: RPT rows per thread (tile = 256 RPT rows of one
// problem), so every source staged in LDS serves RPT rows.
constexpr int kVariants = 3;
constexpr int kRptOf[kVariants] = {1, 2, 4};

// MQ: the multiquadric sqrt(|dx|^2 + c^2) with fused multiply-adds (the
// batched RBF systems' kernel), instead of the generic profile.
template <int D, bool FP32, int RPT, bool MQ>
__global__ __launch_bounds__(kRows) void k_dense(const int2 *__restrict__ tiles, int ntiles,
                                                 const int64_t *__restrict__ off,
                                                 const double *__restrict__ pos,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y,
                                                 const uint8_t *__restrict__ active,
                                                 KernelDesc k, double alpha, int overwrite) {
  __shared__ double sp[kRows][D];
  __shared__ double sx[kRows];
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const int p = tiles[t].x;
  if (active && !active[p]) return;
  const int64_t base = off[p];
  const int64_t n = off[p + 1] - base;
  double xi[RPT][D];
  bool act[RPT];
  double acc[RPT];
  float a32[RPT], c32[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t i = int64_t(tiles[t].y) + threadIdx.x + q * kRows;
    act[q] = i < n;
#pragma unroll
    for (int d = 0; d < D; ++d) xi[q][d] = act[q] ? pos[(base + i) * D + d] : 0.0;
    acc[q] = 0.0;
    a32[q] = 0.0f;
    c32[q] = 0.0f;
  }
  bool any = false;
#pragma unroll
  for (int q = 0; q < RPT; ++q) any |= act[q];
  for (int64_t c0 = 0; c0 < n; c0 += kRows) {
    const int m = int(n - c0 < kRows ? n - c0 : kRows);
    __syncthreads();
    if (int(threadIdx.x) < m) {
      const int64_t j = base + c0 + threadIdx.x;
#pragma unroll
      for (int d = 0; d < D; ++d) sp[threadIdx.x][d] = pos[j * D + d];
      sx[threadIdx.x] = x[j];
    }
    __syncthreads();
    if (!any) continue;
    for (int j = 0; j < m; ++j) {
      double pj[D];
#pragma unroll
      for (int d = 0; d < D; ++d) pj[d] = sp[j][d];
      const double xj = sx[j];
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (MQ && FP32) {
          float s = float(k.param2);
#pragma unroll
          for (int d = D - 1; d >= 0; --d) {
            const float dx = float(pj[d] - xi[q][d]);
            s = fmaf(dx, dx, s);
          }
          abh::kb_add(a32[q], c32[q], __builtin_amdgcn_sqrtf(s) * float(xj));
        } else if (MQ) {
          double s = k.param2;
#pragma unroll
          for (int d = D - 1; d >= 0; --d) {
            const double dx = pj[d] - xi[q][d];
            s = fma(dx, dx, s);
          }
          acc[q] = fma(abh::sqrt_pos(s), xj, acc[q]);  // s >= c^2 >= 1e-280
        } else if (FP32) {
          float s = 0.0f;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const float dx = float(pj[d] - xi[q][d]);
            s = s + dx * dx;
          }
          abh::kb_add(a32[q], c32[q], abh::phi32(k, s) * float(xj));
        } else {
          double s = 0.0;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const double dx = pj[d] - xi[q][d];
            s = s + dx * dx;
          }
          acc[q] += abh::phi64(k, s) * xj;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    if (!act[q]) continue;
    const int64_t i = int64_t(tiles[t].y) + threadIdx.x + q * kRows;
    const double v = FP32 ? double(a32[q]) + double(c32[q]) : acc[q];
    double *yp = &y[base + i];
    *yp = (overwrite ? 0.0 : *yp) + alpha * v;
  }
}

// r = b - K x for every row of the batch in double-double arithmetic: each
// term K_ij x_j split exactly into its rounded product and the product's
// error (fma), the terms summed with Knuth's two-sum into a (hi, lo) pair,
// and b - (hi + lo) formed the same way before one final rounding.  K_ij is
// the FP64 value the operator's own FP64 apply uses (the multiquadric form
// with fused multiply-adds, or the generic profile), so r is the residual of
// the FP64 matrix, free of the summation's rounding.  One row per thread,
// sources staged in LDS as in k_dense.
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}

template <int D, bool MQ>
__global__ __launch_bounds__(kRows) void k_dense_resid_dd(const int2 *__restrict__ tiles,
                                                          int ntiles,
                                                          const int64_t *__restrict__ off,
                                                          const double *__restrict__ pos,
                                                          const double *__restrict__ x,
                                                          const double *__restrict__ b,
                                                          double *__restrict__ r,
                                                          const uint8_t *__restrict__ active,
                                                          KernelDesc k) {
  __shared__ double sp[kRows][D];
  __shared__ double sx[kRows];
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const int p = tiles[t].x;
  if (active && !active[p]) return;
  const int64_t base = off[p];
  const int64_t n = off[p + 1] - base;
  const int64_t i = int64_t(tiles[t].y) + threadIdx.x;
  const bool act = i < n;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = act ? pos[(base + i) * D + d] : 0.0;
  double hi = 0.0, lo = 0.0;
  for (int64_t c0 = 0; c0 < n; c0 += kRows) {
    const int m = int(n - c0 < kRows ? n - c0 : kRows);
    __syncthreads();
    if (int(threadIdx.x) < m) {
      const int64_t j = base + c0 + threadIdx.x;
#pragma unroll
      for (int d = 0; d < D; ++d) sp[threadIdx.x][d] = pos[j * D + d];
      sx[threadIdx.x] = x[j];
    }
    __syncthreads();
    if (!act) continue;
    for (int j = 0; j < m; ++j) {
      double kij;
      if (MQ) {
        double s = k.param2;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
          const double dx = sp[j][d] - xi[d];
          s = fma(dx, dx, s);
        }
        kij = abh::sqrt_pos(s);
      } else {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const double dx = sp[j][d] - xi[d];
          s = s + dx * dx;
        }
        kij = abh::phi64(k, s);
      }
      const double xj = sx[j];
      const double pr = kij * xj;
      const double pe = fma(kij, xj, -pr);
      double sum, err;
      two_sum(hi, pr, sum, err);
      hi = sum;
      lo += err + pe;
    }
  }
  if (act) {
    double s, e;
    two_sum(b[base + i], -hi, s, e);
    r[base + i] = s + (e - lo);
  }
}

// Rectangular operator y += K x over separate row / column point sets,
// up to kRectRhs right-hand sides per pass (column-major x, y).  Block
// (bx, by) takes rows bx * kRows .. and columns by * cspan ..; with several
// column ranges each writes its partial sums to part[by][r][row], which
// k_rect_reduce adds to y in range order (deterministic).
constexpr int kRectRhs = 16;

template <int D, bool FP32>
__global__ __launch_bounds__(kRows) void k_dense_rect(int64_t nrows, const double *__restrict__ rp,
                                                      int64_t ncols, const double *__restrict__ cp,
                                                      KernelDesc k, int nrhs,
                                                      const double *__restrict__ x, int64_t ldx,
                                                      double *__restrict__ y, int64_t ldy,
                                                      int64_t cspan, double *__restrict__ part) {
  __shared__ double sp[kRows][D];
  __shared__ double sx[kRectRhs][kRows];
  const int64_t i = int64_t(blockIdx.x) * kRows + threadIdx.x;
  const bool act = i < nrows;
  double xi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = act ? rp[i * D + d] : 0.0;
  double acc[kRectRhs];
  float a32[kRectRhs], c32[kRectRhs];
#pragma unroll
  for (int r = 0; r < kRectRhs; ++r) {
    acc[r] = 0.0;
    a32[r] = 0.0f;
    c32[r] = 0.0f;
  }
  const int64_t cb = int64_t(blockIdx.y) * cspan;
  const int64_t ce = cb + cspan < ncols ? cb + cspan : ncols;
  for (int64_t c0 = cb; c0 < ce; c0 += kRows) {
    const int m = int(ce - c0 < kRows ? ce - c0 : kRows);
    __syncthreads();
    if (int(threadIdx.x) < m) {
      const int64_t j = c0 + threadIdx.x;
#pragma unroll
      for (int d = 0; d < D; ++d) sp[threadIdx.x][d] = cp[j * D + d];
      for (int r = 0; r < nrhs; ++r) sx[r][threadIdx.x] = x[int64_t(r) * ldx + j];
    }
    __syncthreads();
    if (!act) continue;
    for (int j = 0; j < m; ++j) {
      if (FP32) {
        float s = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const float dx = float(sp[j][d] - xi[d]);
          s = s + dx * dx;
        }
        const float f = abh::phi32(k, s);
#pragma unroll
        for (int r = 0; r < kRectRhs; ++r)
          if (r < nrhs) abh::kb_add(a32[r], c32[r], f * float(sx[r][j]));
      } else {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const double dx = sp[j][d] - xi[d];
          s = s + dx * dx;
        }
        const double f = abh::phi64(k, s);
#pragma unroll
        for (int r = 0; r < kRectRhs; ++r)
          if (r < nrhs) acc[r] += f * sx[r][j];
      }
    }
  }
  if (act)
#pragma unroll
    for (int r = 0; r < kRectRhs; ++r)
      if (r < nrhs) {
        const double v = FP32 ? double(a32[r]) + double(c32[r]) : acc[r];
        if (part) part[(int64_t(blockIdx.y) * nrhs + r) * nrows + i] = v;
        else y[int64_t(r) * ldy + i] += v;
      }
}

// y[r][i] += sum over the column ranges s (in order) of part[s][r][i]
__global__ __launch_bounds__(256) void k_rect_reduce(int64_t nrows, int nrhs, int ns,
                                                     const double *__restrict__ part,
                                                     double *__restrict__ y, int64_t ldy) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= nrows * nrhs) return;
  const int r = int(t / nrows);
  const int64_t i = t % nrows;
  double v = 0.0;
  for (int s = 0; s < ns; ++s) v += part[(int64_t(s) * nrhs + r) * nrows + i];
  y[int64_t(r) * ldy + i] += v;
}

}  // namespace

// Static scratch of abh_dense_apply_noalloc (the column split's partial
// sums), one copy per device for the life of the process.  Host code uses
// the device address that hipGetSymbolAddress returns for it (kept in the
// context), never the array's name.
constexpr size_t kNoAllocScratchDoubles = size_t(4) << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace abh {

double *dense_scratch_address() {
  void *p = nullptr;
  ABH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  ABH_REQUIRE(p != nullptr, ABH_ERR_HIP, "no device address for the dense apply scratch");
  return static_cast<double *>(p);
}

}  // namespace abh

struct abh_dense_batch {
  abh_context *ctx = nullptr;
  int dim = 2;
  abh::BatchLayout L;
  KernelDesc k{};
  const double *pos = nullptr;
  DevBuf tiles[kVariants], off, active;
  int ntiles[kVariants] = {0, 0, 0};
  int choice[2] = {-1, -1};        // tuned variant per precision (FP64, FP32c)
  float tune_ms[2][kVariants] = {};

  template <int RPT>
  void launch(int v, const double *x, double *y, bool fp32, double alpha, bool overwrite,
              const uint8_t *act) {
    hipStream_t s = ctx->stream;
    const int2 *tl = tiles[v].as<int2>();
    const int64_t *o = off.as<int64_t>();
    const int nt = ntiles[v];
    const int ow = overwrite ? 1 : 0;
    const bool mq =
        k.type == ABH_KERNEL_MULTIQUADRIC && !k.row_s && !k.col_s && k.param2 >= 1e-280;
#define ABH_DENSE_LAUNCH2(DD, MQ)                                                               \
  if (fp32)                                                                                     \
    k_dense<DD, true, RPT, MQ><<<nt, kRows, 0, s>>>(tl, nt, o, pos, x, y, act, k, alpha, ow);   \
  else                                                                                          \
    k_dense<DD, false, RPT, MQ><<<nt, kRows, 0, s>>>(tl, nt, o, pos, x, y, act, k, alpha, ow);
#define ABH_DENSE_LAUNCH(DD)                                                                    \
  if (mq) { ABH_DENSE_LAUNCH2(DD, true) }                                                       \
  else { ABH_DENSE_LAUNCH2(DD, false) }
    if (dim == 1) { ABH_DENSE_LAUNCH(1) }
    else if (dim == 2) { ABH_DENSE_LAUNCH(2) }
    else { ABH_DENSE_LAUNCH(3) }
#undef ABH_DENSE_LAUNCH
#undef ABH_DENSE_LAUNCH2
    abh::check_launch();
  }

  // r = b - K x in double-double (k_dense_resid_dd) over the one-row tiles
  void residual_dd(const double *x, const double *b, double *r, const uint8_t *act) {
    if (ntiles[0] == 0) return;
    hipStream_t s = ctx->stream;
    const int2 *tl = tiles[0].as<int2>();
    const int64_t *o = off.as<int64_t>();
    const int nt = ntiles[0];
    const bool mq =
        k.type == ABH_KERNEL_MULTIQUADRIC && !k.row_s && !k.col_s && k.param2 >= 1e-280;
#define ABH_RES_DD(DD)                                                                      \
  if (mq) k_dense_resid_dd<DD, true><<<nt, kRows, 0, s>>>(tl, nt, o, pos, x, b, r, act, k);  \
  else k_dense_resid_dd<DD, false><<<nt, kRows, 0, s>>>(tl, nt, o, pos, x, b, r, act, k);
    if (dim == 1) { ABH_RES_DD(1) }
    else if (dim == 2) { ABH_RES_DD(2) }
    else { ABH_RES_DD(3) }
#undef ABH_RES_DD
    abh::check_launch();
  }

  void run(int v, const double *x, double *y, bool fp32, double alpha, bool overwrite,
           const uint8_t *act) {
    if (ntiles[v] == 0) return;
    if (v == 0) launch<1>(v, x, y, fp32, alpha, overwrite, act);
    else if (v == 1) launch<2>(v, x, y, fp32, alpha, overwrite, act);
    else launch<4>(v, x, y, fp32, alpha, overwrite, act);
  }

  // Autotuner: times every variant once (after a warm-up) on the batch
  // itself and keeps the fastest for this precision.
  int tune(bool fp32) {
    const int pi = fp32 ? 1 : 0;
    if (choice[pi] >= 0) return choice[pi];
    const int64_t n = L.total();
    DevBuf xs, ys;
    xs.ensure(size_t(n) * 8 + 64);
    ys.ensure(size_t(n) * 8 + 64);
    hipStream_t s = ctx->stream;
    ABH_HIP(hipMemsetAsync(xs.p, 0, size_t(n) * 8, s));
    hipEvent_t e0, e1;
    ABH_HIP(hipEventCreate(&e0));
    ABH_HIP(hipEventCreate(&e1));
    int best = 0;
    for (int v = 0; v < kVariants; ++v) {
      run(v, xs.as<double>(), ys.as<double>(), fp32, 1.0, true, nullptr);
      ABH_HIP(hipEventRecord(e0, s));
      run(v, xs.as<double>(), ys.as<double>(), fp32, 1.0, true, nullptr);
      ABH_HIP(hipEventRecord(e1, s));
      ABH_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      ABH_HIP(hipEventElapsedTime(&ms, e0, e1));
      tune_ms[pi][v] = ms;
      if (ms < tune_ms[pi][best]) best = v;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    choice[pi] = best;
    return best;
  }

  void apply(const double *x, double *y, bool fp32, double alpha, bool overwrite,
             const uint8_t *act) {
    run(tune(fp32), x, y, fp32, alpha, overwrite, act);
  }
};

namespace {

// y += K x for the rectangular operator.  part_ext == nullptr: the column
// split's partial sums live in the context's grow-only buffer; otherwise in
// part_ext (part_cap doubles), and a split that needs more is refused.
void dense_apply_impl(abh_context *ctx, int dim, int64_t nrows, const double *row_pos,
                      int64_t ncols, const double *col_pos, const abh_kernel *kernel, int nrhs,
                      const double *x, int64_t ldx, double *y, int64_t ldy,
                      abh_precision precision, double *part_ext, size_t part_cap) {
  ABH_REQUIRE(dim >= 1 && dim <= 3, ABH_ERR_INVALID_ARG, "dim must be 1..3");
  ABH_REQUIRE(nrows >= 0 && ncols >= 0 && nrhs >= 0, ABH_ERR_INVALID_ARG, "negative size");
  ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
              ABH_ERR_UNSUPPORTED, "unknown kernel type");
  ABH_REQUIRE(precision == ABH_FP64 || precision == ABH_FP32_COMPENSATED, ABH_ERR_UNSUPPORTED,
              "dense precision must be ABH_FP64 or ABH_FP32_COMPENSATED");
  if (nrows == 0 || nrhs == 0) return;
  ABH_REQUIRE(row_pos && y && (ncols == 0 || (col_pos && x)), ABH_ERR_INVALID_ARG,
              "null pointer");
  ABH_REQUIRE(ldx >= ncols && ldy >= nrows, ABH_ERR_INVALID_ARG,
              "leading dimension smaller than the operator");
  KernelDesc k{};
  k.type = kernel->type;
  k.param = kernel->param;
  k.param2 = kernel->param * kernel->param;
  abh::set_inverses(k);
  const bool f32 = precision == ABH_FP32_COMPENSATED;
  hipStream_t s = ctx->stream;
  const int64_t grows = abh::ceil_div(nrows, kRows);
  // few row blocks (a tall operator has plenty): split the columns into
  // ranges of at least 2,048 so that about 8 blocks per CU run
  int64_t ns = 1;
  const int64_t want = int64_t(ctx->n_cu) * 8;
  if (grows < want && ncols > 0)
    ns = std::max<int64_t>(1, std::min(abh::ceil_div(want, grows), ncols / 2048));
  const int64_t cspan =
      ncols > 0 ? abh::ceil_div(abh::ceil_div(ncols, ns), int64_t(kRows)) * kRows : 1;
  ns = ncols > 0 ? abh::ceil_div(ncols, cspan) : 1;
  double *part = nullptr;
  if (ns > 1) {
    // part[s][r][i] for the ranges s, the right-hand sides r of one pass and
    // the rows i; every entry is written by k_dense_rect before k_rect_reduce
    // reads it, so the buffer needs no clearing
    const size_t need = size_t(ns) * size_t(std::min(nrhs, kRectRhs)) * size_t(nrows);
    if (part_ext) {
      ABH_REQUIRE(need <= part_cap, ABH_ERR_CAPACITY,
                  "column split needs " + std::to_string(need * 8) +
                      " bytes of partial sums, above the " + std::to_string(part_cap * 8) +
                      "-byte static scratch of the allocation-free entry point (use "
                      "abh_dense_apply, or fewer right-hand sides per call)");
      part = part_ext;
    } else {
      ctx->dense_part.ensure(need * sizeof(double));
      part = ctx->dense_part.as<double>();
    }
  }
  const dim3 grid{unsigned(grows), unsigned(ns), 1u};
  for (int r0 = 0; r0 < nrhs; r0 += kRectRhs) {
    const int nr = std::min(kRectRhs, nrhs - r0);
    const double *xr = x + int64_t(r0) * ldx;
    double *yr = y + int64_t(r0) * ldy;
#define ABH_RECT(DD)                                                                            \
  if (f32)                                                                                      \
    k_dense_rect<DD, true><<<grid, kRows, 0, s>>>(nrows, row_pos, ncols, col_pos, k, nr, xr, ldx, \
                                                  yr, ldy, cspan, part);                        \
  else                                                                                          \
    k_dense_rect<DD, false><<<grid, kRows, 0, s>>>(nrows, row_pos, ncols, col_pos, k, nr, xr,   \
                                                   ldx, yr, ldy, cspan, part);
    if (dim == 1) { ABH_RECT(1) }
    else if (dim == 2) { ABH_RECT(2) }
    else { ABH_RECT(3) }
#undef ABH_RECT
    abh::check_launch();
    if (part) {
      k_rect_reduce<<<unsigned(abh::ceil_div(nrows * nr, int64_t(256))), 256, 0, s>>>(
          nrows, nr, int(ns), part, yr, ldy);
      abh::check_launch();
    }
  }
}

}  // namespace

extern "C" {

abh_status abh_dense_batch_create(abh_context *ctx, int dim, int nprob, const int64_t *offsets,
                                  const double *pos, const abh_kernel *kernel,
                                  abh_dense_batch **out) {
  if (!ctx || !out || !offsets || !kernel) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_dense_batch *op = new (std::nothrow) abh_dense_batch();
  if (!op) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    ABH_REQUIRE(dim >= 1 && dim <= 3, ABH_ERR_INVALID_ARG, "dim must be 1..3");
    ABH_REQUIRE(nprob >= 1, ABH_ERR_INVALID_ARG, "nprob must be >= 1");
    ABH_REQUIRE(offsets[0] == 0, ABH_ERR_INVALID_ARG, "offsets[0] must be 0");
    ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
                ABH_ERR_UNSUPPORTED, "unknown kernel type");
    op->ctx = ctx;
    op->dim = dim;
    op->pos = pos;
    op->k.type = kernel->type;
    op->k.param = kernel->param;
    op->k.param2 = kernel->param * kernel->param;
    abh::set_inverses(op->k);
    op->L.build(ctx, offsets, nprob);
    ABH_REQUIRE(op->L.total() == 0 || pos, ABH_ERR_INVALID_ARG, "pos is null");
    // tiles, largest problem first
    std::vector<int> order(nprob);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return offsets[a + 1] - offsets[a] > offsets[b + 1] - offsets[b];
    });
    hipStream_t s = ctx->stream;
    std::vector<int2> tl[kVariants];
    for (int v = 0; v < kVariants; ++v) {
      const int64_t rows = int64_t(kRows) * kRptOf[v];
      for (int p : order)
        for (int64_t r = 0; r < offsets[p + 1] - offsets[p]; r += rows)
          tl[v].push_back(make_int2(p, int(r)));
      op->ntiles[v] = int(tl[v].size());
      op->tiles[v].ensure(tl[v].size() * sizeof(int2) + 16);
      if (!tl[v].empty())
        ABH_HIP(hipMemcpyAsync(op->tiles[v].p, tl[v].data(), tl[v].size() * sizeof(int2),
                               hipMemcpyHostToDevice, s));
    }
    op->off.ensure(size_t(nprob + 1) * 8);
    op->active.ensure(size_t(nprob) + 16);
    ABH_HIP(hipMemcpyAsync(op->off.p, offsets, size_t(nprob + 1) * 8, hipMemcpyHostToDevice, s));
    ABH_HIP(hipStreamSynchronize(s));
  });
  if (st != ABH_OK) {
    delete op;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = op;
  return ABH_OK;
}

abh_status abh_dense_batch_destroy(abh_dense_batch *op) {
  if (!op) return ABH_OK;
  abh_context *ctx = op->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete op;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_dense_batch_apply(abh_dense_batch *op, const double *x, double *y,
                                 abh_precision precision) {
  if (!op) return ABH_ERR_INVALID_ARG;
  return abh::guard(op->ctx, [&] {
    ABH_REQUIRE(precision == ABH_FP64 || precision == ABH_FP32_COMPENSATED, ABH_ERR_UNSUPPORTED,
                "dense precision must be ABH_FP64 or ABH_FP32_COMPENSATED");
    ABH_REQUIRE(op->L.total() == 0 || (x && y), ABH_ERR_INVALID_ARG, "x or y is null");
    op->apply(x, y, precision == ABH_FP32_COMPENSATED, 1.0, false, nullptr);
  });
}

abh_status abh_dense_batch_gmres(abh_dense_batch *op, const double *b, double *x,
                                 const abh_gmres_options *opts, abh_gmres_stats *stats) {
  return abh_dense_batch_gmres_precond(op, nullptr, b, x, opts, stats);
}

abh_status abh_dense_batch_schwarz_create(abh_dense_batch *op, const double *low,
                                          const double *high, double n_particles_in_leaf,
                                          double diag_shift, abh_schwarz **out) {
  if (!op || !low || !high || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_context *ctx = op->ctx;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    ABH_REQUIRE(n_particles_in_leaf > 0, ABH_ERR_INVALID_ARG,
                "n_particles_in_leaf must be positive");
    for (int d = 0; d < op->dim; ++d)
      ABH_REQUIRE(high[d] > low[d], ABH_ERR_INVALID_ARG, "high must exceed low");
    *out = abh::schwarz_create_batch(ctx, op->dim, op->L.off, op->pos, low, high,
                                     n_particles_in_leaf, op->k, diag_shift);
  });
  if (st != ABH_OK) return st;
  abh::ctx_retain(ctx);
  return ABH_OK;
}

abh_status abh_dense_batch_gmres_precond(abh_dense_batch *op, const abh_schwarz *prec,
                                         const double *b, double *x,
                                         const abh_gmres_options *opts, abh_gmres_stats *stats) {
  if (!op || !opts) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = op->ctx;
  return abh::guard(ctx, [&] {
    ABH_REQUIRE(!prec || abh::schwarz_size(prec) == op->L.total(), ABH_ERR_INVALID_ARG,
                "preconditioner size differs from the batch");
    ABH_REQUIRE(opts->restart >= 1 && opts->max_iters >= 1 && opts->tol > 0, ABH_ERR_INVALID_ARG,
                "invalid GMRES options");
    ABH_REQUIRE(opts->precision == ABH_FP64 || opts->precision == ABH_FP32_COMPENSATED,
                ABH_ERR_UNSUPPORTED, "GMRES operator precision must be FP64 or FP32_COMPENSATED");
    ABH_REQUIRE(op->L.total() == 0 || (b && x), ABH_ERR_INVALID_ARG, "b or x is null");
    ABH_REQUIRE(!opts->graph || !prec || abh::schwarz_context(prec) == ctx, ABH_ERR_INVALID_ARG,
                "a graph-captured solve needs the preconditioner on the operator's context");
    const bool fp32 = opts->precision == ABH_FP32_COMPENSATED;
    abh::BatchOp f_op = [&](const double *in, double *out, const uint8_t *mask) {
      op->apply(in, out, fp32, 1.0, true, mask);
    };
    abh::BatchOp f_true = [&](const double *in, double *out, const uint8_t *mask) {
      op->apply(in, out, false, 1.0, true, mask);
    };
    abh::BatchPrec f_prec = [&](const double *in, double *out) {
      abh::schwarz_apply(prec, in, out);
    };
    // tiles tuned before any graph capture (the tuner allocates and times)
    op->tune(fp32);
    op->tune(false);
    abh::GmresResult res;
    abh::batch_gmres(ctx, op->L, f_op, f_true, prec ? &f_prec : nullptr, b, x, *opts,
                     opts->graph != 0, res);
    if (stats)
      for (int p = 0; p < op->L.nprob; ++p) stats[p] = res.stats[p];
  });
}

abh_status abh_dense_batch_residual_dd(abh_dense_batch *op, const double *b, const double *x,
                                       double *r) {
  if (!op) return ABH_ERR_INVALID_ARG;
  return abh::guard(op->ctx, [&] {
    ABH_REQUIRE(op->L.total() == 0 || (b && x && r), ABH_ERR_INVALID_ARG, "b, x or r is null");
    op->residual_dd(x, b, r, nullptr);
  });
}

abh_status abh_dense_batch_refine(abh_dense_batch *op, const abh_schwarz *prec,
                                  const double *b, double *x, const abh_gmres_options *inner,
                                  double tol, int max_outer, abh_gmres_stats *stats) {
  if (!op || !inner) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = op->ctx;
  return abh::guard(ctx, [&] {
    ABH_REQUIRE(!prec || abh::schwarz_size(prec) == op->L.total(), ABH_ERR_INVALID_ARG,
                "preconditioner size differs from the batch");
    ABH_REQUIRE(inner->restart >= 1 && inner->max_iters >= 1 && inner->tol > 0,
                ABH_ERR_INVALID_ARG, "invalid inner GMRES options");
    ABH_REQUIRE(inner->precision == ABH_FP64 || inner->precision == ABH_FP32_COMPENSATED,
                ABH_ERR_UNSUPPORTED, "GMRES operator precision must be FP64 or FP32_COMPENSATED");
    ABH_REQUIRE(tol > 0 && max_outer >= 1, ABH_ERR_INVALID_ARG, "invalid refinement options");
    const abh::BatchLayout &L = op->L;
    const int np = L.nprob;
    const int64_t n = L.total();
    ABH_REQUIRE(n == 0 || (b && x), ABH_ERR_INVALID_ARG, "b or x is null");
    hipStream_t s = ctx->stream;
    ABH_REQUIRE(!inner->graph || !prec || abh::schwarz_context(prec) == ctx,
                ABH_ERR_INVALID_ARG,
                "a graph-captured solve needs the preconditioner on the operator's context");
    const bool fp32 = inner->precision == ABH_FP32_COMPENSATED;
    const uint8_t *act = op->active.as<uint8_t>();
    std::vector<uint8_t> active(np, 1);
    auto set_act = [&](const uint8_t *a) {
      ABH_HIP(hipMemcpyAsync(op->active.p, a, size_t(np), hipMemcpyHostToDevice, s));
      ABH_HIP(hipStreamSynchronize(s));
    };
    abh::BatchOp f_op = [&](const double *in, double *out, const uint8_t *mask) {
      op->apply(in, out, fp32, 1.0, true, mask);
    };
    abh::BatchOp f_true = [&](const double *in, double *out, const uint8_t *mask) {
      op->apply(in, out, false, 1.0, true, mask);
    };
    abh::BatchPrec f_prec = [&](const double *in, double *out) {
      abh::schwarz_apply(prec, in, out);
    };
    op->tune(fp32);
    op->tune(false);
    const int ld = 2;
    DevBuf r, d, partial, hdev, sdev;
    r.ensure(size_t(n) * 8 + 64);
    d.ensure(size_t(n) * 8 + 64);
    partial.ensure(size_t(std::max(L.nchunks, 1)) * ld * 8 + 64);
    hdev.ensure(size_t(np) * ld * 8 + 64);
    sdev.ensure(size_t(np) * 8 + 64);
    std::vector<double> hh(size_t(np) * ld), bn(np), rn(np);
    auto norms = [&](const double *v, std::vector<double> &out) {
      abh::batch_dots(ctx, L, v, n, 1, v, partial.as<double>(), hdev.as<double>(), ld);
      ABH_HIP(hipMemcpyAsync(hh.data(), hdev.p, hh.size() * 8, hipMemcpyDeviceToHost, s));
      ABH_HIP(hipStreamSynchronize(s));
      for (int p = 0; p < np; ++p) out[p] = std::sqrt(std::max(0.0, hh[size_t(p) * ld]));
    };
    std::vector<abh_gmres_stats> st(np, abh_gmres_stats{0, 0, 0.0});
    norms(b, bn);
    for (int outer = 0; outer < max_outer; ++outer) {
      // the accurate residual of the problems still refining
      set_act(active.data());
      ABH_HIP(hipMemsetAsync(r.p, 0, size_t(n) * 8, s));  // finished problems' rows stay 0
      op->residual_dd(x, b, r.as<double>(), act);
      norms(r.as<double>(), rn);
      bool any = false;
      std::vector<double> keep(np, 0.0);
      for (int p = 0; p < np; ++p) {
        if (!active[p]) continue;
        st[p].residual = bn[p] > 0 ? rn[p] / bn[p] : 0.0;
        if (bn[p] == 0.0 || rn[p] <= tol * bn[p]) {
          st[p].converged = 1;
          active[p] = 0;
          continue;
        }
        keep[p] = 1.0;
        any = true;
      }
      if (!any || outer == max_outer - 1) break;
      // K d = r for the active problems (the others' right-hand sides are
      // zeroed, so the inner solve leaves them at once), then x += d
      ABH_HIP(hipMemcpyAsync(sdev.p, keep.data(), size_t(np) * 8, hipMemcpyHostToDevice, s));
      abh::batch_scale(ctx, L, r.as<double>(), sdev.as<double>(), r.as<double>());
      ABH_HIP(hipMemsetAsync(d.p, 0, size_t(n) * 8, s));
      abh::GmresResult res;
      abh::batch_gmres(ctx, L, f_op, f_true, prec ? &f_prec : nullptr, r.as<double>(),
                       d.as<double>(), *inner, inner->graph != 0, res);
      for (int p = 0; p < np; ++p) st[p].iterations += res.stats[p].iterations;
      std::vector<double> one(size_t(np) * ld, 0.0);
      for (int p = 0; p < np; ++p) one[size_t(p) * ld] = 1.0;
      ABH_HIP(hipMemcpyAsync(hdev.p, one.data(), one.size() * 8, hipMemcpyHostToDevice, s));
      abh::batch_combine(ctx, L, d.as<double>(), 0, 1, hdev.as<double>(), ld, 1.0, x);
      ABH_HIP(hipStreamSynchronize(s));  // keep, one are host sources of the copies above
    }
    if (stats)
      for (int p = 0; p < np; ++p) stats[p] = st[p];
  });
}

abh_status abh_dense_apply(abh_context *ctx, int dim, int64_t nrows, const double *row_pos,
                           int64_t ncols, const double *col_pos, const abh_kernel *kernel,
                           int nrhs, const double *x, int64_t ldx, double *y, int64_t ldy,
                           abh_precision precision) {
  if (!ctx || !kernel) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] {
    dense_apply_impl(ctx, dim, nrows, row_pos, ncols, col_pos, kernel, nrhs, x, ldx, y, ldy,
                     precision, nullptr, 0);
  });
}

abh_status abh_dense_apply_noalloc(abh_context *ctx, int dim, int64_t nrows,
                                   const double *row_pos, int64_t ncols, const double *col_pos,
                                   const abh_kernel *kernel, int nrhs, const double *x,
                                   int64_t ldx, double *y, int64_t ldy, abh_precision precision) {
  if (!ctx || !kernel) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] {
    ABH_REQUIRE(ctx->dense_scratch, ABH_ERR_INVALID_ARG, "context has no dense scratch");
    dense_apply_impl(ctx, dim, nrows, row_pos, ncols, col_pos, kernel, nrhs, x, ldx, y, ldy,
                     precision, ctx->dense_scratch, kNoAllocScratchDoubles);
  });
}

abh_status abh_dense_batch_tune(abh_dense_batch *op, abh_precision precision, int *rows_per_thread,
                                float *variant_ms) {
  if (!op) return ABH_ERR_INVALID_ARG;
  return abh::guard(op->ctx, [&] {
    ABH_REQUIRE(precision == ABH_FP64 || precision == ABH_FP32_COMPENSATED, ABH_ERR_UNSUPPORTED,
                "dense precision must be ABH_FP64 or ABH_FP32_COMPENSATED");
    const bool fp32 = precision == ABH_FP32_COMPENSATED;
    if (rows_per_thread && *rows_per_thread > 0) {  // forced choice
      int v = -1;
      for (int q = 0; q < kVariants; ++q)
        if (kRptOf[q] == *rows_per_thread) v = q;
      ABH_REQUIRE(v >= 0, ABH_ERR_INVALID_ARG, "rows_per_thread must be 1, 2 or 4");
      op->choice[fp32 ? 1 : 0] = v;
      return;
    }
    const int v = op->tune(fp32);
    if (rows_per_thread) *rows_per_thread = kRptOf[v];
    if (variant_ms)
      for (int q = 0; q < kVariants; ++q) variant_ms[q] = op->tune_ms[fp32 ? 1 : 0][q];
  });
}

}  // extern "C"
