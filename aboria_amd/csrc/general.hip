// General sparse kernel operator: a radius per row (create_sparse_operator
// with a radius function, src/Operators.h:472-482, KernelSparse<..., FRadius,
// ...> src/Kernels.h:578-751) and matrix-valued kernels of BlockRows x
// BlockCols entries (the Eigen block the reference's functor may return,
// src/detail/Kernels.h:336-367).  Each row walks its neighbours with the
// device query of include/aboria_hip_device.hpp (the reference's visit rule
// and inclusive FP64 test), one thread per row: this is the general form
// beside the tuned constant-radius scalar operator of search.hip.
//
// Entry (ii, jj) of the block of pair (i, j) is
//   phi(|dx|) * (w0[e] + w_row[e] * a_i + w_col[e] * b_j),  e = ii * BC + jj,
// with a_i / b_j the row / column scalars (0 when not given).  Apply adds,
// per row and neighbour in visit order, the block times the column's
// segment of x to the row's segment of y (lhs.segment<BR>(i*BR) += f *
// rhs.segment<BC>(j*BC), src/Kernels.h:740-750), accumulating each block
// product before adding it.
#include <algorithm>

#include "../../include/aboria_hip_device.hpp"
#include "cells.hpp"
#include "kernel_fn.hpp"

using abh::DevBuf;
using abh::KernelDesc;

namespace {

constexpr int kMaxBlock = 4;

struct BlockDesc {
  KernelDesc k;
  int br, bc;
  double w0[16], wr[16], wc[16];
};

__device__ __forceinline__ double entry(const BlockDesc &kb, int e, double phi, double a,
                                        double b) {
  return phi * ((kb.w0[e] + kb.wr[e] * a) + kb.wc[e] * b);
}

template <int D>
__global__ __launch_bounds__(256) void k_gen_apply(abh_query_view q, int64_t nrows,
                                                   const double *__restrict__ rp,
                                                   double radius,
                                                   const double *__restrict__ rrad,
                                                   BlockDesc kb, const double *__restrict__ rs,
                                                   const double *__restrict__ cs,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nrows) return;
  const double r = rrad ? rrad[i] : radius;
  const double a = rs ? rs[i] : 0.0;
  double acc[kMaxBlock] = {0.0, 0.0, 0.0, 0.0};
  double c[D];
#pragma unroll
  for (int d = 0; d < D; ++d) c[d] = rp[i * D + d];
  abh::for_each_neighbour_d<D>(q, c, r, [&](uint32_t j, const double *, double d2) {
    const double phi = abh::phi64(kb.k, d2);
    const double b = cs ? cs[j] : 0.0;
    for (int ii = 0; ii < kb.br; ++ii) {
      double t = 0.0;
      for (int jj = 0; jj < kb.bc; ++jj)
        t = t + entry(kb, ii * kb.bc + jj, phi, a, b) * x[int64_t(j) * kb.bc + jj];
      acc[ii] = acc[ii] + t;
    }
  });
  for (int ii = 0; ii < kb.br; ++ii) y[i * kb.br + ii] += acc[ii];
}

template <int D>
__global__ __launch_bounds__(256) void k_gen_count(abh_query_view q, int64_t nrows,
                                                   const double *__restrict__ rp, double radius,
                                                   const double *__restrict__ rrad,
                                                   int64_t *__restrict__ cnt) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nrows) return;
  const double r = rrad ? rrad[i] : radius;
  double c[D];
#pragma unroll
  for (int d = 0; d < D; ++d) c[d] = rp[i * D + d];
  int64_t n = 0;
  abh::for_each_neighbour_d<D>(q, c, r, [&](uint32_t, const double *, double) { ++n; });
  cnt[i] = n;
}

// CSR of the scalar matrix: block row i * BR + ii holds, per neighbour j in
// visit order, the entries of columns j * BC .. j * BC + BC - 1
template <int D>
__global__ __launch_bounds__(256) void k_gen_fill(abh_query_view q, int64_t nrows,
                                                  const double *__restrict__ rp, double radius,
                                                  const double *__restrict__ rrad, BlockDesc kb,
                                                  const double *__restrict__ rs,
                                                  const double *__restrict__ cs,
                                                  const int64_t *__restrict__ poff,
                                                  int64_t *__restrict__ row_ptr,
                                                  int32_t *__restrict__ col,
                                                  double *__restrict__ val) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nrows) return;
  const double r = rrad ? rrad[i] : radius;
  const double a = rs ? rs[i] : 0.0;
  const int64_t p0 = poff[i], np = poff[i + 1] - p0;
  const int64_t per = int64_t(kb.br) * kb.bc;
  for (int ii = 0; ii < kb.br; ++ii) row_ptr[i * kb.br + ii] = p0 * per + ii * np * kb.bc;
  if (i == nrows - 1) row_ptr[nrows * kb.br] = poff[nrows] * per;
  double c[D];
#pragma unroll
  for (int d = 0; d < D; ++d) c[d] = rp[i * D + d];
  int64_t k = 0;
  abh::for_each_neighbour_d<D>(q, c, r, [&](uint32_t j, const double *, double d2) {
    const double phi = abh::phi64(kb.k, d2);
    const double b = cs ? cs[j] : 0.0;
    for (int ii = 0; ii < kb.br; ++ii)
      for (int jj = 0; jj < kb.bc; ++jj) {
        const int64_t e = p0 * per + ii * np * kb.bc + k * kb.bc + jj;
        col[e] = int32_t(int64_t(j) * kb.bc + jj);
        val[e] = entry(kb, ii * kb.bc + jj, phi, a, b);
      }
    ++k;
  });
}

abh_query_view make_view(const abh_cells *c) {
  abh_query_view v{};
  v.dim = c->dim;
  for (int d = 0; d < 3; ++d) {
    const bool on = d < c->dim;
    v.periodic[d] = on && c->periodic[d];
    v.size[d] = on ? int32_t(c->size[d]) : 1;
    v.end_bucket[d] = v.size[d] - 1;
    v.low[d] = on ? c->lo[d] : 0.0;
    v.high[d] = on ? c->hi[d] : 1.0;
    v.side[d] = on ? c->side[d] : 1.0;
    v.inv_side[d] = on ? c->inv[d] : 1.0;
  }
  v.pos = c->pos.as<double>();
  v.bucket_begin = c->begin.as<uint32_t>();
  v.bucket_end = c->end.as<uint32_t>();
  v.n = c->n_alive;
  v.n_buckets = c->nb;
  return v;
}

}  // namespace

struct abh_sparse_block_op {
  abh_context *ctx = nullptr;
  abh_query_view view{};
  int64_t nrows = 0, ncols = 0;
  const double *row_pos = nullptr, *row_radius = nullptr, *rs = nullptr, *cs = nullptr;
  double radius = 0.0;
  BlockDesc kb{};
  int64_t nnz = 0;  // neighbour pairs
  DevBuf poff, total;
};

namespace {

template <class F> void by_dim(int dim, F &&f) {
  if (dim == 1) f(std::integral_constant<int, 1>());
  else if (dim == 2) f(std::integral_constant<int, 2>());
  else f(std::integral_constant<int, 3>());
}

// pair counts per row, scanned into poff (nrows + 1); returns the total
int64_t count_pairs(abh_sparse_block_op *op) {
  abh_context *ctx = op->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = op->nrows;
  op->poff.ensure(size_t(n + 1) * 8 + 64);
  op->total.ensure(64);
  int64_t *po = op->poff.as<int64_t>();
  ABH_HIP(hipMemsetAsync(po + n, 0, 8, st));
  if (n > 0)
    by_dim(op->view.dim, [&](auto DD) {
      k_gen_count<decltype(DD)::value><<<unsigned(abh::ceil_div(n, 256)), 256, 0, st>>>(
          op->view, n, op->row_pos, op->radius, op->row_radius, po);
    });
  abh::check_launch();
  abh::scan_exclusive(ctx, po, po, n + 1, op->total.as<int64_t>());
  int64_t t = 0;
  ABH_HIP(hipMemcpyAsync(&t, op->total.p, 8, hipMemcpyDeviceToHost, st));
  ABH_HIP(hipStreamSynchronize(st));
  return t;
}

}  // namespace

extern "C" {

abh_status abh_cells_query_view(const abh_cells *cells, abh_query_view *view) {
  if (!cells || !view) return ABH_ERR_INVALID_ARG;
  *view = make_view(cells);
  return ABH_OK;
}

abh_status abh_sparse_block_op_create(abh_context *ctx, const abh_cells *cols, int64_t nrows,
                                      const double *row_pos, double radius,
                                      const double *row_radius, const abh_block_kernel *kernel,
                                      const double *row_scalar, const double *col_scalar,
                                      abh_sparse_block_op **out) {
  if (!ctx || !cols || !kernel || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_sparse_block_op *op = new (std::nothrow) abh_sparse_block_op();
  if (!op) return ABH_ERR_OUT_OF_MEMORY;
  op->ctx = ctx;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    ABH_REQUIRE(nrows >= 0 && (nrows == 0 || row_pos), ABH_ERR_INVALID_ARG, "bad rows");
    ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
                ABH_ERR_UNSUPPORTED, "unknown kernel type");
    ABH_REQUIRE(kernel->block_rows >= 1 && kernel->block_rows <= kMaxBlock &&
                    kernel->block_cols >= 1 && kernel->block_cols <= kMaxBlock,
                ABH_ERR_UNSUPPORTED, "block rows and columns must be 1..4");
    ABH_REQUIRE(row_radius || radius >= 0.0, ABH_ERR_INVALID_ARG, "negative radius");
    op->view = make_view(cols);
    op->nrows = nrows;
    op->ncols = cols->n_alive;
    op->row_pos = row_pos;
    op->row_radius = row_radius;
    op->radius = radius;
    op->rs = row_scalar;
    op->cs = col_scalar;
    BlockDesc &kb = op->kb;
    kb.k.type = kernel->type;
    kb.k.param = kernel->param;
    kb.k.param2 = kernel->param * kernel->param;
    abh::set_inverses(kb.k);
    kb.br = kernel->block_rows;
    kb.bc = kernel->block_cols;
    for (int e = 0; e < 16; ++e) {
      kb.w0[e] = kernel->w0[e];
      kb.wr[e] = kernel->w_row[e];
      kb.wc[e] = kernel->w_col[e];
    }
    op->nnz = count_pairs(op);
  });
  if (st != ABH_OK) {
    delete op;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = op;
  return ABH_OK;
}

abh_status abh_sparse_block_op_destroy(abh_sparse_block_op *op) {
  if (!op) return ABH_OK;
  abh_context *ctx = op->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete op;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_sparse_block_op_pairs(const abh_sparse_block_op *op, int64_t *npairs) {
  if (!op || !npairs) return ABH_ERR_INVALID_ARG;
  *npairs = op->nnz;
  return ABH_OK;
}

abh_status abh_sparse_block_op_apply(abh_sparse_block_op *op, const double *x, double *y) {
  if (!op) return ABH_ERR_INVALID_ARG;
  return abh::guard(op->ctx, [&] {
    if (op->nrows == 0) return;
    ABH_REQUIRE(y && (op->ncols == 0 || x), ABH_ERR_INVALID_ARG, "null pointer");
    hipStream_t st = op->ctx->stream;
    by_dim(op->view.dim, [&](auto DD) {
      k_gen_apply<decltype(DD)::value><<<unsigned(abh::ceil_div(op->nrows, 256)), 256, 0, st>>>(
          op->view, op->nrows, op->row_pos, op->radius, op->row_radius, op->kb, op->rs, op->cs,
          x, y);
    });
    abh::check_launch();
  });
}

abh_status abh_sparse_block_op_assemble(abh_sparse_block_op *op, int64_t *row_ptr, int32_t *col,
                                        double *val, int64_t capacity) {
  if (!op) return ABH_ERR_INVALID_ARG;
  return abh::guard(op->ctx, [&] {
    const int64_t per = int64_t(op->kb.br) * op->kb.bc;
    ABH_REQUIRE(capacity >= op->nnz * per, ABH_ERR_CAPACITY, "col / val capacity below the entries");
    ABH_REQUIRE(row_ptr, ABH_ERR_INVALID_ARG, "row_ptr is null");
    hipStream_t st = op->ctx->stream;
    if (op->nrows == 0) {
      ABH_HIP(hipMemsetAsync(row_ptr, 0, 8, st));
      return;
    }
    ABH_REQUIRE(op->nnz == 0 || (col && val), ABH_ERR_INVALID_ARG, "col or val is null");
    by_dim(op->view.dim, [&](auto DD) {
      k_gen_fill<decltype(DD)::value><<<unsigned(abh::ceil_div(op->nrows, 256)), 256, 0, st>>>(
          op->view, op->nrows, op->row_pos, op->radius, op->row_radius, op->kb, op->rs, op->cs,
          op->poff.as<int64_t>(), row_ptr, col, val);
    });
    abh::check_launch();
  });
}

}  // extern "C"
