// GMRES on a square sparse kernel operator plus a diagonal shift, right
// preconditioned by the Schwarz or the approximate-inverse preconditioner.
// The Arnoldi process, restarts and FP64 true residual are the batched
// engine's (batch.hip) with one problem.
#include <functional>

#include "batch.hpp"
#include "solvers.hpp"

namespace {

// out += shift * in
__global__ void k_axpy(double *__restrict__ out, const double *__restrict__ in, int64_t n,
                       double shift) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = out[i] + shift * in[i];
}

}  // namespace

extern "C" {

abh_status abh_sparse_op_gmres(abh_sparse_op *op, double diag_shift, const abh_schwarz *schwarz,
                               const abh_ainv *ainv, const double *b, double *x,
                               const abh_gmres_options *opts, abh_gmres_stats *stats) {
  if (!op || !opts) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = abh::sparse_context(op);
  return abh::guard(ctx, [&] {
    const int64_t n = abh::sparse_rows(op);
    ABH_REQUIRE(abh::sparse_cols(op) == n, ABH_ERR_INVALID_ARG, "operator is not square");
    ABH_REQUIRE(!(schwarz && ainv), ABH_ERR_INVALID_ARG, "at most one preconditioner");
    ABH_REQUIRE(!schwarz || abh::schwarz_size(schwarz) == n, ABH_ERR_INVALID_ARG,
                "preconditioner size differs from the operator");
    ABH_REQUIRE(!ainv || abh::ainv_size(ainv) == n, ABH_ERR_INVALID_ARG,
                "preconditioner size differs from the operator");
    ABH_REQUIRE(opts->restart >= 1 && opts->max_iters >= 1 && opts->tol > 0, ABH_ERR_INVALID_ARG,
                "invalid GMRES options");
    ABH_REQUIRE(opts->precision == ABH_FP64 || opts->precision == ABH_FP32_COMPENSATED,
                ABH_ERR_UNSUPPORTED, "GMRES operator precision must be FP64 or FP32_COMPENSATED");
    ABH_REQUIRE(n == 0 || (b && x), ABH_ERR_INVALID_ARG, "b or x is null");
    ABH_REQUIRE(opts->graph == 0, ABH_ERR_UNSUPPORTED,
                "graph-captured GMRES cycles are for dense batches");
    const int64_t off[2] = {0, n};
    abh::BatchLayout L;
    L.build(ctx, off, 1);
    hipStream_t s = ctx->stream;
    const bool fp32 = opts->precision == ABH_FP32_COMPENSATED;
    auto apply = [&](const double *in, double *out, bool f32) {
      ABH_HIP(hipMemsetAsync(out, 0, size_t(n) * 8, s));
      abh::sparse_apply(op, in, out, f32);
      if (diag_shift != 0.0 && n > 0) {
        k_axpy<<<unsigned(abh::ceil_div(n, 256)), 256, 0, s>>>(out, in, n, diag_shift);
        abh::check_launch();
      }
    };
    abh::BatchOp f_op = [&](const double *in, double *out, const uint8_t *) {
      apply(in, out, fp32);
    };
    abh::BatchOp f_true = [&](const double *in, double *out, const uint8_t *) {
      apply(in, out, false);
    };
    abh::BatchPrec f_prec = [&](const double *in, double *out) {
      if (schwarz) abh::schwarz_apply(schwarz, in, out);
      else abh::ainv_apply(ainv, in, out);
    };
    abh::GmresResult res;
    abh::batch_gmres(ctx, L, f_op, f_true, (schwarz || ainv) ? &f_prec : nullptr, b, x, *opts,
                     false, res);
    if (stats) stats[0] = res.stats[0];
  });
}

}  // extern "C"
