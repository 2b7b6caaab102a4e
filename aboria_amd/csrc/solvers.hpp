// Internal entry points shared by the operator, preconditioner and Krylov
// translation units (search.hip, schwarz.hip, learned.hip, krylov.hip).
#pragma once

#include <vector>

#include "common.hpp"
#include "kernel_fn.hpp"

namespace abh {

// y += K x for one right-hand side (search.hip).
void sparse_apply(abh_sparse_op *op, const double *x, double *y, bool fp32);
int64_t sparse_rows(const abh_sparse_op *op);
int64_t sparse_cols(const abh_sparse_op *op);
abh_context *sparse_context(const abh_sparse_op *op);

// Device address, on the current device, of the static scratch of the
// allocation-free dense apply (dense.hip); throws abh::Error.
double *dense_scratch_address();

// x = M^-1 b, restricted additive Schwarz (schwarz.hip).
void schwarz_apply(const abh_schwarz *s, const double *b, double *x);
int64_t schwarz_size(const abh_schwarz *s);
abh_context *schwarz_context(const abh_schwarz *s);
// Batched restricted additive Schwarz over independent particle sets
// (schwarz.hip): signed factorisation, no kernel truncation.  Throws abh::Error.
abh_schwarz *schwarz_create_batch(abh_context *ctx, int dim, const std::vector<int64_t> &off,
                                  const double *pos_dev, const double *lo, const double *hi,
                                  double n_leaf, const KernelDesc &k, double shift);

// Strided-batched FP64 GEMM on the matrix cores, column-major
// (dgemm.hip): C[g] = alpha op(A[g]) op(B[g]) + beta C[g] for g < batch;
// beta == 0 does not read C.  lower_only skips the 64 x 64 tiles of C that
// lie strictly above its diagonal.  Enqueued on st; throws abh::Error.
void dgemm_strided_batched(hipStream_t st, bool ta, bool tb, int m, int n, int k, double alpha,
                           const double *A, int lda, int64_t sA, const double *B, int ldb,
                           int64_t sB, double beta, double *C, int ldc, int64_t sC, int batch,
                           bool lower_only = false);

// x = W b with W the approximate-inverse rows (learned.hip).
void ainv_apply(const abh_ainv *a, const double *b, double *x);
int64_t ainv_size(const abh_ainv *a);

}  // namespace abh
