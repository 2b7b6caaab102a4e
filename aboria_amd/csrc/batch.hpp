// Batched vectors (many independent problems stored back to back) and the
// restarted GMRES(m) engine that runs every problem of a batch in lockstep.
#pragma once

#include <functional>
#include <vector>

#include "common.hpp"

namespace abh {

// A run of at most kChunk consecutive entries of one problem.
struct Chunk {
  int32_t prob;
  int32_t pad;
  int64_t start;
  int64_t len;
};

struct BatchLayout {
  int nprob = 0;
  std::vector<int64_t> off;  // nprob + 1
  std::vector<int32_t> chunk_ptr_h;
  int nchunks = 0;
  DevBuf chunks, chunk_ptr;
  int64_t total() const { return off.empty() ? 0 : off.back(); }
  void build(abh_context *ctx, const int64_t *offsets, int np);
};

// Vector kernels over a batch (batch.hip).
// partial[c * ld + i] = <V_i, w> restricted to chunk c, then h[p * ld + i]
// sums the chunks of problem p in order (deterministic).
void batch_dots(abh_context *ctx, const BatchLayout &L, const double *V, int64_t vstride, int nv,
                const double *w, double *partial, double *h, int ld);
// out[j] += sign * sum_i coef[p(j) * ld + i] * V_i[j]
void batch_combine(abh_context *ctx, const BatchLayout &L, const double *V, int64_t vstride,
                   int nv, const double *coef, int ld, double sign, double *out);
// out[j] = in[j] * s[p(j)]
void batch_scale(abh_context *ctx, const BatchLayout &L, const double *in, const double *s,
                 double *out);

// Restarted GMRES(m), right preconditioned, every problem of the batch in
// lockstep.  op(in, out, mask): out = A in for the problems whose mask byte
// (device, one per problem) is set, other rows unspecified; prec(in, out):
// out = M^-1 in.  The Arnoldi basis is orthogonalised by classical
// Gram-Schmidt applied twice (CGS2) in FP64.  The steps of a restart cycle
// run on the device without host round trips: the CGS2 coefficients, the
// norms, the Givens reduction of each problem's Hessenberg column and its
// convergence test are kernels over the batch, and a problem that leaves
// the cycle is masked out of the operator and the updates.  With graph the
// steps of a cycle run as HIP graphs of a few steps each, captured on first
// use and replayed every cycle, the host checking between them whether any
// problem is still iterating (op and prec must then be capturable: kernels
// and device copies only); without it the host checks after every step.  Each restart recomputes the true
// residual with op_true (FP64 operator), for the problems still iterating;
// a problem whose true residual grew over a cycle is put back to its
// previous iterate and stopped, and with opt.stall_restarts > 0 one whose
// residual fell by less than 10% on that many restarts in a row stops too.
struct GmresResult {
  std::vector<abh_gmres_stats> stats;
};

using BatchOp = std::function<void(const double *, double *, const uint8_t *)>;
using BatchPrec = std::function<void(const double *, double *)>;

void batch_gmres(abh_context *ctx, const BatchLayout &L, const BatchOp &op,
                 const BatchOp &op_true, const BatchPrec *prec, const double *b, double *x,
                 const abh_gmres_options &opt, bool graph, GmresResult &res);

}  // namespace abh
