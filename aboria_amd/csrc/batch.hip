// Batched vector kernels and the lockstep restarted GMRES(m) engine.
#include <algorithm>
#include <cmath>

#include "batch.hpp"

namespace abh {

namespace {

constexpr int kChunk = 4096;
constexpr int kThreads = 256;
constexpr int kSegment = 5;  // Arnoldi steps per captured graph

__global__ __launch_bounds__(kThreads) void k_multidot(const Chunk *__restrict__ ch,
                                                       const double *__restrict__ V,
                                                       int64_t vstride, int nv,
                                                       const double *__restrict__ w,
                                                       double *__restrict__ partial, int ld) {
  __shared__ double red[kThreads / 64];
  const Chunk c = ch[blockIdx.x];
  for (int i = 0; i < nv; ++i) {
    const double *vi = V + int64_t(i) * vstride;
    double s = 0.0;
    for (int64_t t = threadIdx.x; t < c.len; t += kThreads) s += vi[c.start + t] * w[c.start + t];
    s = wave_sum(s);
    if (lane_id() == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0;
      for (int q = 0; q < kThreads / 64; ++q) tot += red[q];
      partial[int64_t(blockIdx.x) * ld + i] = tot;
    }
    __syncthreads();
  }
}

__global__ void k_chunk_reduce(const int32_t *__restrict__ chunk_ptr, int nprob, int nv,
                               const double *__restrict__ partial, double *__restrict__ h,
                               int ld) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nprob * nv) return;
  const int p = t / nv, i = t % nv;
  double s = 0.0;
  for (int c = chunk_ptr[p]; c < chunk_ptr[p + 1]; ++c) s += partial[int64_t(c) * ld + i];
  h[int64_t(p) * ld + i] = s;
}

__global__ __launch_bounds__(kThreads) void k_combine(const Chunk *__restrict__ ch,
                                                      const double *__restrict__ V,
                                                      int64_t vstride, int nv,
                                                      const double *__restrict__ coef, int ld,
                                                      double sign, double *__restrict__ out) {
  const Chunk c = ch[blockIdx.x];
  const double *cp = coef + int64_t(c.prob) * ld;
  for (int64_t t = threadIdx.x; t < c.len; t += kThreads) {
    double s = 0.0;
    for (int i = 0; i < nv; ++i) s += cp[i] * V[int64_t(i) * vstride + c.start + t];
    out[c.start + t] += sign * s;
  }
}

__global__ __launch_bounds__(kThreads) void k_scale(const Chunk *__restrict__ ch,
                                                    const double *__restrict__ in,
                                                    const double *__restrict__ s,
                                                    double *__restrict__ out) {
  const Chunk c = ch[blockIdx.x];
  const double f = s[c.prob];
  for (int64_t t = threadIdx.x; t < c.len; t += kThreads)
    out[c.start + t] = in[c.start + t] * f;
}

// zero the coefficients of the problems outside the cycle and accumulate
// the two CGS passes: htot = (first ? 0 : htot) + h
__global__ void k_mask_acc(int np, int nv, int ld, const uint8_t *__restrict__ cyc,
                           double *__restrict__ h, double *__restrict__ htot, int first) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * nv) return;
  const int p = t / nv, i = t % nv;
  const int64_t q = int64_t(p) * ld + i;
  const double v = cyc[p] ? h[q] : 0.0;
  h[q] = v;
  htot[q] = first ? v : htot[q] + v;
}

// hk1 = ||w||, sc = 1 / hk1 for the problems in the cycle (0 otherwise)
__global__ void k_norm_scale(int np, int ld, const uint8_t *__restrict__ cyc,
                             const double *__restrict__ h, double *__restrict__ hk1,
                             double *__restrict__ sc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double v = sqrt(fmax(0.0, h[int64_t(p) * ld]));
  hk1[p] = v;
  sc[p] = (cyc[p] && v > 0.0) ? 1.0 / v : 0.0;
}

// a restart cycle begins for the active problems with iterations left:
// v_0 = r / beta, g = beta e_1
__global__ void k_cycle_start(int np, int m, const uint8_t *__restrict__ active,
                              const int32_t *__restrict__ iters, int max_iters,
                              const double *__restrict__ beta, uint8_t *__restrict__ cyc,
                              double *__restrict__ sc, double *__restrict__ g,
                              int32_t *__restrict__ kp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const bool on = active[p] && iters[p] < max_iters;
  cyc[p] = on ? 1 : 0;
  sc[p] = on ? 1.0 / beta[p] : 0.0;
  kp[p] = 0;
  double *gp = g + int64_t(p) * (m + 1);
  for (int i = 0; i <= m; ++i) gp[i] = 0.0;
  gp[0] = beta[p];
}

// Arnoldi step k of every problem in the cycle: column k of the Hessenberg
// matrix (H column-major, (m+1) x m per problem) from the CGS2 coefficients
// and ||w||, the previous Givens rotations applied to it, a new rotation
// zeroing H[k+1][k], the rotated right-hand side g; the problem leaves the
// cycle when |g[k+1]| <= tol ||b||, at a breakdown, or out of iterations.
__global__ void k_givens(int np, int k, int m, int ld, uint8_t *__restrict__ cyc,
                         const double *__restrict__ htot, const double *__restrict__ hk1,
                         double *__restrict__ H, double *__restrict__ cs, double *__restrict__ sn,
                         double *__restrict__ g, int32_t *__restrict__ kp,
                         int32_t *__restrict__ iters, const double *__restrict__ bnorm, double tol,
                         int max_iters) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np || !cyc[p]) return;
  double *Hc = H + (int64_t(p) * m + k) * (m + 1);
  for (int i = 0; i <= k; ++i) Hc[i] = htot[int64_t(p) * ld + i];
  Hc[k + 1] = hk1[p];
  double *c = cs + int64_t(p) * m, *s = sn + int64_t(p) * m;
  for (int i = 0; i < k; ++i) {
    const double t = c[i] * Hc[i] + s[i] * Hc[i + 1];
    Hc[i + 1] = -s[i] * Hc[i] + c[i] * Hc[i + 1];
    Hc[i] = t;
  }
  const double den = hypot(Hc[k], Hc[k + 1]);
  c[k] = den > 0.0 ? Hc[k] / den : 1.0;
  s[k] = den > 0.0 ? Hc[k + 1] / den : 0.0;
  Hc[k] = den;
  Hc[k + 1] = 0.0;
  double *gp = g + int64_t(p) * (m + 1);
  gp[k + 1] = -s[k] * gp[k];
  gp[k] = c[k] * gp[k];
  kp[p] = k + 1;
  const int it = iters[p] + 1;
  iters[p] = it;
  if (fabs(gp[k + 1]) <= tol * bnorm[p] || hk1[p] == 0.0 || it >= max_iters) cyc[p] = 0;
}

// y = H^-1 g over the problem's kp steps (upper triangular), zero beyond
__global__ void k_hsolve(int np, int m, int ld, const int32_t *__restrict__ kp,
                         const double *__restrict__ H, const double *__restrict__ g,
                         double *__restrict__ y) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int kk = kp[p];
  double *yp = y + int64_t(p) * ld;
  for (int i = kk; i < ld; ++i) yp[i] = 0.0;
  for (int i = kk - 1; i >= 0; --i) {
    double t = g[int64_t(p) * (m + 1) + i];
    for (int j = i + 1; j < kk; ++j) t -= H[(int64_t(p) * m + j) * (m + 1) + i] * yp[j];
    const double d = H[(int64_t(p) * m + i) * (m + 1) + i];
    yp[i] = d != 0.0 ? t / d : 0.0;
  }
}

// r = b - r
__global__ void k_rsub(int64_t n, const double *__restrict__ b, double *__restrict__ r) {
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (j < n) r[j] = b[j] - r[j];
}

// x += z
__global__ void k_add(int64_t n, const double *__restrict__ z, double *__restrict__ x) {
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (j < n) x[j] += z[j];
}

// x += back_p (xprev - x): a problem whose residual grew returns to xprev
__global__ void k_restore(const Chunk *__restrict__ ch, const double *__restrict__ back,
                          const double *__restrict__ xprev, double *__restrict__ x) {
  const Chunk c = ch[blockIdx.x];
  if (back[c.prob] == 0.0) return;
  for (int64_t t = threadIdx.x; t < c.len; t += kThreads) x[c.start + t] = xprev[c.start + t];
}

}  // namespace

void BatchLayout::build(abh_context *ctx, const int64_t *offsets, int np) {
  nprob = np;
  off.assign(offsets, offsets + np + 1);
  std::vector<Chunk> chs;
  chunk_ptr_h.assign(size_t(np) + 1, 0);
  for (int p = 0; p < np; ++p) {
    ABH_REQUIRE(off[p + 1] >= off[p], ABH_ERR_INVALID_ARG, "offsets must be non-decreasing");
    for (int64_t s = off[p]; s < off[p + 1]; s += kChunk)
      chs.push_back(Chunk{p, 0, s, std::min<int64_t>(kChunk, off[p + 1] - s)});
    chunk_ptr_h[p + 1] = int32_t(chs.size());
  }
  nchunks = int(chs.size());
  chunks.ensure(chs.size() * sizeof(Chunk) + 16);
  chunk_ptr.ensure(chunk_ptr_h.size() * 4 + 16);
  if (!chs.empty())
    ABH_HIP(hipMemcpyAsync(chunks.p, chs.data(), chs.size() * sizeof(Chunk),
                           hipMemcpyHostToDevice, ctx->stream));
  ABH_HIP(hipMemcpyAsync(chunk_ptr.p, chunk_ptr_h.data(), chunk_ptr_h.size() * 4,
                         hipMemcpyHostToDevice, ctx->stream));
  ABH_HIP(hipStreamSynchronize(ctx->stream));
}

void batch_dots(abh_context *ctx, const BatchLayout &L, const double *V, int64_t vstride, int nv,
                const double *w, double *partial, double *h, int ld) {
  if (L.nchunks == 0 || nv == 0) return;
  k_multidot<<<L.nchunks, kThreads, 0, ctx->stream>>>(L.chunks.as<Chunk>(), V, vstride, nv, w,
                                                      partial, ld);
  check_launch();
  const int t = L.nprob * nv;
  k_chunk_reduce<<<(t + 255) / 256, 256, 0, ctx->stream>>>(L.chunk_ptr.as<int32_t>(), L.nprob,
                                                            nv, partial, h, ld);
  check_launch();
}

void batch_combine(abh_context *ctx, const BatchLayout &L, const double *V, int64_t vstride,
                   int nv, const double *coef, int ld, double sign, double *out) {
  if (L.nchunks == 0 || nv == 0) return;
  k_combine<<<L.nchunks, kThreads, 0, ctx->stream>>>(L.chunks.as<Chunk>(), V, vstride, nv, coef,
                                                     ld, sign, out);
  check_launch();
}

void batch_scale(abh_context *ctx, const BatchLayout &L, const double *in, const double *s,
                 double *out) {
  if (L.nchunks == 0) return;
  k_scale<<<L.nchunks, kThreads, 0, ctx->stream>>>(L.chunks.as<Chunk>(), in, s, out);
  check_launch();
}

void batch_gmres(abh_context *ctx, const BatchLayout &L, const BatchOp &op,
                 const BatchOp &op_true, const BatchPrec *prec, const double *b, double *x,
                 const abh_gmres_options &opt, bool graph, GmresResult &res) {
  CaptureStream swap;
  if (graph) swap.begin(ctx);
  hipStream_t s = ctx->stream;
  const int np = L.nprob;
  const int64_t n = L.total();
  const int m = std::max(1, opt.restart);
  const int ld = m + 2;
  const int max_iters = std::max(1, opt.max_iters);
  const unsigned gp = unsigned((np + 255) / 256), gn = unsigned(std::max<int64_t>(1, (n + 255) / 256));
  DevBuf V, w, z, partial, hdev, htot, xprev, Hd, csd, snd, gd, kpd, itd, hk1, sc, bnd, betad,
      cyc, act, back;
  V.ensure(size_t(m + 1) * size_t(n) * 8 + 64);
  xprev.ensure(size_t(n) * 8 + 64);
  w.ensure(size_t(n) * 8 + 64);
  z.ensure(size_t(n) * 8 + 64);
  partial.ensure(size_t(std::max(L.nchunks, 1)) * ld * 8);
  hdev.ensure(size_t(np) * ld * 8 + 64);
  htot.ensure(size_t(np) * ld * 8 + 64);
  Hd.ensure(size_t(np) * m * (m + 1) * 8 + 64);
  csd.ensure(size_t(np) * m * 8 + 64);
  snd.ensure(size_t(np) * m * 8 + 64);
  gd.ensure(size_t(np) * (m + 1) * 8 + 64);
  kpd.ensure(size_t(np) * 4 + 64);
  itd.ensure(size_t(np) * 4 + 64);
  hk1.ensure(size_t(np) * 8 + 64);
  sc.ensure(size_t(np) * 8 + 64);
  bnd.ensure(size_t(np) * 8 + 64);
  betad.ensure(size_t(np) * 8 + 64);
  cyc.ensure(size_t(np) + 64);
  act.ensure(size_t(np) + 64);
  back.ensure(size_t(np) * 8 + 64);
  double *Vp = V.as<double>();
  auto vec = [&](int i) { return Vp + int64_t(i) * n; };
  std::vector<double> hh(size_t(np) * ld), bnorm(np), beta(np);
  res.stats.assign(np, abh_gmres_stats{0, 0, 0.0});
  std::vector<uint8_t> active(np, 1), cyc_h(np, 0);
  std::vector<int32_t> iters(np, 0), kp(np, 0);
  // host -> device copies of host vectors reused right after: wait for them
  auto up = [&](void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    ABH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    ABH_HIP(hipStreamSynchronize(s));
  };
  auto norms = [&](const double *v, std::vector<double> &out) {
    batch_dots(ctx, L, v, n, 1, v, partial.as<double>(), hdev.as<double>(), ld);
    ABH_HIP(hipMemcpyAsync(hh.data(), hdev.p, size_t(np) * ld * 8, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < np; ++p) out[p] = std::sqrt(std::max(0.0, hh[size_t(p) * ld]));
  };
  // r = b - A x (FP64) for the problems still iterating (the rows of the
  // others are left unspecified)
  auto residual = [&](double *r) {
    up(act.p, active.data(), size_t(np));
    op_true(x, r, act.as<uint8_t>());
    if (n > 0) {
      k_rsub<<<gn, 256, 0, s>>>(n, b, r);
      check_launch();
    }
  };
  // Arnoldi step k on the device
  auto step = [&](int k) {
    const double *zin = vec(k);
    if (prec) {
      (*prec)(vec(k), z.as<double>());
      zin = z.as<double>();
    }
    op(zin, w.as<double>(), cyc.as<uint8_t>());
    for (int pass = 0; pass < 2; ++pass) {  // CGS2 against v_0..v_k
      batch_dots(ctx, L, Vp, n, k + 1, w.as<double>(), partial.as<double>(), hdev.as<double>(), ld);
      if (np > 0) {
        k_mask_acc<<<unsigned((np * (k + 1) + 255) / 256), 256, 0, s>>>(
            np, k + 1, ld, cyc.as<uint8_t>(), hdev.as<double>(), htot.as<double>(), pass == 0);
        check_launch();
      }
      batch_combine(ctx, L, Vp, n, k + 1, hdev.as<double>(), ld, -1.0, w.as<double>());
    }
    batch_dots(ctx, L, w.as<double>(), n, 1, w.as<double>(), partial.as<double>(),
               hdev.as<double>(), ld);
    if (np > 0) {
      k_norm_scale<<<gp, 256, 0, s>>>(np, ld, cyc.as<uint8_t>(), hdev.as<double>(),
                                      hk1.as<double>(), sc.as<double>());
      check_launch();
    }
    batch_scale(ctx, L, w.as<double>(), sc.as<double>(), vec(k + 1));
    if (np > 0) {
      k_givens<<<gp, 256, 0, s>>>(np, k, m, ld, cyc.as<uint8_t>(), htot.as<double>(),
                                  hk1.as<double>(), Hd.as<double>(), csd.as<double>(),
                                  snd.as<double>(), gd.as<double>(), kpd.as<int32_t>(),
                                  itd.as<int32_t>(), bnd.as<double>(), opt.tol, max_iters);
      check_launch();
    }
  };

  norms(b, bnorm);
  up(bnd.p, bnorm.data(), size_t(np) * 8);
  up(itd.p, iters.data(), size_t(np) * 4);
  residual(w.as<double>());
  norms(w.as<double>(), beta);
  std::vector<int> stalls(np, 0);
  if (n > 0) ABH_HIP(hipMemcpyAsync(xprev.p, x, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  for (int p = 0; p < np; ++p) {
    res.stats[p].residual = bnorm[p] > 0 ? beta[p] / bnorm[p] : 0.0;
    if (bnorm[p] == 0.0 || beta[p] <= opt.tol * bnorm[p]) {
      active[p] = 0;
      res.stats[p].converged = 1;
    }
  }
  // graph mode: the cycle's steps in segments of kSegment, each captured on
  // first use and replayed, the host checking between segments whether a
  // problem is still iterating
  const int nseg = (m + kSegment - 1) / kSegment;
  std::vector<hipGraph_t> gr(size_t(nseg), nullptr);
  std::vector<hipGraphExec_t> ge(size_t(nseg), nullptr);
  struct GraphFree {
    std::vector<hipGraph_t> &g;
    std::vector<hipGraphExec_t> &e;
    ~GraphFree() {
      for (hipGraphExec_t x : e)
        if (x) (void)hipGraphExecDestroy(x);
      for (hipGraph_t x : g)
        if (x) (void)hipGraphDestroy(x);
    }
  } graph_free{gr, ge};
  auto any_cycling = [&]() {
    ABH_HIP(hipMemcpyAsync(cyc_h.data(), cyc.p, size_t(np), hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    bool more = false;
    for (int p = 0; p < np; ++p) more |= cyc_h[p] != 0;
    return more;
  };
  for (;;) {
    bool any = false;
    for (int p = 0; p < np; ++p) any |= active[p] && iters[p] < max_iters;
    if (!any) break;
    up(act.p, active.data(), size_t(np));
    up(betad.p, beta.data(), size_t(np) * 8);
    k_cycle_start<<<gp, 256, 0, s>>>(np, m, act.as<uint8_t>(), itd.as<int32_t>(), max_iters,
                                     betad.as<double>(), cyc.as<uint8_t>(), sc.as<double>(),
                                     gd.as<double>(), kpd.as<int32_t>());
    check_launch();
    batch_scale(ctx, L, w.as<double>(), sc.as<double>(), vec(0));
    if (graph) {
      for (int q = 0; q < nseg; ++q) {
        const int k0 = q * kSegment, k1 = std::min(m, k0 + kSegment);
        if (!ge[size_t(q)]) {
          CaptureGuard cg{s};
          ABH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          cg.open = true;
          for (int k = k0; k < k1; ++k) step(k);
          cg.open = false;
          ABH_HIP(hipStreamEndCapture(s, &gr[size_t(q)]));
          ABH_HIP(hipGraphInstantiate(&ge[size_t(q)], gr[size_t(q)], nullptr, nullptr, 0));
        }
        ABH_HIP(hipGraphLaunch(ge[size_t(q)], s));
        if (!any_cycling()) break;
      }
    } else {
      for (int k = 0; k < m; ++k) {
        step(k);
        if (!any_cycling()) break;
      }
    }
    // x += M^-1 V y with H y = g (upper triangular, size kp)
    k_hsolve<<<gp, 256, 0, s>>>(np, m, ld, kpd.as<int32_t>(), Hd.as<double>(), gd.as<double>(),
                                hdev.as<double>());
    check_launch();
    ABH_HIP(hipMemcpyAsync(kp.data(), kpd.p, size_t(np) * 4, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipMemcpyAsync(iters.data(), itd.p, size_t(np) * 4, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    int kmax = 0;
    for (int p = 0; p < np; ++p) {
      kmax = std::max(kmax, kp[p]);
      res.stats[p].iterations = iters[p];
    }
    if (prec) {
      ABH_HIP(hipMemsetAsync(w.p, 0, size_t(n) * 8, s));
      batch_combine(ctx, L, Vp, n, kmax, hdev.as<double>(), ld, 1.0, w.as<double>());
      (*prec)(w.as<double>(), z.as<double>());
      if (n > 0) {
        k_add<<<gn, 256, 0, s>>>(n, z.as<double>(), x);
        check_launch();
      }
    } else {
      batch_combine(ctx, L, Vp, n, kmax, hdev.as<double>(), ld, 1.0, x);
    }
    // true residual in FP64
    const std::vector<double> beta_prev = beta;
    residual(w.as<double>());
    norms(w.as<double>(), beta);
    // a problem whose true residual grew over the cycle (rounding has taken
    // over) goes back to its previous iterate and stops
    std::vector<double> bk(np, 0.0);
    bool any_back = false;
    for (int p = 0; p < np; ++p) {
      if (!active[p]) {
        beta[p] = beta_prev[p];  // not recomputed
        continue;
      }
      if (beta[p] > beta_prev[p]) {
        bk[p] = 1.0;
        any_back = true;
        beta[p] = beta_prev[p];
        active[p] = 0;
        continue;
      }
      res.stats[p].residual = bnorm[p] > 0 ? beta[p] / bnorm[p] : 0.0;
      if (beta[p] <= opt.tol * bnorm[p]) {
        res.stats[p].converged = 1;
        active[p] = 0;
      } else if (opt.stall_restarts > 0) {
        stalls[p] = beta[p] > 0.9 * beta_prev[p] ? stalls[p] + 1 : 0;
        if (stalls[p] >= opt.stall_restarts) active[p] = 0;  // stagnated
      }
    }
    if (any_back && L.nchunks > 0) {
      up(back.p, bk.data(), size_t(np) * 8);
      k_restore<<<L.nchunks, kThreads, 0, s>>>(L.chunks.as<Chunk>(), back.as<double>(),
                                               xprev.as<double>(), x);
      check_launch();
    }
    if (n > 0) ABH_HIP(hipMemcpyAsync(xprev.p, x, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  }
  ABH_HIP(hipStreamSynchronize(s));
}

}  // namespace abh
