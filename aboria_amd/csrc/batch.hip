// Batched vector kernels and the lockstep restarted GMRES(m) engine.
#include <algorithm>
#include <cmath>

#include "batch.hpp"

namespace abh {

namespace {

constexpr int kChunk = 4096;
constexpr int kThreads = 256;

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

void batch_gmres(abh_context *ctx, const BatchLayout &L,
                 const std::function<void(const double *, double *)> &op,
                 const std::function<void(const double *, double *)> &op_true,
                 const std::function<void(const double *, double *)> *prec,
                 const std::function<void(const uint8_t *)> &set_active, const double *b,
                 double *x, const abh_gmres_options &opt, GmresResult &res) {
  hipStream_t s = ctx->stream;
  // small host -> device copies wait for completion: the host vectors they
  // read are reused right after
  auto up = [&](void *dst, const void *src, size_t bytes) {
    ABH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    ABH_HIP(hipStreamSynchronize(s));
  };
  const int np = L.nprob;
  const int64_t n = L.total();
  const int m = std::max(1, opt.restart);
  const int ld = m + 2;
  DevBuf V, w, z, partial, hdev, sdev, xprev;
  V.ensure(size_t(m + 1) * size_t(n) * 8 + 64);
  xprev.ensure(size_t(n) * 8 + 64);
  w.ensure(size_t(n) * 8 + 64);
  z.ensure(size_t(n) * 8 + 64);
  partial.ensure(size_t(std::max(L.nchunks, 1)) * ld * 8);
  hdev.ensure(size_t(np) * ld * 8 + 64);
  sdev.ensure(size_t(np) * 8 + 64);
  double *Vp = V.as<double>();
  auto vec = [&](int i) { return Vp + int64_t(i) * n; };
  std::vector<double> hh(size_t(np) * ld), bnorm(np), beta(np);
  res.stats.assign(np, abh_gmres_stats{0, 0, 0.0});
  std::vector<uint8_t> active(np, 1), cyc(np, 0);
  // per problem Hessenberg (column-major (m+1) x m), Givens, rhs g
  std::vector<double> H(size_t(np) * (m + 1) * m), cs(size_t(np) * m), sn(size_t(np) * m),
      g(size_t(np) * (m + 1));
  std::vector<int> kp(np, 0);

  auto norms = [&](const double *v, std::vector<double> &out) {
    batch_dots(ctx, L, v, n, 1, v, partial.as<double>(), hdev.as<double>(), ld);
    ABH_HIP(hipMemcpyAsync(hh.data(), hdev.p, size_t(np) * ld * 8, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < np; ++p) out[p] = std::sqrt(std::max(0.0, hh[size_t(p) * ld]));
  };
  // r = b - A x (FP64) for the problems still iterating (the rows of the
  // others are left unspecified)
  auto residual = [&](double *r) {
    set_active(active.data());
    op_true(x, r);
    std::vector<double> neg(np, -1.0);
    up(sdev.p, neg.data(), size_t(np) * 8);
    batch_scale(ctx, L, r, sdev.as<double>(), r);
    std::vector<double> one(size_t(np) * ld, 0.0);
    for (int p = 0; p < np; ++p) one[size_t(p) * ld] = 1.0;
    up(hdev.p, one.data(), one.size() * 8);
    batch_combine(ctx, L, b, 0, 1, hdev.as<double>(), ld, 1.0, r);
  };

  norms(b, bnorm);
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
  const int max_iters = std::max(1, opt.max_iters);
  for (;;) {
    bool any = false;
    for (int p = 0; p < np; ++p) any |= active[p] && res.stats[p].iterations < max_iters;
    if (!any) break;
    // v0 = r / beta
    std::vector<double> sc(np, 0.0);
    for (int p = 0; p < np; ++p) {
      cyc[p] = active[p] && res.stats[p].iterations < max_iters;
      sc[p] = cyc[p] ? 1.0 / beta[p] : 0.0;
      kp[p] = 0;
      std::fill(g.begin() + size_t(p) * (m + 1), g.begin() + size_t(p + 1) * (m + 1), 0.0);
      g[size_t(p) * (m + 1)] = beta[p];
    }
    up(sdev.p, sc.data(), size_t(np) * 8);
    batch_scale(ctx, L, w.as<double>(), sdev.as<double>(), vec(0));
    for (int k = 0; k < m; ++k) {
      set_active(cyc.data());
      const double *zin = vec(k);
      if (prec) {
        (*prec)(vec(k), z.as<double>());
        zin = z.as<double>();
      }
      op(zin, w.as<double>());
      // CGS2 against v_0..v_k
      std::vector<double> htot(size_t(np) * (k + 1), 0.0);
      for (int pass = 0; pass < 2; ++pass) {
        batch_dots(ctx, L, Vp, n, k + 1, w.as<double>(), partial.as<double>(), hdev.as<double>(),
                   ld);
        ABH_HIP(hipMemcpyAsync(hh.data(), hdev.p, size_t(np) * ld * 8, hipMemcpyDeviceToHost, s));
        ABH_HIP(hipStreamSynchronize(s));
        for (int p = 0; p < np; ++p)
          for (int i = 0; i <= k; ++i) {
            double &v = hh[size_t(p) * ld + i];
            if (!cyc[p]) v = 0.0;
            htot[size_t(p) * (k + 1) + i] += v;
          }
        up(hdev.p, hh.data(), size_t(np) * ld * 8);
        batch_combine(ctx, L, Vp, n, k + 1, hdev.as<double>(), ld, -1.0, w.as<double>());
      }
      std::vector<double> hk1(np);
      norms(w.as<double>(), hk1);
      for (int p = 0; p < np; ++p) sc[p] = (cyc[p] && hk1[p] > 0.0) ? 1.0 / hk1[p] : 0.0;
      up(sdev.p, sc.data(), size_t(np) * 8);
      batch_scale(ctx, L, w.as<double>(), sdev.as<double>(), vec(k + 1));
      bool more = false;
      for (int p = 0; p < np; ++p) {
        if (!cyc[p]) continue;
        double *Hc = &H[(size_t(p) * m + k) * (m + 1)];
        for (int i = 0; i <= k; ++i) Hc[i] = htot[size_t(p) * (k + 1) + i];
        Hc[k + 1] = hk1[p];
        double *c = &cs[size_t(p) * m], *sg = &sn[size_t(p) * m];
        for (int i = 0; i < k; ++i) {
          const double t = c[i] * Hc[i] + sg[i] * Hc[i + 1];
          Hc[i + 1] = -sg[i] * Hc[i] + c[i] * Hc[i + 1];
          Hc[i] = t;
        }
        const double den = std::hypot(Hc[k], Hc[k + 1]);
        c[k] = den > 0 ? Hc[k] / den : 1.0;
        sg[k] = den > 0 ? Hc[k + 1] / den : 0.0;
        Hc[k] = den;
        Hc[k + 1] = 0.0;
        double *gp = &g[size_t(p) * (m + 1)];
        gp[k + 1] = -sg[k] * gp[k];
        gp[k] = c[k] * gp[k];
        kp[p] = k + 1;
        res.stats[p].iterations += 1;
        const double est = std::fabs(gp[k + 1]);
        if (est <= opt.tol * bnorm[p] || hk1[p] == 0.0 || res.stats[p].iterations >= max_iters)
          cyc[p] = 0;
        more |= cyc[p] != 0;
      }
      if (!more) break;
    }
    // x += M^-1 V y with H y = g (upper triangular, size kp)
    std::vector<double> y(size_t(np) * ld, 0.0);
    int kmax = 0;
    for (int p = 0; p < np; ++p) {
      const int kk = kp[p];
      kmax = std::max(kmax, kk);
      for (int i = kk - 1; i >= 0; --i) {
        double t = g[size_t(p) * (m + 1) + i];
        for (int j = i + 1; j < kk; ++j) t -= H[(size_t(p) * m + j) * (m + 1) + i] * y[size_t(p) * ld + j];
        const double d = H[(size_t(p) * m + i) * (m + 1) + i];
        y[size_t(p) * ld + i] = d != 0.0 ? t / d : 0.0;
      }
    }
    up(hdev.p, y.data(), y.size() * 8);
    if (prec) {
      ABH_HIP(hipMemsetAsync(w.p, 0, size_t(n) * 8, s));
      batch_combine(ctx, L, Vp, n, kmax, hdev.as<double>(), ld, 1.0, w.as<double>());
      (*prec)(w.as<double>(), z.as<double>());
      std::vector<double> one(size_t(np) * ld, 0.0);
      for (int p = 0; p < np; ++p) one[size_t(p) * ld] = 1.0;
      up(hdev.p, one.data(), one.size() * 8);
      batch_combine(ctx, L, z.as<double>(), 0, 1, hdev.as<double>(), ld, 1.0, x);
    } else {
      batch_combine(ctx, L, Vp, n, kmax, hdev.as<double>(), ld, 1.0, x);
    }
    // true residual in FP64
    const std::vector<double> beta_prev = beta;
    residual(w.as<double>());
    norms(w.as<double>(), beta);
    // a problem whose true residual grew over the cycle (rounding has taken
    // over) goes back to its previous iterate and stops
    std::vector<double> back(np, 0.0);
    bool any_back = false;
    for (int p = 0; p < np; ++p) {
      if (!active[p]) {
        beta[p] = beta_prev[p];  // not recomputed
        continue;
      }
      if (beta[p] > beta_prev[p]) {
        back[p] = 1.0;
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
    if (any_back) {  // x += back_p (xprev - x); z is free here
      ABH_HIP(hipMemcpyAsync(z.p, xprev.p, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
      std::vector<double> one(size_t(np) * ld, 0.0);
      for (int p = 0; p < np; ++p) one[size_t(p) * ld] = 1.0;
      up(hdev.p, one.data(), one.size() * 8);
      batch_combine(ctx, L, x, 0, 1, hdev.as<double>(), ld, -1.0, z.as<double>());
      up(sdev.p, back.data(), size_t(np) * 8);
      batch_scale(ctx, L, z.as<double>(), sdev.as<double>(), z.as<double>());
      batch_combine(ctx, L, z.as<double>(), 0, 1, hdev.as<double>(), ld, 1.0, x);
    }
    if (n > 0) ABH_HIP(hipMemcpyAsync(xprev.p, x, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  }
  ABH_HIP(hipStreamSynchronize(s));
}

}  // namespace abh
