// Restricted additive Schwarz preconditioner over a cell list.
//
// Analyse (host, O(3^dim N)): one domain per non-empty bucket; particles of
// the 3^dim block of buckets around it inside the bucket bounds are the
// interior, the rest the buffer (src/Preconditioners.h:1169-1230).
//
// Factorise (device): domains are sorted by size and processed in groups of
// equal padded size n (a multiple of 64).  Per group every local matrix
// A_g = K + shift I (identity in the padding) is assembled in HBM,
// factorised by a right-looking blocked Cholesky (64-wide panels: the
// diagonal block in LDS, panel and trailing updates as strided-batched
// FP64 MFMA GEMMs, dgemm.hip), and the interior columns of A_g^-1 (= the
// interior rows, A_g is symmetric) are obtained by a blocked forward and
// backward substitution on the unit columns.  They stay resident: R_g is
// n x ni column-major, so row a of the local inverse is the contiguous
// column a.
//
// For the dense batch (every problem an independent particle set) the
// domains come from a per-problem bucket grid instead of a cell list, the
// kernel is not truncated, and the factorisation is signed, A = L S L^T with
// S = diag(+-1), since multiquadric matrices are indefinite.
//
// Apply (device, HBM-bound): x = b, then for every domain and interior row a
// x[idx[a]] = <R_g[:, a], b[idx]>, one wavefront per row, the gathered
// b[idx] staged once per workgroup in LDS.  A domain whose factorisation
// broke down (counted in failed_domains) is skipped: its rows keep b.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <vector>

#include "cells.hpp"
#include "kernel_fn.hpp"
#include "schwarz_analysis.hpp"
#include "solvers.hpp"

using abh::DevBuf;
using abh::KernelDesc;

namespace {

constexpr int kNb = 64;            // Cholesky panel width
constexpr int kApplyRows = 64;     // interior rows per apply workgroup
constexpr int kApplyLds = 7936;    // max domain size staged in LDS by apply (62 KiB)

struct DomDev {
  int64_t idx_off;  // first entry in the index array
  int64_t r_off;    // first double of the domain's inverse rows
  int32_t m;        // interior + buffer
  int32_t ni;       // interior
  int32_t ld;       // padded size n of its group (leading dimension of R)
  int32_t failed;   // factorisation broke down: apply leaves its rows at b
};

// A[g] (n x n, column-major) = K + shift I over the domain's particles,
// identity on the padding.
template <int D>
__global__ __launch_bounds__(256) void k_assemble(const DomDev *__restrict__ doms,
                                                  const int32_t *__restrict__ idx,
                                                  const double *__restrict__ pos, KernelDesc k,
                                                  double r2, double shift, int n,
                                                  double *__restrict__ A) {
  __shared__ double pi[64][D], pj[64][D];
  const int g = blockIdx.z;
  const DomDev dm = doms[g];
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  const int t = threadIdx.x;
  if (t < 64) {
    const int i = i0 + t;
#pragma unroll
    for (int d = 0; d < D; ++d) pi[t][d] = i < dm.m ? pos[int64_t(idx[dm.idx_off + i]) * D + d] : 0.0;
  } else if (t < 128) {
    const int j = j0 + t - 64;
#pragma unroll
    for (int d = 0; d < D; ++d)
      pj[t - 64][d] = j < dm.m ? pos[int64_t(idx[dm.idx_off + j]) * D + d] : 0.0;
  }
  __syncthreads();
  double *Ag = A + int64_t(g) * n * n;
  const int li = t & 63;
  const int i = i0 + li;
  for (int lj = t >> 6; lj < 64; lj += 4) {
    const int j = j0 + lj;
    double v;
    if (i < dm.m && j < dm.m) {
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double dx = pj[lj][d] - pi[li][d];
        s = s + dx * dx;
      }
      v = (r2 < 0.0 || s <= r2) ? abh::phi64(k, s) : 0.0;
      if (i == j) v = v + shift;
    } else {
      v = i == j ? 1.0 : 0.0;
    }
    Ag[int64_t(j) * n + i] = v;
  }
}

// Cholesky of the 64 x 64 diagonal block at (o, o) of every matrix of the
// group, in LDS; writes L back and L^-1 to dinv (64 x 64 column-major).
// Signed mode (symmetric matrices with negative eigenvalues, such as
// multiquadric ones): A = L S L^T with S = diag(+-1); the signs go to
// sg[g * sg_stride + o + j] and S L^-1 (the panel's right factor) to dinvs.
__global__ __launch_bounds__(256) void k_potrf_diag(double *__restrict__ A, int n, int o,
                                                    double *__restrict__ dinv,
                                                    int64_t dinv_stride,
                                                    uint8_t *__restrict__ failed, int signed_mode,
                                                    double *__restrict__ sg, int64_t sg_stride,
                                                    double *__restrict__ dinvs) {
  __shared__ double L[kNb][kNb + 1];
  __shared__ double X[kNb][kNb + 1];
  __shared__ double S[kNb];
  __shared__ int bad;
  const int g = blockIdx.x;
  const int t = threadIdx.x;
  double *Ag = A + int64_t(g) * n * n + int64_t(o) * n + o;
  for (int e = t; e < kNb * kNb; e += 256) {
    const int r = e & 63, c = e >> 6;
    L[r][c] = Ag[int64_t(c) * n + r];
  }
  if (t == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < kNb; ++j) {
    if (t == 0) {
      double d = L[j][j];
      double sj = 1.0;
      if (signed_mode) {
        if (!(d != 0.0) || !isfinite(d)) {
          bad = 1;
          d = 1.0;
        }
        sj = d > 0.0 ? 1.0 : -1.0;
        d = fabs(d);
      } else if (!(d > 0.0)) {
        bad = 1;
        d = 1.0;
      }
      S[j] = sj;
      L[j][j] = sqrt(d);
    }
    __syncthreads();
    const double djj = L[j][j] * S[j];
    if (t > j && t < kNb) L[t][j] = L[t][j] / djj;
    __syncthreads();
    // trailing update of the lower triangle, columns j+1..63
    const double sj = S[j];
    for (int e = t; e < kNb * kNb; e += 256) {
      const int r = e & 63, c = e >> 6;
      if (c > j && r >= c) L[r][c] = L[r][c] - (L[r][j] * L[c][j]) * sj;
    }
    __syncthreads();
  }
  // X = L^-1, one column per thread (forward substitution)
  if (t < kNb) {
    const int c = t;
    for (int r = 0; r < c; ++r) X[r][c] = 0.0;
    X[c][c] = 1.0 / L[c][c];
    for (int r = c + 1; r < kNb; ++r) {
      double s = 0.0;
      for (int q = c; q < r; ++q) s = s + L[r][q] * X[q][c];
      X[r][c] = -s / L[r][r];
    }
  }
  __syncthreads();
  double *Dg = dinv + int64_t(g) * dinv_stride;
  for (int e = t; e < kNb * kNb; e += 256) {
    const int r = e & 63, c = e >> 6;
    Ag[int64_t(c) * n + r] = r >= c ? L[r][c] : 0.0;
    Dg[c * kNb + r] = X[r][c];
    if (signed_mode) dinvs[int64_t(g) * dinv_stride + c * kNb + r] = X[r][c] * S[r];
  }
  if (signed_mode && t < kNb) sg[int64_t(g) * sg_stride + o + t] = S[t];
  if (t == 0 && bad) failed[g] = 1;
}

// dst[g] (rows x cols, ld ldd) = src[g] (ld lds) times cs[g * cs_stride + c]
// per column when cs is given; batched strided copy
__global__ void k_copy_block(const double *__restrict__ src, int64_t src_stride, int lds,
                             double *__restrict__ dst, int64_t dst_stride, int ldd, int rows,
                             int cols, const double *__restrict__ cs = nullptr,
                             int64_t cs_stride = 0) {
  const int g = blockIdx.y;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(rows) * cols) return;
  const int r = int(e % rows), c = int(e / rows);
  double v = src[g * src_stride + int64_t(c) * lds + r];
  if (cs) v = v * cs[g * cs_stride + c];
  dst[g * dst_stride + int64_t(c) * ldd + r] = v;
}

// Y[g][r][c] *= sg[g * sg_stride + r] (n x ni, column-major)
__global__ void k_scale_rows(double *__restrict__ Y, int64_t stride, int n, int ni,
                             const double *__restrict__ sg, int64_t sg_stride) {
  const int g = blockIdx.y;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(n) * ni) return;
  Y[g * stride + e] = Y[g * stride + e] * sg[g * sg_stride + int(e % n)];
}

// Y[g] (n x ni) = the first ni columns of the identity
__global__ void k_unit_cols(double *__restrict__ Y, int64_t stride, int n, int ni) {
  const int g = blockIdx.y;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(n) * ni) return;
  const int r = int(e % n), c = int(e / n);
  Y[g * stride + e] = r == c ? 1.0 : 0.0;
}

struct ApplyChunk {
  int32_t dom;
  int32_t row0;
};

template <bool LDS>
__global__ __launch_bounds__(256) void k_ras_apply(const ApplyChunk *__restrict__ chunks,
                                                   const DomDev *__restrict__ doms,
                                                   const int32_t *__restrict__ idx,
                                                   const double *__restrict__ R,
                                                   const double *__restrict__ b,
                                                   double *__restrict__ x) {
  extern __shared__ double bl[];
  const ApplyChunk ch = chunks[blockIdx.x];
  const DomDev dm = doms[ch.dom];
  if (dm.failed) return;  // uniform over the workgroup
  const int32_t *id = idx + dm.idx_off;
  if (LDS) {
    for (int j = threadIdx.x; j < dm.m; j += 256) bl[j] = b[id[j]];
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rows = min(kApplyRows, dm.ni - ch.row0);
  for (int a = wave; a < rows; a += 4) {
    const double *Ra = R + dm.r_off + int64_t(ch.row0 + a) * dm.ld;
    double s = 0.0;
    int j = lane;
    // four row loads in flight per step, summed in the same order
    for (; j + 192 < dm.m; j += 256) {
      const double r0 = Ra[j], r1 = Ra[j + 64], r2 = Ra[j + 128], r3 = Ra[j + 192];
      const double b0 = LDS ? bl[j] : b[id[j]];
      const double b1 = LDS ? bl[j + 64] : b[id[j + 64]];
      const double b2 = LDS ? bl[j + 128] : b[id[j + 128]];
      const double b3 = LDS ? bl[j + 192] : b[id[j + 192]];
      s = s + r0 * b0;
      s = s + r1 * b1;
      s = s + r2 * b2;
      s = s + r3 * b3;
    }
    for (; j < dm.m; j += 64) s = s + Ra[j] * (LDS ? bl[j] : b[id[j]]);
    s = abh::wave_sum(s);
    if (lane == 0) x[id[ch.row0 + a]] = s;
  }
}

}  // namespace

struct abh_schwarz {
  abh_context *ctx = nullptr;
  int dim = 2;
  int64_t n = 0;
  std::vector<int64_t> off;   // host domain offsets (n_domains + 1)
  std::vector<int32_t> idx;   // host domain lists
  std::vector<int32_t> nint;  // host interior counts
  int64_t failed = 0;
  int max_m = 0;
  DevBuf d_idx, d_doms, d_chunks, R;
  int nchunks = 0;
  size_t factor_bytes = 0;
  bool batch = false;  // built per problem of a dense batch (signed factors)
  // the coarse-grid domain (set_coarse_grid_n): one domain of sampled
  // particles, all interior, factorised like the others and applied after
  // them, its rows overwriting theirs
  abh_schwarz *coarse = nullptr;
  ~abh_schwarz() { delete coarse; }
};

namespace {

void take(abh_schwarz *s, abh::SchwarzDomains &dm) {
  s->off.swap(dm.off);
  s->idx.swap(dm.idx);
  s->nint.swap(dm.nint);
}

abh::GridView grid_view(const abh_cells *c) {
  abh::GridView g;
  g.dim = c->dim;
  for (int d = 0; d < c->dim; ++d) {
    g.size[d] = c->size[d];
    g.side[d] = c->side[d];
    g.lo[d] = c->lo[d];
    g.hi[d] = c->hi[d];
    g.inv[d] = c->inv[d];
  }
  return g;
}

void analyse_into(abh_schwarz *s, const abh_cells *c, const std::vector<double> &pos,
             const std::vector<uint32_t> &beg, const std::vector<uint32_t> &end) {
  const abh::GridView g = grid_view(c);
  abh::SchwarzDomains dm;
  const char *err = abh::analyse_cells(g, pos, beg, end, dm);
  ABH_REQUIRE(!err, ABH_ERR_INVALID_ARG, err ? err : "");
  take(s, dm);
}

// the coarse-grid domain of n x ... x n lattice cells as a one-domain
// preconditioner over the same particles
void coarse_into(abh_schwarz *s, const abh_cells *c, const std::vector<uint32_t> &beg,
                 const std::vector<uint32_t> &end, int n, uint64_t seed) {
  std::vector<int32_t> idx;
  const char *err = abh::coarse_grid(grid_view(c), beg, end, n, seed, idx);
  ABH_REQUIRE(!err, ABH_ERR_INVALID_ARG, err ? err : "");
  std::unique_ptr<abh_schwarz> q(new abh_schwarz());
  q->ctx = s->ctx;
  q->dim = s->dim;
  q->n = s->n;
  q->off = {0, int64_t(idx.size())};
  q->nint = {int32_t(idx.size())};
  q->max_m = int(idx.size());
  q->idx.swap(idx);
  delete s->coarse;
  s->coarse = q.release();
}

void analyse_batch_into(abh_schwarz *s, int D, const std::vector<int64_t> &boff,
                   const std::vector<double> &pos, const double *lo, const double *hi,
                   double n_leaf) {
  abh::SchwarzDomains dm;
  const char *err = abh::analyse_batch(D, boff, pos, lo, hi, n_leaf, dm);
  ABH_REQUIRE(!err, ABH_ERR_INVALID_ARG, err ? err : "");
  take(s, dm);
}

// sgn: local matrices may be indefinite; factorise A = L S L^T (S = +-1)
// instead of requiring positive pivots.
void factorise(abh_schwarz *s, const double *pos, const KernelDesc &k, double r2, double shift,
               bool sgn) {
  abh_context *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int nd = int(s->nint.size());
  if (nd == 0) return;
  // domains by padded size, then interior count, largest first; a group
  // shares the padded size and the interior count of its first member
  std::vector<int> order(nd);
  std::iota(order.begin(), order.end(), 0);
  auto msize = [&](int d) { return int(s->off[d + 1] - s->off[d]); };
  auto npad = [&](int d) { return int(abh::ceil_div(msize(d), kNb) * kNb); };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (npad(a) != npad(b)) return npad(a) > npad(b);
    return s->nint[a] > s->nint[b];
  });
  size_t free_b = 0, total_b = 0;
  ABH_HIP(hipMemGetInfo(&free_b, &total_b));
  // host staging of the asynchronous copies below; the drain is destroyed
  // before them, so an error thrown while a copy is pending cannot free its
  // source or destination under it
  std::vector<DomDev> dd(nd), dg(nd);
  std::vector<ApplyChunk> ch;
  std::vector<uint8_t> fh(nd);
  abh::StreamDrain drain{st};
  // inverse rows: n_pad x ni_pad per domain
  struct Group {
    int first, count, n, ni;
  };
  std::vector<Group> groups;
  int64_t r_total = 0;
  {
    int i = 0;
    while (i < nd) {
      const int n = npad(order[i]);
      // working set per member: A (n^2) + panel copy (64 n) + dinv (n 64)
      // (+ S L^-1 and the signs in signed mode)
      const double per = double(n) * n * 8.0 + double(n) * kNb * (sgn ? 24.0 : 16.0) +
                         (sgn ? double(n) * 8.0 : 0.0);
      const double budget = std::min(8.0e9, 0.35 * double(free_b));
      int gmax = int(std::max(1.0, std::min(256.0, budget / per)));
      Group gr{i, 0, n, 0};
      gr.ni = s->nint[order[i]];
      while (i < nd && gr.count < gmax && npad(order[i]) == n) {
        ++gr.count;
        ++i;
      }
      for (int q = 0; q < gr.count; ++q) {
        const int d = order[gr.first + q];
        dd[d].idx_off = s->off[d];
        dd[d].r_off = r_total + int64_t(q) * n * gr.ni;
        dd[d].m = msize(d);
        dd[d].ni = s->nint[d];
        dd[d].ld = n;
        dd[d].failed = 0;
      }
      r_total += int64_t(gr.count) * n * gr.ni;
      groups.push_back(gr);
    }
  }
  s->factor_bytes = size_t(r_total) * 8;
  s->R.ensure(s->factor_bytes + 64);
  s->d_idx.ensure(s->idx.size() * 4 + 64);
  ABH_HIP(hipMemcpyAsync(s->d_idx.p, s->idx.data(), s->idx.size() * 4, hipMemcpyHostToDevice, st));
  // domain table in group order (the group's members are contiguous)
  for (int i = 0; i < nd; ++i) dg[i] = dd[order[i]];
  DevBuf d_dg;
  d_dg.ensure(size_t(nd) * sizeof(DomDev) + 64);
  ABH_HIP(hipMemcpyAsync(d_dg.p, dg.data(), size_t(nd) * sizeof(DomDev), hipMemcpyHostToDevice,
                         st));
  // apply table in domain order
  s->d_doms.ensure(size_t(nd) * sizeof(DomDev) + 64);
  ABH_HIP(hipMemcpyAsync(s->d_doms.p, dd.data(), size_t(nd) * sizeof(DomDev),
                         hipMemcpyHostToDevice, st));
  for (int d = 0; d < nd; ++d)
    for (int r = 0; r < s->nint[d]; r += kApplyRows) ch.push_back(ApplyChunk{d, r});
  s->nchunks = int(ch.size());
  s->d_chunks.ensure(ch.size() * sizeof(ApplyChunk) + 64);
  ABH_HIP(hipMemcpyAsync(s->d_chunks.p, ch.data(), ch.size() * sizeof(ApplyChunk),
                         hipMemcpyHostToDevice, st));
  using abh::dgemm_strided_batched;

  int gmax = 0, nmax = 0, nimax = 0;
  for (const Group &gr : groups) {
    gmax = std::max(gmax, gr.count);
    nmax = std::max(nmax, gr.n);
    nimax = std::max(nimax, gr.ni);
  }
  DevBuf A, W, Tm, Dinv, DinvS, Sg, fail;
  A.ensure(size_t(gmax) * nmax * nmax * 8);
  W.ensure(size_t(gmax) * nmax * kNb * 8);
  Tm.ensure(size_t(gmax) * kNb * nimax * 8);
  Dinv.ensure(size_t(gmax) * (nmax / kNb) * kNb * kNb * 8);
  if (sgn) {
    DinvS.ensure(size_t(gmax) * (nmax / kNb) * kNb * kNb * 8);
    Sg.ensure(size_t(gmax) * nmax * 8);
  }
  fail.ensure(size_t(nd) + 64);
  ABH_HIP(hipMemsetAsync(fail.p, 0, size_t(nd), st));
  const int32_t *didx = s->d_idx.as<int32_t>();

  for (const Group &gr : groups) {
    const int G = gr.count, n = gr.n, ni = gr.ni, K = n / kNb;
    const int64_t sA = int64_t(n) * n;
    const int64_t sD = int64_t(K) * kNb * kNb;
    const int64_t sW = int64_t(n) * kNb;
    const int64_t sT = int64_t(kNb) * ni;
    const int64_t sY = int64_t(n) * ni;
    double *Ap = A.as<double>(), *Wp = W.as<double>(), *Tp = Tm.as<double>();
    double *Dp = Dinv.as<double>();
    // signed mode: panel right factor S L^-1, signs per member (stride n)
    double *DSp = sgn ? DinvS.as<double>() : Dp;
    double *Sp = sgn ? Sg.as<double>() : nullptr;
    double *Y = s->R.as<double>() + dd[order[gr.first]].r_off;
    uint8_t *fl = fail.as<uint8_t>() + gr.first;
    const DomDev *gd = d_dg.as<DomDev>() + gr.first;
    {
      dim3 grid(n / 64, n / 64, G);
      if (s->dim == 1) k_assemble<1><<<grid, 256, 0, st>>>(gd, didx, pos, k, r2, shift, n, Ap);
      else if (s->dim == 2) k_assemble<2><<<grid, 256, 0, st>>>(gd, didx, pos, k, r2, shift, n, Ap);
      else k_assemble<3><<<grid, 256, 0, st>>>(gd, didx, pos, k, r2, shift, n, Ap);
      abh::check_launch();
    }
    // blocked Cholesky, lower (signed: A = L S L^T)
    for (int kb = 0; kb < K; ++kb) {
      const int o = kb * kNb;
      k_potrf_diag<<<G, 256, 0, st>>>(Ap, n, o, Dp + int64_t(kb) * kNb * kNb, sD, fl,
                                      sgn ? 1 : 0, Sp, n, DSp + int64_t(kb) * kNb * kNb);
      abh::check_launch();
      const int r = n - o - kNb;
      if (r == 0) break;
      double *P = Ap + o + kNb + int64_t(o) * n;
      // panel P = A[o+64:, o:o+64] L_kk^-T (S_k)
      {
        dim3 grid(unsigned(abh::ceil_div(int64_t(r) * kNb, 256)), G);
        k_copy_block<<<grid, 256, 0, st>>>(P, sA, n, Wp, sW, n, r, kNb);
        abh::check_launch();
      }
      dgemm_strided_batched(st, false, true, r, kNb, kNb, 1.0, Wp, n, sW,
                            DSp + int64_t(kb) * kNb * kNb, kNb, sD, 0.0, P, n, sA, G);
      // trailing A[o+64:, o+64:] -= P (S_k) P^T, lower triangle (the upper
      // one is never read again)
      const double *Pl = P;
      if (sgn) {
        dim3 grid(unsigned(abh::ceil_div(int64_t(r) * kNb, 256)), G);
        k_copy_block<<<grid, 256, 0, st>>>(P, sA, n, Wp, sW, n, r, kNb, Sp + o, n);
        abh::check_launch();
        Pl = Wp;
      }
      dgemm_strided_batched(st, false, true, r, r, kNb, -1.0, Pl, n, sgn ? sW : sA, P, n, sA,
                            1.0, Ap + o + kNb + int64_t(o + kNb) * n, n, sA, G, true);
    }
    // Y = L^-1 E (E = first ni unit columns); Y is lower trapezoidal
    {
      dim3 grid(unsigned(abh::ceil_div(int64_t(n) * ni, 256)), G);
      k_unit_cols<<<grid, 256, 0, st>>>(Y, sY, n, ni);
      abh::check_launch();
    }
    for (int kb = 0; kb < K; ++kb) {
      const int o = kb * kNb;
      const int nc = std::min(ni, o + kNb);  // columns >= o+64 are still zero
      {
        dim3 grid(unsigned(abh::ceil_div(int64_t(kNb) * nc, 256)), G);
        k_copy_block<<<grid, 256, 0, st>>>(Y + o, sY, n, Tp, sT, kNb, kNb, nc);
        abh::check_launch();
      }
      dgemm_strided_batched(st, false, false, kNb, nc, kNb, 1.0, Dp + int64_t(kb) * kNb * kNb,
                            kNb, sD, Tp, kNb, sT, 0.0, Y + o, n, sY, G);
      const int r = n - o - kNb;
      if (r > 0)
        dgemm_strided_batched(st, false, false, r, nc, kNb, -1.0, Ap + o + kNb + int64_t(o) * n,
                              n, sA, Y + o, n, sY, 1.0, Y + o + kNb, n, sY, G);
    }
    // Z = L^-T (S) Y, in place
    if (sgn) {
      dim3 grid(unsigned(abh::ceil_div(int64_t(n) * ni, 256)), G);
      k_scale_rows<<<grid, 256, 0, st>>>(Y, sY, n, ni, Sp, n);
      abh::check_launch();
    }
    for (int kb = K - 1; kb >= 0; --kb) {
      const int o = kb * kNb;
      const int r = n - o - kNb;
      if (r > 0)
        dgemm_strided_batched(st, true, false, kNb, ni, r, -1.0, Ap + o + kNb + int64_t(o) * n,
                              n, sA, Y + o + kNb, n, sY, 1.0, Y + o, n, sY, G);
      {
        dim3 grid(unsigned(abh::ceil_div(int64_t(kNb) * ni, 256)), G);
        k_copy_block<<<grid, 256, 0, st>>>(Y + o, sY, n, Tp, sT, kNb, kNb, ni);
        abh::check_launch();
      }
      dgemm_strided_batched(st, true, false, kNb, ni, kNb, 1.0, Dp + int64_t(kb) * kNb * kNb,
                            kNb, sD, Tp, kNb, sT, 0.0, Y + o, n, sY, G);
    }
  }
  ABH_HIP(hipMemcpyAsync(fh.data(), fail.p, size_t(nd), hipMemcpyDeviceToHost, st));
  ABH_HIP(hipStreamSynchronize(st));
  s->failed = std::accumulate(fh.begin(), fh.end(), int64_t(0));
  if (s->failed) {
    // a broken-down domain contributes nothing (its interior rows keep b)
    // rather than non-finite values
    for (int i = 0; i < nd; ++i) dd[order[i]].failed = fh[i] ? 1 : 0;
    ABH_HIP(hipMemcpyAsync(s->d_doms.p, dd.data(), size_t(nd) * sizeof(DomDev),
                           hipMemcpyHostToDevice, st));
    ABH_HIP(hipStreamSynchronize(st));
  }
}

}  // namespace

namespace abh {

// every domain's interior rows x[i] = (A_dom^-1 b_dom)_i
void apply_domains(const abh_schwarz *s, const double *b, double *x) {
  hipStream_t st = s->ctx->stream;
  if (s->nchunks == 0) return;
  const ApplyChunk *ch = s->d_chunks.as<ApplyChunk>();
  const DomDev *dm = s->d_doms.as<DomDev>();
  const int32_t *id = s->d_idx.as<int32_t>();
  const double *R = s->R.as<double>();
  if (s->max_m <= kApplyLds)
    k_ras_apply<true><<<s->nchunks, 256, size_t(s->max_m) * 8, st>>>(ch, dm, id, R, b, x);
  else
    k_ras_apply<false><<<s->nchunks, 256, 0, st>>>(ch, dm, id, R, b, x);
  check_launch();
}

void schwarz_apply(const abh_schwarz *s, const double *b, double *x) {
  hipStream_t st = s->ctx->stream;
  if (s->n == 0) return;
  ABH_HIP(hipMemcpyAsync(x, b, size_t(s->n) * 8, hipMemcpyDeviceToDevice, st));
  apply_domains(s, b, x);
  if (s->coarse) apply_domains(s->coarse, b, x);  // solved last (src/Preconditioners.h:1489-1520)
}

int64_t schwarz_size(const abh_schwarz *s) { return s->n; }
abh_context *schwarz_context(const abh_schwarz *s) { return s->ctx; }

abh_schwarz *schwarz_create_batch(abh_context *ctx, int dim, const std::vector<int64_t> &off,
                                  const double *pos_dev, const double *lo, const double *hi,
                                  double n_leaf, const KernelDesc &k, double shift) {
  std::unique_ptr<abh_schwarz> s(new abh_schwarz());
  s->ctx = ctx;
  s->dim = dim;
  s->n = off.back();
  ABH_REQUIRE(s->n < (int64_t(1) << 31), ABH_ERR_INVALID_ARG, "batch too large for int32 indices");
  std::vector<double> pos(size_t(s->n) * dim);
  hipStream_t hs = ctx->stream;
  if (s->n > 0)
    ABH_HIP(hipMemcpyAsync(pos.data(), pos_dev, pos.size() * 8, hipMemcpyDeviceToHost, hs));
  ABH_HIP(hipStreamSynchronize(hs));
  analyse_batch_into(s.get(), dim, off, pos, lo, hi, n_leaf);
  for (size_t d = 0; d + 1 < s->off.size(); ++d)
    s->max_m = std::max(s->max_m, int(s->off[d + 1] - s->off[d]));
  s->batch = true;
  factorise(s.get(), pos_dev, k, -1.0, shift, true);
  return s.release();
}

}  // namespace abh

extern "C" {

abh_status abh_schwarz_create_coarse(abh_context *ctx, const abh_cells *cells,
                                     const abh_kernel *kernel, double radius, double diag_shift,
                                     int coarse_grid_n, uint64_t seed, abh_schwarz **out) {
  if (!ctx || !cells || !kernel || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_schwarz *s = new (std::nothrow) abh_schwarz();
  if (!s) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
                ABH_ERR_UNSUPPORTED, "unknown kernel type");
    ABH_REQUIRE(coarse_grid_n >= 0 && coarse_grid_n <= 4096, ABH_ERR_INVALID_ARG,
                "coarse grid size must be in [0, 4096]");
    for (int d = 0; d < cells->dim; ++d)
      ABH_REQUIRE(!cells->periodic[d], ABH_ERR_UNSUPPORTED,
                  "Schwarz preconditioner needs a non-periodic cell list");
    s->ctx = ctx;
    s->dim = cells->dim;
    s->n = cells->n_alive;
    const int D = cells->dim;
    std::vector<double> pos(size_t(s->n) * D);
    std::vector<uint32_t> beg(size_t(cells->nb)), end(size_t(cells->nb));
    hipStream_t hs = ctx->stream;
    abh::StreamDrain drain{hs};
    if (s->n > 0) {
      ABH_HIP(hipMemcpyAsync(pos.data(), cells->pos.p, pos.size() * 8, hipMemcpyDeviceToHost, hs));
      ABH_HIP(hipMemcpyAsync(beg.data(), cells->begin.p, beg.size() * 4, hipMemcpyDeviceToHost, hs));
      ABH_HIP(hipMemcpyAsync(end.data(), cells->end.p, end.size() * 4, hipMemcpyDeviceToHost, hs));
    }
    ABH_HIP(hipStreamSynchronize(hs));
    if (s->n > 0) {
      analyse_into(s, cells, pos, beg, end);
      if (coarse_grid_n > 0) coarse_into(s, cells, beg, end, coarse_grid_n, seed);
    }
    for (size_t d = 0; d + 1 < s->off.size(); ++d)
      s->max_m = std::max(s->max_m, int(s->off[d + 1] - s->off[d]));
    KernelDesc k{};
    k.type = kernel->type;
    k.param = kernel->param;
    k.param2 = kernel->param * kernel->param;
    abh::set_inverses(k);
    k.w0 = 1.0;
    const double r2 = radius > 0.0 ? radius * radius : -1.0;
    factorise(s, cells->pos.as<double>(), k, r2, diag_shift, false);
    if (s->coarse) factorise(s->coarse, cells->pos.as<double>(), k, r2, diag_shift, false);
  });
  if (st != ABH_OK) {
    delete s;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = s;
  return ABH_OK;
}

abh_status abh_schwarz_create(abh_context *ctx, const abh_cells *cells, const abh_kernel *kernel,
                              double radius, double diag_shift, abh_schwarz **out) {
  return abh_schwarz_create_coarse(ctx, cells, kernel, radius, diag_shift, 0, 0, out);
}

abh_status abh_schwarz_factorize(abh_schwarz *s, const abh_cells *cells,
                                 const abh_kernel *kernel, double radius, double diag_shift) {
  if (!s || !cells || !kernel) return ABH_ERR_INVALID_ARG;
  return abh::guard(s->ctx, [&] {
    ABH_REQUIRE(!s->batch, ABH_ERR_UNSUPPORTED,
                "refactorisation is for preconditioners built on a cell list");
    ABH_REQUIRE(cells->ctx == s->ctx && cells->dim == s->dim && cells->n_alive == s->n,
                ABH_ERR_INVALID_ARG, "cell list does not match the analysed pattern");
    ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
                ABH_ERR_UNSUPPORTED, "unknown kernel type");
    KernelDesc k{};
    k.type = kernel->type;
    k.param = kernel->param;
    k.param2 = kernel->param * kernel->param;
    abh::set_inverses(k);
    k.w0 = 1.0;
    const double r2 = radius > 0.0 ? radius * radius : -1.0;
    s->failed = 0;
    factorise(s, cells->pos.as<double>(), k, r2, diag_shift, false);
    if (s->coarse) {
      s->coarse->failed = 0;
      factorise(s->coarse, cells->pos.as<double>(), k, r2, diag_shift, false);
    }
  });
}

abh_status abh_schwarz_destroy(abh_schwarz *s) {
  if (!s) return ABH_OK;
  abh_context *ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete s;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_schwarz_get_info(const abh_schwarz *s, abh_schwarz_info *info) {
  if (!s || !info) return ABH_ERR_INVALID_ARG;
  abh_schwarz_info r{};
  r.n = s->n;
  r.n_domains = int64_t(s->nint.size()) + (s->coarse ? 1 : 0);
  r.min_size = r.n_domains ? INT64_MAX : 0;
  r.min_interior = r.n_domains ? INT64_MAX : 0;
  auto add = [&](int64_t m, int64_t ni) {
    r.min_size = std::min(r.min_size, m);
    r.max_size = std::max(r.max_size, m);
    r.min_interior = std::min<int64_t>(r.min_interior, ni);
    r.max_interior = std::max<int64_t>(r.max_interior, ni);
  };
  for (size_t d = 0; d < s->nint.size(); ++d) add(s->off[d + 1] - s->off[d], s->nint[d]);
  r.factor_bytes = int64_t(s->factor_bytes);
  r.failed_domains = s->failed;
  if (s->coarse) {
    add(int64_t(s->coarse->idx.size()), s->coarse->nint[0]);
    r.factor_bytes += int64_t(s->coarse->factor_bytes);
    r.failed_domains += s->coarse->failed;
  }
  *info = r;
  return ABH_OK;
}

abh_status abh_schwarz_get_domains(const abh_schwarz *s, int64_t *offsets, int32_t *indices,
                                   int32_t *n_interior) {
  if (!s) return ABH_ERR_INVALID_ARG;
  const size_t nd = s->nint.size();
  if (offsets) std::copy(s->off.begin(), s->off.end(), offsets);
  if (indices) std::copy(s->idx.begin(), s->idx.end(), indices);
  if (n_interior) std::copy(s->nint.begin(), s->nint.end(), n_interior);
  if (s->coarse) {  // the coarse domain last, all interior
    const abh_schwarz *q = s->coarse;
    if (offsets) offsets[nd + 1] = s->off[nd] + int64_t(q->idx.size());
    if (indices) std::copy(q->idx.begin(), q->idx.end(), indices + s->off[nd]);
    if (n_interior) n_interior[nd] = q->nint[0];
  }
  return ABH_OK;
}

abh_status abh_schwarz_apply(const abh_schwarz *s, const double *b, double *x) {
  if (!s) return ABH_ERR_INVALID_ARG;
  return abh::guard(s->ctx, [&] {
    ABH_REQUIRE(s->n == 0 || (b && x), ABH_ERR_INVALID_ARG, "b or x is null");
    ABH_REQUIRE(b != x || s->n == 0, ABH_ERR_INVALID_ARG, "b and x must be distinct");
    abh::schwarz_apply(s, b, x);
  });
}

}  // extern "C"
