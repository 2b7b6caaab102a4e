// Cell-pair traversal of a CellListOrdered: the sequence of bucket pairs
// bucket_pair_iterator yields (get_neighbouring_buckets(query),
// src/Search.h:498-770, :857-860), materialised on the device, and the
// neighbour counts of the reference's symmetric pair loop over it
// (tests/neighbours.h:892-951).
//
// Sequence (the iterator's order): first every bucket i but the last in
// lattice order against the buckets j of its 3^D block (clipped to the grid)
// that come after i; then, for each periodic quadrant q after 0 in the
// lattice order of {-1,0,1}^D (periodic dimensions only), every bucket i of
// the face that q's ghosts touch against the clipped 3^D block around its
// ghost image i + q * size, the pair's position offset being q * (high -
// low).  Each source (phase, i) owns a contiguous run of pairs: a count
// kernel sizes the runs, an exclusive scan places them, a fill kernel
// writes them -- in the iterator's order.
//
// Counts: one wavefront per bucket pair, lanes over the particles a of
// bucket i (64 at a time), every particle b of bucket j tested against all
// of them: |(x_a + offset) - x_b|^2 < r^2 (the offset added first, squares
// summed in dimension order, as the reference computes it); a lane adds its
// hits to counts[a], and each b gets the ballot's population count -- one
// integer atomic per particle per pair, so the result is deterministic.
// Then one wavefront per bucket for the pairs inside it (each particle
// counts itself, a < b within r adds one to both).
#include <algorithm>
#include <vector>

#include "cells.hpp"

namespace {

constexpr int kMaxPhases = 14;  // domain-domain + 13 quadrants (3D)

struct Phases {
  int n;                         // number of phases
  int64_t src_begin[kMaxPhases + 1];
  int32_t q[kMaxPhases][3];      // periodic quadrant (0 for domain-domain)
  int32_t rs[kMaxPhases][3];     // regular buckets: start per dimension
  int32_t rc[kMaxPhases][3];     // regular buckets: count per dimension
  int32_t size[3];
  int32_t dim;
};

struct Block {
  int32_t s[3], e[3];  // inclusive clipped range
};

__device__ __forceinline__ void source_of(const Phases &P, int64_t s, int &ph, int32_t *iv) {
  ph = 0;
  while (ph + 1 < P.n && s >= P.src_begin[ph + 1]) ++ph;
  int64_t r = s - P.src_begin[ph];
  for (int d = 2; d >= 0; --d) {
    iv[d] = P.rs[ph][d] + int32_t(r % P.rc[ph][d]);
    r /= P.rc[ph][d];
  }
}

// clipped 3^D block around bucket b (dimensions past dim have size 1)
__device__ __forceinline__ Block block_of(const Phases &P, const int32_t *b) {
  Block k;
  for (int d = 0; d < 3; ++d) {
    const int32_t end = P.size[d] - 1;
    k.s[d] = max(b[d] - 1, 0);
    k.e[d] = min(b[d] + 1, end);
  }
  return k;
}

__device__ __forceinline__ int32_t collapse3(const Phases &P, const int32_t *v) {
  return (v[0] * P.size[1] + v[1]) * P.size[2] + v[2];
}

template <bool FILL>
__global__ __launch_bounds__(256) void k_pairs(Phases P, int64_t nsrc,
                                               int64_t *__restrict__ counts,
                                               int32_t *__restrict__ bi,
                                               int32_t *__restrict__ bj,
                                               int8_t *__restrict__ img) {
  const int64_t s = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (s >= nsrc) return;
  int ph;
  int32_t iv[3];
  source_of(P, s, ph, iv);
  int32_t base[3];
  for (int d = 0; d < 3; ++d) base[d] = iv[d] + P.q[ph][d] * P.size[d];
  const Block k = block_of(P, base);
  const int32_t i = collapse3(P, iv);
  int64_t out = FILL ? counts[s] : 0;
  int64_t c = 0;
  int32_t v[3];
  for (v[0] = k.s[0]; v[0] <= k.e[0]; ++v[0])
    for (v[1] = k.s[1]; v[1] <= k.e[1]; ++v[1])
      for (v[2] = k.s[2]; v[2] <= k.e[2]; ++v[2]) {
        const int32_t j = collapse3(P, v);
        if (ph == 0 && j <= i) continue;  // domain-domain: j after i
        if (FILL) {
          bi[out] = i;
          bj[out] = j;
          for (int d = 0; d < P.dim; ++d) img[out * P.dim + d] = int8_t(P.q[ph][d]);
          ++out;
        } else {
          ++c;
        }
      }
  if (!FILL) counts[s] = c;
}

template <int D>
__global__ __launch_bounds__(256) void k_pair_counts(int64_t npairs, const int32_t *__restrict__ bi,
                                                     const int32_t *__restrict__ bj,
                                                     const int8_t *__restrict__ img,
                                                     const uint32_t *__restrict__ beg,
                                                     const uint32_t *__restrict__ end,
                                                     const double *__restrict__ pos, double L0,
                                                     double L1, double L2, double r2,
                                                     int32_t *__restrict__ counts) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6;
  if (p >= npairs) return;  // uniform per wavefront
  const double Ls[3] = {L0, L1, L2};
  double off[D];
#pragma unroll
  for (int d = 0; d < D; ++d) off[d] = double(img[p * D + d]) * Ls[d];
  const uint32_t ib = uint32_t(bi[p]), jb = uint32_t(bj[p]);
  const uint32_t a0 = beg[ib], a1 = end[ib], b0 = beg[jb], b1 = end[jb];
  for (uint32_t ac = a0; ac < a1; ac += 64) {
    const uint32_t a = ac + lane;
    const bool act = a < a1;
    double pa[D];
#pragma unroll
    for (int d = 0; d < D; ++d) pa[d] = act ? pos[size_t(a) * D + d] + off[d] : 0.0;
    int hits = 0;
    for (uint32_t b = b0; b < b1; ++b) {
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double dx = pa[d] - pos[size_t(b) * D + d];
        s = s + dx * dx;
      }
      const bool hit = act && s < r2;
      hits += hit ? 1 : 0;
      const int nb = __popcll(__ballot(hit));
      if (lane == 0 && nb) atomicAdd(&counts[b], nb);
    }
    if (act && hits) atomicAdd(&counts[a], hits);
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_self_counts(int64_t nb, const uint32_t *__restrict__ beg,
                                                     const uint32_t *__restrict__ end,
                                                     const double *__restrict__ pos, double r2,
                                                     int32_t *__restrict__ counts) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6;
  if (k >= nb) return;
  const uint32_t a0 = beg[k], a1 = end[k];
  for (uint32_t ac = a0; ac < a1; ac += 64) {
    const uint32_t a = ac + lane;
    const bool act = a < a1;
    double pa[D];
#pragma unroll
    for (int d = 0; d < D; ++d) pa[d] = act ? pos[size_t(a) * D + d] : 0.0;
    int hits = act ? 1 : 0;  // itself
    for (uint32_t b = ac + 1; b < a1; ++b) {
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double dx = pa[d] - pos[size_t(b) * D + d];
        s = s + dx * dx;
      }
      const bool hit = act && b > a && s < r2;
      hits += hit ? 1 : 0;
      const int nbits = __popcll(__ballot(hit));
      if (lane == 0 && nbits) atomicAdd(&counts[b], nbits);
    }
    if (act) atomicAdd(&counts[a], hits);
  }
}

Phases make_phases(const abh_cells *c) {
  Phases P{};
  P.dim = c->dim;
  for (int d = 0; d < 3; ++d) P.size[d] = d < c->dim ? int32_t(c->size[d]) : 1;
  // phase 0: every bucket but the last
  P.n = 1;
  P.src_begin[0] = 0;
  for (int d = 0; d < 3; ++d) {
    P.q[0][d] = 0;
    P.rs[0][d] = 0;
    P.rc[0][d] = P.size[d];
  }
  int64_t total = std::max<int64_t>(c->nb - 1, 0);
  // periodic quadrants after 0 in the lattice order of {-1,0,1} per
  // periodic dimension ({0} otherwise), last dimension fastest
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    const bool per = d < c->dim && c->periodic[d];
    lo[d] = per ? -1 : 0;
    hi[d] = per ? 1 : 0;
  }
  bool past_zero = false;
  for (int q0 = lo[0]; q0 <= hi[0]; ++q0)
    for (int q1 = lo[1]; q1 <= hi[1]; ++q1)
      for (int q2 = lo[2]; q2 <= hi[2]; ++q2) {
        const int q[3] = {q0, q1, q2};
        if (q0 == 0 && q1 == 0 && q2 == 0) {
          past_zero = true;
          continue;
        }
        if (!past_zero) continue;
        const int ph = P.n++;
        P.src_begin[ph] = total;
        int64_t cnt = 1;
        for (int d = 0; d < 3; ++d) {
          P.q[ph][d] = q[d];
          P.rs[ph][d] = q[d] < 0 ? P.size[d] - 1 : 0;
          P.rc[ph][d] = q[d] != 0 ? 1 : P.size[d];
          cnt *= P.rc[ph][d];
        }
        total += cnt;
      }
  P.src_begin[P.n] = total;
  return P;
}

}  // namespace

namespace abh {

// Writes the pair sequence into the cell list's buffers (or only counts it
// when fill is false); returns the number of pairs.  Synchronises.
int64_t bucket_pairs(abh_cells *c, bool fill) {
  abh_context *ctx = c->ctx;
  hipStream_t st = ctx->stream;
  const Phases P = make_phases(c);
  const int64_t nsrc = P.src_begin[P.n];
  c->pair_src.ensure(size_t(nsrc + 1) * 8 + 64);
  int64_t *cnt = c->pair_src.as<int64_t>();
  ABH_HIP(hipMemsetAsync(cnt + nsrc, 0, 8, st));
  if (nsrc > 0) {
    k_pairs<false><<<unsigned(ceil_div(nsrc, 256)), 256, 0, st>>>(P, nsrc, cnt, nullptr, nullptr,
                                                                  nullptr);
    check_launch();
  }
  c->pair_total.ensure(64);
  int64_t *tot = c->pair_total.as<int64_t>();
  scan_exclusive(ctx, cnt, cnt, nsrc + 1, tot);
  int64_t npairs = 0;
  ABH_HIP(hipMemcpyAsync(&npairs, tot, 8, hipMemcpyDeviceToHost, st));
  ABH_HIP(hipStreamSynchronize(st));
  c->npairs = npairs;
  if (!fill) return npairs;
  c->pair_i.ensure(size_t(npairs) * 4 + 64);
  c->pair_j.ensure(size_t(npairs) * 4 + 64);
  c->pair_img.ensure(size_t(npairs) * c->dim + 64);
  if (nsrc > 0 && npairs > 0) {
    k_pairs<true><<<unsigned(ceil_div(nsrc, 256)), 256, 0, st>>>(
        P, nsrc, cnt, c->pair_i.as<int32_t>(), c->pair_j.as<int32_t>(), c->pair_img.as<int8_t>());
    check_launch();
  }
  return npairs;
}

}  // namespace abh

extern "C" {

abh_status abh_cells_bucket_pairs(abh_cells *cells, int64_t capacity, int32_t *bi, int32_t *bj,
                                  int8_t *image, int64_t *npairs) {
  if (!cells) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = cells->ctx;
  return abh::guard(ctx, [&] {
    const bool want = bi || bj || image;
    const int64_t n = abh::bucket_pairs(cells, want);
    if (npairs) *npairs = n;
    if (!want) return;
    ABH_REQUIRE(bi && bj && image, ABH_ERR_INVALID_ARG, "bi, bj and image must all be given");
    ABH_REQUIRE(capacity >= n, ABH_ERR_CAPACITY, "pair buffers too small");
    hipStream_t st = ctx->stream;
    if (n == 0) return;
    ABH_HIP(hipMemcpyAsync(bi, cells->pair_i.p, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
    ABH_HIP(hipMemcpyAsync(bj, cells->pair_j.p, size_t(n) * 4, hipMemcpyDeviceToDevice, st));
    ABH_HIP(hipMemcpyAsync(image, cells->pair_img.p, size_t(n) * cells->dim,
                           hipMemcpyDeviceToDevice, st));
  });
}

abh_status abh_cells_pair_counts(abh_cells *cells, double radius, int32_t *counts) {
  if (!cells) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = cells->ctx;
  return abh::guard(ctx, [&] {
    ABH_REQUIRE(radius >= 0.0, ABH_ERR_INVALID_ARG, "radius must be non-negative");
    const int64_t n = cells->n_alive;
    if (n == 0) return;
    ABH_REQUIRE(counts, ABH_ERR_INVALID_ARG, "counts is null");
    for (int d = 0; d < cells->dim; ++d)
      ABH_REQUIRE(cells->side[d] >= radius, ABH_ERR_INVALID_ARG,
                  "bucket side below the radius: the 3^dim pair traversal would miss pairs");
    hipStream_t st = ctx->stream;
    const int64_t npairs = abh::bucket_pairs(cells, true);
    ABH_HIP(hipMemsetAsync(counts, 0, size_t(n) * 4, st));
    const double r2 = radius * radius;
    double L[3] = {0, 0, 0};
    for (int d = 0; d < cells->dim; ++d) L[d] = cells->hi[d] - cells->lo[d];
    const double *pos = cells->pos.as<double>();
    const uint32_t *beg = cells->begin.as<uint32_t>(), *end = cells->end.as<uint32_t>();
    const int32_t *pi = cells->pair_i.as<int32_t>(), *pj = cells->pair_j.as<int32_t>();
    const int8_t *im = cells->pair_img.as<int8_t>();
#define ABH_PC(DD)                                                                              \
  do {                                                                                          \
    if (npairs > 0)                                                                             \
      k_pair_counts<DD><<<unsigned(abh::ceil_div(npairs, 4)), 256, 0, st>>>(                    \
          npairs, pi, pj, im, beg, end, pos, L[0], L[1], L[2], r2, counts);                     \
    k_self_counts<DD><<<unsigned(abh::ceil_div(cells->nb, 4)), 256, 0, st>>>(cells->nb, beg, end, \
                                                                            pos, r2, counts);   \
  } while (0)
    if (cells->dim == 1) ABH_PC(1);
    else if (cells->dim == 2) ABH_PC(2);
    else ABH_PC(3);
#undef ABH_PC
    abh::check_launch();
  });
}

}  // extern "C"
