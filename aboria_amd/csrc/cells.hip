// Context / memory entry points and the CellListOrdered build.
//
// Build pipeline on one stream (neighbour_search_base::update_positions,
// src/NeighbourSearchBase.h:350-495, and CellListOrdered::update_positions_impl,
// src/CellListOrdered.h:190-259):
//   1. k_enforce_domain: periodic wrap / kill, alive flags as uint32
//   2. exclusive scan of the flags -> compacted slot of every alive particle;
//      the alive count is read back (the reference reads m_alive_sum.back())
//   3. bucket sizing on the host (set_domain_impl, src/CellListOrdered.h:132-186)
//   4. k_compact_bucket: scatter_if + point_to_bucket_index fused
//   5. stable LSD radix sort of (bucket, alive index) on the bits that occur
//   6. bucket offsets: counts per bucket from the sorted keys, exclusive scan
//   7. gather of the positions into sorted order (Particles::reorder)
// All FP64 arithmetic is compiled without contraction so bucket indices are
// the reference's bit for bit.
#include <cmath>
#include <cstring>

#include "cells.hpp"
#include "solvers.hpp"

using abh::DevBuf;
using abh::GridParams;

namespace abh {
void context_free(abh_context *ctx) {
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->scan_blocks.release();
  ctx->sort_keys.release();
  ctx->sort_vals.release();
  ctx->sort_hist.release();
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}
}  // namespace abh

namespace {

constexpr int kThreads = 256;
constexpr int kMaxWrap = 64;  // periodic wrap iterations before a particle is dropped

using abh::BuildCounters;

template <int D>
__global__ __launch_bounds__(kThreads) void k_enforce_domain(double *__restrict__ pos,
                                                             uint8_t *__restrict__ alive,
                                                             uint32_t *__restrict__ flags,
                                                             int64_t n, GridParams g,
                                                             BuildCounters *counters) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  bool ok = alive ? alive[i] != 0 : true;
  double r[D];
#pragma unroll
  for (int d = 0; d < D; ++d) r[d] = pos[i * D + d];
  bool wrap_fail = false;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (!isfinite(r[d])) {
      ok = false;
    } else if (g.periodic[d]) {
      // enforce_domain_lambda (src/NeighbourSearchBase.h:218-224): repeated
      // +/- L, bounded here so absurd inputs cannot spin a wave.
      int it = 0;
      while (r[d] < g.lo[d] && it < kMaxWrap) { r[d] += g.L[d]; ++it; }
      while (r[d] >= g.hi[d] && it < kMaxWrap) { r[d] -= g.L[d]; ++it; }
      if (r[d] < g.lo[d] || r[d] >= g.hi[d]) { ok = false; wrap_fail = true; }
    } else if (r[d] < g.lo[d] || r[d] >= g.hi[d]) {
      ok = false;
    }
  }
  // only a periodic dimension can change; the others are left unwritten
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (g.periodic[d]) pos[i * D + d] = r[d];
  if (alive) alive[i] = ok ? 1 : 0;
  flags[i] = ok ? 1u : 0u;
  if (wrap_fail) atomicAdd(&counters->n_wrap_fail, 1u);
}

// point_to_bucket_index::find_bucket_index + collapse_index_vector
// (src/detail/SpatialUtil.h:49-60, 122-138); modulo-2^32 like the reference.
template <int D> __device__ __forceinline__ uint32_t bucket_of(const GridParams &g, const double *p) {
  uint32_t index = 0, mult = 1;
#pragma unroll
  for (int i = D - 1; i >= 0; --i) {
    if (i != D - 1) mult *= uint32_t(g.size[i + 1]);
    const int v = int(floor((p[i] - g.lo[i]) * g.inv[i]));
    index += mult * uint32_t(v);
  }
  return index;
}

template <int D>
__global__ __launch_bounds__(kThreads) void k_compact_bucket(const double *__restrict__ pos,
                                                             const uint32_t *__restrict__ flags,
                                                             const uint32_t *__restrict__ offs,
                                                             int64_t n, GridParams g,
                                                             int32_t *__restrict__ order,
                                                             uint32_t *__restrict__ keys) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n || !flags[i]) return;
  double p[D];
#pragma unroll
  for (int d = 0; d < D; ++d) p[d] = pos[i * D + d];
  const uint32_t k = offs[i];
  order[k] = int32_t(i);
  keys[k] = bucket_of<D>(g, p);
}

// Bucket offsets from counts (src/CellListOrdered.h:231-239 takes lower /
// upper bounds of every bucket id in the sorted keys; with counts c_b,
// lower_bound(b) = sum_{b' < b} c_b' and upper_bound(b) = that + c_b):
// k_bucket_counts adds, per run of equal keys within a wavefront, the run's
// length to count[key] once (keys >= nb lie past every bucket and are not
// counted); an exclusive scan of the counts gives begin, end = begin + count.
__global__ __launch_bounds__(kThreads) void k_fill_u32(uint32_t *__restrict__ p, int64_t n,
                                                       uint32_t v) {
  const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ __launch_bounds__(kThreads) void k_bucket_counts(const uint32_t *__restrict__ keys,
                                                            int64_t n, int64_t nb,
                                                            uint32_t *__restrict__ count) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = i < n;
  const uint32_t k = valid ? keys[i] : 0xffffffffu;
  const uint32_t kp = __shfl_up(k, 1, 64);
  const bool head = valid && (lane == 0 || kp != k);
  const uint64_t heads = __ballot(head);
  const uint64_t vmask = __ballot(valid);
  if (head && int64_t(k) < nb) {
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes <= lane
    const uint64_t after = heads & ~upto;
    const int next = after ? __ffsll((long long)after) - 1 : 64 - __clzll((long long)vmask);
    atomicAdd(&count[k], uint32_t(next - lane));
  }
}

// end[b] = begin[b] + count[b], the counts held in end
__global__ __launch_bounds__(kThreads) void k_bucket_ends(const uint32_t *__restrict__ begin,
                                                          int64_t nb, uint32_t *__restrict__ end) {
  const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (t < nb) end[t] = begin[t] + end[t];
}

template <int D> void launch_enforce(abh_context *ctx, const GridParams &g, int64_t n,
                                     double *pos, uint8_t *alive, uint32_t *flags,
                                     BuildCounters *ctr) {
  k_enforce_domain<D><<<unsigned(abh::ceil_div(n, kThreads)), kThreads, 0, ctx->stream>>>(
      pos, alive, flags, n, g, ctr);
}
template <int D> void launch_compact(abh_cells *c, int64_t n, const double *pos,
                                     const uint32_t *flags, const uint32_t *offs) {
  k_compact_bucket<D><<<unsigned(abh::ceil_div(n, kThreads)), kThreads, 0, c->ctx->stream>>>(
      pos, flags, offs, n, c->params(), c->order.as<int32_t>(), c->keys.as<uint32_t>());
}

// CellListOrdered::set_domain_impl (src/CellListOrdered.h:132-186).
void size_grid(abh_cells *c, size_t n) {
  if (!(n < 0.5 * double(c->sized_with_n) || n > 2 * c->sized_with_n)) return;
  c->sized_with_n = n;
  const int D = c->dim;
  if (c->n_leaf > double(n)) {
    for (int d = 0; d < D; ++d) c->size[d] = 1;
  } else {
    double volume = 1;
    for (int d = 0; d < D; ++d) volume *= (c->hi[d] - c->lo[d]);
    const double box_volume = c->n_leaf / double(n) * volume;
    const double box_side = std::pow(box_volume, 1.0 / D);
    for (int d = 0; d < D; ++d) {
      c->size[d] = static_cast<unsigned>(std::floor((c->hi[d] - c->lo[d]) / box_side));
      if (c->size[d] == 0) c->size[d] = 1;
    }
  }
  int64_t nb = 1;
  for (int d = 0; d < D; ++d) {
    c->side[d] = (c->hi[d] - c->lo[d]) / c->size[d];
    c->inv[d] = 1.0 / c->side[d];
    nb *= int64_t(c->size[d]);
  }
  ABH_REQUIRE(nb < (int64_t(1) << 31), ABH_ERR_INVALID_ARG, "too many buckets");
  c->nb = nb;
}

int bits_for(uint64_t v) {
  int b = 0;
  while (b < 32 && (v >> b) != 0) ++b;
  return b;
}

}  // namespace

namespace abh {
void enforce_domain(abh_context *ctx, int dim, const GridParams &g, int64_t n, double *pos,
                    uint8_t *alive, uint32_t *flags, BuildCounters *counters) {
  if (n <= 0) return;
  if (dim == 1) launch_enforce<1>(ctx, g, n, pos, alive, flags, counters);
  else if (dim == 2) launch_enforce<2>(ctx, g, n, pos, alive, flags, counters);
  else launch_enforce<3>(ctx, g, n, pos, alive, flags, counters);
  check_launch();
}
}  // namespace abh

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

abh_status abh_context_create(int device, void *stream, abh_context **out) {
  if (!out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_context *ctx = new (std::nothrow) abh_context();
  if (!ctx) return ABH_ERR_OUT_OF_MEMORY;
  ctx->device = device;
  abh_status st = abh::guard(ctx, [&] {
    // NULL is HIP's null stream (torch's default stream), used as is
    ctx->stream = static_cast<hipStream_t>(stream);
    ctx->own_stream = false;
    hipDeviceProp_t prop;
    ABH_HIP(hipGetDeviceProperties(&prop, device));
    ctx->n_cu = prop.multiProcessorCount;
    // the device's copy of the allocation-free dense apply's scratch
    ctx->dense_scratch = abh::dense_scratch_address();
  });
  if (st != ABH_OK) {
    delete ctx;
    return st;
  }
  *out = ctx;
  return ABH_OK;
}

abh_status abh_context_destroy(abh_context *ctx) {
  if (!ctx || ctx->closed) return ABH_OK;
  ctx->closed = true;
  if (ctx->refs == 0) abh::context_free(ctx);
  return ABH_OK;
}

abh_status abh_context_sync(abh_context *ctx) {
  if (!ctx) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] { ABH_HIP(hipStreamSynchronize(ctx->stream)); });
}

const char *abh_status_string(abh_status s) {
  switch (s) {
    case ABH_OK: return "ok";
    case ABH_ERR_INVALID_ARG: return "invalid argument";
    case ABH_ERR_HIP: return "HIP runtime error";
    case ABH_ERR_OUT_OF_MEMORY: return "out of device memory";
    case ABH_ERR_CAPACITY: return "output buffer too small";
    case ABH_ERR_UNSUPPORTED: return "unsupported";
    case ABH_ERR_DOMAIN: return "position outside domain";
    case ABH_ERR_NOT_CONVERGED: return "not converged";
    case ABH_ERR_NUMERICAL: return "numerical failure";
  }
  return "unknown status";
}

const char *abh_last_error(const abh_context *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

abh_status abh_malloc(abh_context *ctx, void **ptr, size_t bytes) {
  if (!ctx || !ptr) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] { ABH_HIP(hipMalloc(ptr, bytes ? bytes : 1)); });
}
abh_status abh_free(abh_context *ctx, void *ptr) {
  if (!ctx) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] {
    if (ptr) ABH_HIP(hipFree(ptr));
  });
}
abh_status abh_memcpy_h2d(abh_context *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] {
    ABH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    ABH_HIP(hipStreamSynchronize(ctx->stream));
  });
}
abh_status abh_memcpy_d2h(abh_context *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] {
    ABH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    ABH_HIP(hipStreamSynchronize(ctx->stream));
  });
}
abh_status abh_memcpy_d2d(abh_context *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] {
    if (bytes) ABH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  });
}
abh_status abh_memset(abh_context *ctx, void *dst, int value, size_t bytes) {
  if (!ctx) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] { ABH_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream)); });
}

abh_status abh_cells_create(abh_context *ctx, int dim, const double *low, const double *high,
                            const uint8_t *periodic, double n_leaf, abh_cells **out) {
  if (!ctx || !out || !low || !high || !periodic || dim < 1 || dim > 3)
    return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_cells *c = new (std::nothrow) abh_cells();
  if (!c) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    c->ctx = ctx;
    c->dim = dim;
    for (int d = 0; d < dim; ++d) {
      ABH_REQUIRE(high[d] > low[d], ABH_ERR_INVALID_ARG, "domain high must exceed low");
      c->lo[d] = low[d];
      c->hi[d] = high[d];
      c->periodic[d] = periodic[d] != 0;
    }
    ABH_REQUIRE(n_leaf > 0, ABH_ERR_INVALID_ARG, "n_particles_in_leaf must be positive");
    c->n_leaf = n_leaf;
    // set_domain at init_neighbour_search time sizes with no alive particles
    size_grid(c, 0);
    c->counters.ensure(sizeof(BuildCounters));
  });
  if (st != ABH_OK) {
    delete c;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = c;
  return ABH_OK;
}

abh_status abh_cells_destroy(abh_cells *c) {
  if (!c) return ABH_OK;
  abh_context *ctx = c->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete c;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_cells_update(abh_cells *c, int64_t n, double *pos, uint8_t *alive,
                            int64_t *n_alive_out) {
  if (!c) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = c->ctx;
  return abh::guard(ctx, [&] {
    ABH_REQUIRE(n >= 0 && n < (int64_t(1) << 31), ABH_ERR_INVALID_ARG, "n out of range");
    ABH_REQUIRE(n == 0 || pos, ABH_ERR_INVALID_ARG, "pos is null");
    const int D = c->dim;
    hipStream_t s = ctx->stream;
    BuildCounters *ctr = c->counters.as<BuildCounters>();
    ABH_HIP(hipMemsetAsync(ctr, 0, sizeof(BuildCounters), s));
    c->flags.ensure(size_t(n + 1) * 4);
    c->offs.ensure(size_t(n + 1) * 4);
    uint32_t *flags = c->flags.as<uint32_t>();
    uint32_t *offs = c->offs.as<uint32_t>();
    if (n > 0) {
      abh::enforce_domain(ctx, D, c->params(), n, pos, alive, flags, ctr);
    }
    abh::scan_exclusive(ctx, flags, offs, n, &ctr->n_alive);
    BuildCounters h{};
    ABH_HIP(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    const int64_t na = int64_t(h.n_alive);
    c->n_alive = na;
    size_grid(c, size_t(na));
    const int64_t nb = c->nb;
    c->order.ensure(size_t(na + 1) * 4);
    c->keys.ensure(size_t(na + 1) * 4);
    c->begin.ensure(size_t(nb) * 4);
    c->end.ensure(size_t(nb) * 4);
    c->pos.ensure(size_t(na + 1) * D * 8);
    if (na > 0) {
      if (D == 1) launch_compact<1>(c, n, pos, flags, offs);
      else if (D == 2) launch_compact<2>(c, n, pos, flags, offs);
      else launch_compact<3>(c, n, pos, flags, offs);
      abh::check_launch();
      // Positions lie in [lo, hi) after enforcement, so each per-dimension
      // index is at most size[d]; the collapsed key is below 2*nb.
      abh::radix_sort_pairs(ctx, c->keys.as<uint32_t>(), c->order.as<int32_t>(), na,
                            bits_for(uint64_t(2 * nb - 1)));
    }
    {
      uint32_t *cnt = c->end.as<uint32_t>();
      k_fill_u32<<<unsigned(abh::ceil_div(nb, kThreads)), kThreads, 0, s>>>(cnt, nb, 0u);
      abh::check_launch();
      if (na > 0) {
        k_bucket_counts<<<unsigned(abh::ceil_div(na, kThreads)), kThreads, 0, s>>>(
            c->keys.as<uint32_t>(), na, nb, cnt);
        abh::check_launch();
      }
      abh::scan_exclusive(ctx, cnt, c->begin.as<uint32_t>(), nb, nullptr);
      k_bucket_ends<<<unsigned(abh::ceil_div(nb, kThreads)), kThreads, 0, s>>>(
          c->begin.as<uint32_t>(), nb, cnt);
      abh::check_launch();
    }
    if (na > 0)
      abh::gather_bytes(ctx, c->order.as<int32_t>(), na, pos, c->pos.p, size_t(D) * 8);
    if (n_alive_out) *n_alive_out = na;
  });
}

abh_status abh_cells_get_info(const abh_cells *c, abh_cells_info *info) {
  if (!c || !info) return ABH_ERR_INVALID_ARG;
  std::memset(info, 0, sizeof(*info));
  info->dim = c->dim;
  for (int d = 0; d < c->dim; ++d) {
    info->grid[d] = c->size[d];
    info->side[d] = c->side[d];
    info->low[d] = c->lo[d];
    info->high[d] = c->hi[d];
    info->periodic[d] = c->periodic[d] ? 1 : 0;
  }
  info->n_buckets = c->nb;
  info->n_alive = c->n_alive;
  return ABH_OK;
}

const int32_t *abh_cells_order(const abh_cells *c) { return c ? c->order.as<int32_t>() : nullptr; }
const uint32_t *abh_cells_bucket_index(const abh_cells *c) {
  return c ? c->keys.as<uint32_t>() : nullptr;
}
const uint32_t *abh_cells_bucket_begin(const abh_cells *c) {
  return c ? c->begin.as<uint32_t>() : nullptr;
}
const uint32_t *abh_cells_bucket_end(const abh_cells *c) {
  return c ? c->end.as<uint32_t>() : nullptr;
}
const double *abh_cells_positions(const abh_cells *c) { return c ? c->pos.as<double>() : nullptr; }

abh_status abh_gather(abh_context *ctx, const int32_t *order, int64_t n, const void *src,
                      void *dst, size_t elem_bytes) {
  if (!ctx || (n > 0 && (!order || !src || !dst)) || elem_bytes == 0) return ABH_ERR_INVALID_ARG;
  return abh::guard(ctx, [&] { abh::gather_bytes(ctx, order, n, src, dst, elem_bytes); });
}

}  // extern "C"
