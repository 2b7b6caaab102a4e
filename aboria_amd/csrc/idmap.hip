// Find-by-id map of a particle set: the (id, container index) pairs sorted
// by id, and a batched lookup (neighbour_search_base's id map,
// src/NeighbourSearchBase.h:276-298 and :440-486, queried by
// CellListOrderedQuery::find, src/CellListOrdered.h:379-388).
//
// Build: ids are 64-bit; they are sorted by a stable LSD radix sort on the
// low 32 bits, then (when any id has high bits) on the high 32 bits, each
// pass over only the bits the ids use (their bitwise OR, reduced on the
// device).  Equal ids keep container order, so a lookup's lower bound finds
// the lowest index among them, as the reference's stable sort_by_key +
// lower_bound does.  Lookup: one thread per query id, a branch-free binary
// search over the sorted keys (HBM / L2 latency bound).
#include <algorithm>

#include "common.hpp"

namespace {

// lo[k] = low word of ids[k], vals[k] = k; OR of all low / high words
__global__ __launch_bounds__(256) void k_id_split(const int64_t *__restrict__ ids, int64_t n,
                                                  uint32_t *__restrict__ lo,
                                                  int32_t *__restrict__ vals,
                                                  uint32_t *__restrict__ bits_or) {
  uint32_t ol = 0, oh = 0;
  for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * 256) {
    const uint64_t v = uint64_t(ids[k]);
    lo[k] = uint32_t(v);
    vals[k] = int32_t(k);
    ol |= uint32_t(v);
    oh |= uint32_t(v >> 32);
  }
  // wave reduction, then one vector atomic per wave
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ol |= uint32_t(__shfl_xor(int(ol), o, 64));
    oh |= uint32_t(__shfl_xor(int(oh), o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicOr(&bits_or[0], ol);
    atomicOr(&bits_or[1], oh);
  }
}

// hi[k] = high word of ids[vals[k]]
__global__ __launch_bounds__(256) void k_id_high(const int64_t *__restrict__ ids, int64_t n,
                                                 const int32_t *__restrict__ vals,
                                                 uint32_t *__restrict__ hi) {
  const int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (k < n) hi[k] = uint32_t(uint64_t(ids[vals[k]]) >> 32);
}

// out[i] = vals[lower_bound(keys, q[i])] if that key equals q[i], else n
__global__ __launch_bounds__(256) void k_id_find(const int64_t *__restrict__ keys,
                                                 const int32_t *__restrict__ vals, int64_t n,
                                                 const int64_t *__restrict__ q, int64_t nq,
                                                 int32_t *__restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nq) return;
  const uint64_t id = uint64_t(q[i]);
  // lower bound: first position whose key is not below id
  int64_t first = 0, count = n;
  while (count > 0) {
    const int64_t half = count >> 1;
    const bool below = uint64_t(keys[first + half]) < id;
    first = below ? first + half + 1 : first;
    count = below ? count - half - 1 : half;
  }
  out[i] = (first < n && uint64_t(keys[first]) == id) ? vals[first] : int32_t(n);
}

int bit_length(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

}  // namespace

struct abh_id_map {
  abh_context *ctx = nullptr;
  int64_t n = 0;
  abh::DevBuf keys, vals, lo, bits;
};

namespace {

void id_map_build(abh_id_map *m, int64_t n, const int64_t *ids) {
  abh_context *ctx = m->ctx;
  hipStream_t st = ctx->stream;
  ABH_REQUIRE(n >= 0 && n < (int64_t(1) << 31), ABH_ERR_INVALID_ARG,
              "id map size must be in [0, 2^31)");
  ABH_REQUIRE(n == 0 || ids, ABH_ERR_INVALID_ARG, "ids is null");
  m->n = n;
  m->keys.ensure(size_t(n) * 8 + 64);
  m->vals.ensure(size_t(n) * 4 + 64);
  m->lo.ensure(size_t(n) * 4 + 64);
  m->bits.ensure(64);
  if (n == 0) return;
  uint32_t *lo = m->lo.as<uint32_t>();
  int32_t *vals = m->vals.as<int32_t>();
  uint32_t *bits_dev = m->bits.as<uint32_t>();
  ABH_HIP(hipMemsetAsync(bits_dev, 0, 8, st));
  const unsigned blocks = unsigned(std::min<int64_t>(abh::ceil_div(n, 256), 4096));
  k_id_split<<<blocks, 256, 0, st>>>(ids, n, lo, vals, bits_dev);
  abh::check_launch();
  uint32_t bits[2] = {0, 0};
  ABH_HIP(hipMemcpyAsync(bits, bits_dev, 8, hipMemcpyDeviceToHost, st));
  ABH_HIP(hipStreamSynchronize(st));
  abh::radix_sort_pairs(ctx, lo, vals, n, bit_length(bits[0]));
  if (bits[1]) {
    // stable pass on the high words after the low ones: sorted by (hi, lo)
    k_id_high<<<unsigned(abh::ceil_div(n, 256)), 256, 0, st>>>(ids, n, vals, lo);
    abh::check_launch();
    abh::radix_sort_pairs(ctx, lo, vals, n, bit_length(bits[1]));
  }
  abh::gather_bytes(ctx, vals, n, ids, m->keys.p, 8);
}

}  // namespace

extern "C" {

abh_status abh_id_map_create(abh_context *ctx, int64_t n, const int64_t *ids, abh_id_map **out) {
  if (!ctx || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_id_map *m = new (std::nothrow) abh_id_map();
  if (!m) return ABH_ERR_OUT_OF_MEMORY;
  m->ctx = ctx;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    id_map_build(m, n, ids);
  });
  if (st != ABH_OK) {
    delete m;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = m;
  return ABH_OK;
}

abh_status abh_id_map_update(abh_id_map *m, int64_t n, const int64_t *ids) {
  if (!m) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] { id_map_build(m, n, ids); });
}

abh_status abh_id_map_destroy(abh_id_map *m) {
  if (!m) return ABH_OK;
  abh_context *ctx = m->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete m;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_id_map_size(const abh_id_map *m, int64_t *n) {
  if (!m || !n) return ABH_ERR_INVALID_ARG;
  *n = m->n;
  return ABH_OK;
}

const int64_t *abh_id_map_keys(const abh_id_map *m) { return m ? m->keys.as<int64_t>() : nullptr; }
const int32_t *abh_id_map_values(const abh_id_map *m) {
  return m ? m->vals.as<int32_t>() : nullptr;
}

abh_status abh_id_map_find(const abh_id_map *m, int64_t nq, const int64_t *ids, int32_t *index) {
  if (!m) return ABH_ERR_INVALID_ARG;
  return abh::guard(m->ctx, [&] {
    ABH_REQUIRE(nq >= 0, ABH_ERR_INVALID_ARG, "negative query count");
    if (nq == 0) return;
    ABH_REQUIRE(ids && index, ABH_ERR_INVALID_ARG, "null pointer");
    k_id_find<<<unsigned(abh::ceil_div(nq, 256)), 256, 0, m->ctx->stream>>>(
        m->keys.as<int64_t>(), m->vals.as<int32_t>(), m->n, ids, nq, index);
    abh::check_launch();
  });
}

}  // extern "C"
