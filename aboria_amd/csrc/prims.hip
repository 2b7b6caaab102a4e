// Device-wide primitives used by the cell-list build and the search plans:
// exclusive scan (reduce / scan-of-tiles / downsweep, recursive on the tile
// sums), a stable LSD radix sort of (uint32 key, int32 value) pairs with
// 8-bit digits, and a byte-generic gather.  These stand in for the Thrust
// calls of src/detail/Algorithms.h (exclusive_scan :551-580, sort_by_key
// :173-206, gather :718-745).
#include "common.hpp"

namespace abh {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanRounds = 8;
constexpr int kScanTile = kScanThreads * kScanRounds;  // 2048

// Exclusive scan of one value per thread over a 256-thread block.
template <class T> __device__ T block_excl_scan(T v, T &block_total, T *wsum /* LDS[4] */) {
  const int wave = threadIdx.x >> 6;
  const T incl = wave_incl_scan(v);
  if (lane_id() == 63) wsum[wave] = incl;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    const T s = wsum[w];
    if (w < wave) off += s;
    tot += s;
  }
  __syncthreads();
  block_total = tot;
  return off + incl - v;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T *__restrict__ in, int64_t n,
                                                             T *__restrict__ sums) {
  __shared__ T ws[kScanThreads / 64];
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  T s = 0;
#pragma unroll
  for (int r = 0; r < kScanRounds; ++r) {
    const int64_t idx = base + r * kScanThreads + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  s = wave_sum(s);
  if (lane_id() == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) t += ws[w];
    sums[blockIdx.x] = t;
  }
}

// One block scans a whole array of at most kScanTile entries... or any size,
// looping with a carry (used only for n <= kScanTile).
template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_single(const T *in, T *out, int64_t n,
                                                             T *total) {
  __shared__ T ws[kScanThreads / 64];
  T carry = 0;
  for (int64_t base = 0; base < n; base += kScanThreads) {
    const int64_t idx = base + threadIdx.x;
    const T v = idx < n ? in[idx] : T(0);
    T tot;
    const T e = block_excl_scan(v, tot, ws);
    if (idx < n) out[idx] = carry + e;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const T *in, T *out, int64_t n,
                                                           const T *__restrict__ offs) {
  __shared__ T ws[kScanThreads / 64];
  T carry = offs[blockIdx.x];
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  for (int r = 0; r < kScanRounds; ++r) {
    const int64_t idx = base + r * kScanThreads + threadIdx.x;
    const T v = idx < n ? in[idx] : T(0);
    T tot;
    const T e = block_excl_scan(v, tot, ws);
    if (idx < n) out[idx] = carry + e;
    carry += tot;
  }
}

template <class T> size_t scan_tmp_elems(int64_t n) {
  size_t total = 0;
  while (n > kScanTile) {
    n = ceil_div(n, kScanTile);
    total += size_t(n);
  }
  return total;
}

template <class T>
void scan_rec(abh_context *ctx, const T *in, T *out, int64_t n, T *total, T *tmp) {
  if (n <= kScanTile) {
    k_scan_single<T><<<1, kScanThreads, 0, ctx->stream>>>(in, out, n, total);
    check_launch();
    return;
  }
  const int64_t nb = ceil_div(n, kScanTile);
  T *sums = tmp;
  k_scan_reduce<T><<<dim3(unsigned(nb)), kScanThreads, 0, ctx->stream>>>(in, n, sums);
  check_launch();
  scan_rec<T>(ctx, sums, sums, nb, total, tmp + nb);
  k_scan_down<T><<<dim3(unsigned(nb)), kScanThreads, 0, ctx->stream>>>(in, out, n, sums);
  check_launch();
}

template <class T>
void scan_impl(abh_context *ctx, const T *in, T *out, int64_t n, T *total) {
  if (n <= 0) {
    if (total) ABH_HIP(hipMemsetAsync(total, 0, sizeof(T), ctx->stream));
    return;
  }
  const size_t need = scan_tmp_elems<T>(n) * sizeof(T);
  ctx->scan_blocks.ensure(need + 64);
  scan_rec<T>(ctx, in, out, n, total, ctx->scan_blocks.as<T>());
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;  // 4096 keys per tile
constexpr int kBins = 256;

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t *__restrict__ keys,
                                                            int64_t n, int shift,
                                                            uint32_t *__restrict__ hist,
                                                            int ntiles) {
  __shared__ uint32_t h[kBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kSortTile;
  for (int r = 0; r < kSortRounds; ++r) {
    const int64_t idx = base + r * kSortThreads + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & (kBins - 1)], 1u);
  }
  __syncthreads();
  hist[int64_t(threadIdx.x) * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile through LDS.  The tile's keys and values are
// loaded once (16 per thread), counted per digit in LDS, and ranked in index
// order, 256 per round: within a round a wave finds the lanes sharing its
// digit with 8 ballots, and per-wave digit counts order the four waves.
// Each key goes to its place in the tile sorted by digit (in LDS); the tile
// is then written out in that order, so every digit's run is one
// contiguous store instead of one scattered word per key.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint32_t *__restrict__ kin, const int32_t *__restrict__ vin, uint32_t *__restrict__ kout,
    int32_t *__restrict__ vout, int64_t n, int shift, const uint32_t *__restrict__ offs,
    int ntiles) {
  __shared__ uint32_t gbase[kBins];  // global offset of each digit's run of this tile
  __shared__ uint32_t lbase[kBins];  // offset of each digit's run within the tile
  __shared__ uint32_t run[kBins];    // keys of each digit placed so far
  __shared__ uint32_t wcount[kSortThreads / 64][kBins];
  __shared__ uint32_t wsum[kSortThreads / 64];
  __shared__ uint32_t sk[kSortTile];
  __shared__ int32_t sv[kSortTile];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int64_t tile0 = int64_t(blockIdx.x) * kSortTile;
  const int tn = int(min<int64_t>(kSortTile, n - tile0));
  uint32_t key[kSortRounds];
  int32_t val[kSortRounds];
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r) {
    const int e = r * kSortThreads + tid;
    key[r] = e < tn ? kin[tile0 + e] : 0u;
    val[r] = e < tn ? vin[tile0 + e] : 0;
  }
  run[tid] = 0;
  gbase[tid] = offs[int64_t(tid) * ntiles + blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r)
    if (r * kSortThreads + tid < tn) atomicAdd(&run[(key[r] >> shift) & (kBins - 1)], 1u);
  __syncthreads();
  // exclusive scan of the tile's digit counts
  const uint32_t c = run[tid];
  const uint32_t incl = wave_incl_scan(c);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (int w = 0; w < wave; ++w) before += wsum[w];
  lbase[tid] = before + incl - c;
  run[tid] = 0;
  __syncthreads();
  for (int r = 0; r < kSortRounds; ++r) {
    if (r * kSortThreads >= tn) break;  // uniform across the block
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; ++w) wcount[w][tid] = 0;
    __syncthreads();
    const bool valid = r * kSortThreads + tid < tn;
    const uint32_t digit = (key[r] >> shift) & (kBins - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (digit >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = __popcll(peers & lanemask_lt());
    if (valid && rank == 0) wcount[wave][digit] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = lbase[digit] + run[digit] + rank;
      for (int w = 0; w < wave; ++w) pos += wcount[w][digit];
      sk[pos] = key[r];
      sv[pos] = val[r];
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; ++w) add += wcount[w][tid];
    run[tid] += add;
    __syncthreads();
  }
  for (int e = tid; e < tn; e += kSortThreads) {
    const uint32_t k = sk[e];
    const uint32_t d = (k >> shift) & (kBins - 1);
    const uint32_t pos = gbase[d] + uint32_t(e) - lbase[d];
    kout[pos] = k;
    vout[pos] = sv[e];
  }
}

template <int W>
__global__ void k_gather_words(const int32_t *__restrict__ order, int64_t n,
                               const uint32_t *__restrict__ src, uint32_t *__restrict__ dst) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t total = n * W;
  if (t >= total) return;
  const int64_t k = t / W;
  const int w = int(t - k * W);
  dst[t] = src[int64_t(order[k]) * W + w];
}

__global__ void k_gather_generic(const int32_t *__restrict__ order, int64_t n,
                                 const uint8_t *__restrict__ src, uint8_t *__restrict__ dst,
                                 int64_t eb) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * eb) return;
  const int64_t k = t / eb;
  const int64_t w = t - k * eb;
  dst[t] = src[int64_t(order[k]) * eb + w];
}

}  // namespace

void scan_exclusive(abh_context *ctx, const uint32_t *in, uint32_t *out, int64_t n,
                    uint32_t *total) {
  scan_impl<uint32_t>(ctx, in, out, n, total);
}
void scan_exclusive(abh_context *ctx, const int64_t *in, int64_t *out, int64_t n,
                    int64_t *total) {
  scan_impl<int64_t>(ctx, in, out, n, total);
}

void radix_sort_pairs(abh_context *ctx, uint32_t *keys, int32_t *vals, int64_t n, int bits) {
  if (n <= 1 || bits <= 0) return;
  const int passes = (bits + 7) / 8;
  const int ntiles = int(ceil_div(n, kSortTile));
  ctx->sort_keys.ensure(size_t(n) * 4);
  ctx->sort_vals.ensure(size_t(n) * 4);
  ctx->sort_hist.ensure(size_t(ntiles) * kBins * 4);
  uint32_t *k0 = keys, *k1 = ctx->sort_keys.as<uint32_t>();
  int32_t *v0 = vals, *v1 = ctx->sort_vals.as<int32_t>();
  uint32_t *hist = ctx->sort_hist.as<uint32_t>();
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    k_sort_hist<<<ntiles, kSortThreads, 0, ctx->stream>>>(k0, n, shift, hist, ntiles);
    check_launch();
    scan_exclusive(ctx, hist, hist, int64_t(ntiles) * kBins, nullptr);
    k_sort_scatter<<<ntiles, kSortThreads, 0, ctx->stream>>>(k0, v0, k1, v1, n, shift, hist,
                                                             ntiles);
    check_launch();
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  if (k0 != keys) {
    ABH_HIP(hipMemcpyAsync(keys, k0, size_t(n) * 4, hipMemcpyDeviceToDevice, ctx->stream));
    ABH_HIP(hipMemcpyAsync(vals, v0, size_t(n) * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
}

void gather_bytes(abh_context *ctx, const int32_t *order, int64_t n, const void *src, void *dst,
                  size_t eb) {
  if (n <= 0) return;
  const int threads = 256;
  if (eb % 4 == 0 && (uintptr_t(src) % 4 == 0) && (uintptr_t(dst) % 4 == 0)) {
    const int64_t W = int64_t(eb / 4);
    const int64_t total = n * W;
    const unsigned blocks = unsigned(ceil_div(total, threads));
    switch (W) {
      case 1: k_gather_words<1><<<blocks, threads, 0, ctx->stream>>>(order, n, (const uint32_t *)src, (uint32_t *)dst); break;
      case 2: k_gather_words<2><<<blocks, threads, 0, ctx->stream>>>(order, n, (const uint32_t *)src, (uint32_t *)dst); break;
      case 4: k_gather_words<4><<<blocks, threads, 0, ctx->stream>>>(order, n, (const uint32_t *)src, (uint32_t *)dst); break;
      case 6: k_gather_words<6><<<blocks, threads, 0, ctx->stream>>>(order, n, (const uint32_t *)src, (uint32_t *)dst); break;
      default:
        k_gather_generic<<<unsigned(ceil_div(n * int64_t(eb), threads)), threads, 0, ctx->stream>>>(
            order, n, (const uint8_t *)src, (uint8_t *)dst, int64_t(eb));
    }
  } else {
    k_gather_generic<<<unsigned(ceil_div(n * int64_t(eb), threads)), threads, 0, ctx->stream>>>(
        order, n, (const uint8_t *)src, (uint8_t *)dst, int64_t(eb));
  }
  check_launch();
}

}  // namespace abh
