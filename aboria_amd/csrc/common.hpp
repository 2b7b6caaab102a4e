// Shared host/device plumbing of libaboria_hip.so: error handling at the C
// ABI, device buffers, the context, and wavefront-level helpers for gfx950
// (64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <string>
#include <utility>

#include "../../include/aboria_hip.h"

namespace abh {

struct Error {
  abh_status status;
  std::string msg;
};

#define ABH_HIP(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw ::abh::Error{e_ == hipErrorOutOfMemory ? ABH_ERR_OUT_OF_MEMORY : ABH_ERR_HIP, \
                         std::string(#expr) + " failed: " + hipGetErrorString(e_)};   \
  } while (0)

#define ABH_REQUIRE(cond, status, message)                    \
  do {                                                        \
    if (!(cond)) throw ::abh::Error{status, std::string(message)}; \
  } while (0)

// Grow-only device buffer.  Growing frees the old allocation (hipFree
// synchronises the device, so pending kernels that used it have finished).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t want) {
    if (want <= bytes && p) return;
    release();
    size_t alloc = want + want / 4 + 256;
    ABH_HIP(hipMalloc(&p, alloc));
    bytes = alloc;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Synchronises a stream when it goes out of scope.  Declared after the host
// buffers of asynchronous copies, it is destroyed before them, so an error
// thrown while a copy is in flight does not free memory the copy still uses.
struct StreamDrain {
  hipStream_t st;
  StreamDrain(const StreamDrain &) = delete;
  StreamDrain &operator=(const StreamDrain &) = delete;
  ~StreamDrain() { (void)hipStreamSynchronize(st); }
};

}  // namespace abh

struct abh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  // temporaries owned by the device-wide primitives (one user at a time:
  // every primitive is enqueued on the one context stream)
  abh::DevBuf scan_blocks;                     // per-tile sums of scan_exclusive
  abh::DevBuf sort_keys, sort_vals, sort_hist;  // radix sort ping-pong + histograms
  abh::DevBuf dense_part;                       // column-split partial sums (abh_dense_apply)
  int n_cu = 256;
  // device address of the static scratch of abh_dense_apply_noalloc on this
  // context's device (hipGetSymbolAddress at context creation)
  double *dense_scratch = nullptr;
  // Handles created on this context keep it alive: abh_context_destroy only
  // marks it closed while any of them exists, the last one frees it.
  int refs = 0;
  bool closed = false;
};

namespace abh {

// Runs f() and converts internal errors to a status at the C ABI.
template <class F> abh_status guard(abh_context *ctx, F &&f) {
  try {
    if (ctx) ABH_HIP(hipSetDevice(ctx->device));
    f();
    return ABH_OK;
  } catch (const Error &e) {
    if (ctx) ctx->last_error = e.msg;
    return e.status;
  } catch (const std::bad_alloc &) {
    if (ctx) ctx->last_error = "host allocation failed";
    return ABH_ERR_OUT_OF_MEMORY;
  } catch (...) {
    if (ctx) ctx->last_error = "unknown internal error";
    return ABH_ERR_INVALID_ARG;
  }
}

inline void check_launch() { ABH_HIP(hipGetLastError()); }

void context_free(abh_context *ctx);  // cells.hip
inline void ctx_retain(abh_context *ctx) { ++ctx->refs; }
inline void ctx_release(abh_context *ctx) {
  if (--ctx->refs == 0 && ctx->closed) context_free(ctx);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Graph capture needs a stream other than the null stream: for the life of
// this object the context's work goes to a stream of its own (default
// priority, ordered after the caller's stream by an event, and the caller's
// stream after it on exit).
struct CaptureStream {
  abh_context *ctx = nullptr;
  hipStream_t orig = nullptr, mine = nullptr;
  hipEvent_t ev = nullptr;
  CaptureStream() = default;
  CaptureStream(const CaptureStream &) = delete;
  CaptureStream &operator=(const CaptureStream &) = delete;
  void begin(abh_context *c) {
    ctx = c;
    orig = c->stream;
    ABH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    ABH_HIP(hipStreamCreateWithFlags(&mine, hipStreamNonBlocking));
    ABH_HIP(hipEventRecord(ev, orig));
    ABH_HIP(hipStreamWaitEvent(mine, ev, 0));
    c->stream = mine;
  }
  ~CaptureStream() {
    if (!mine) {
      if (ev) (void)hipEventDestroy(ev);
      return;
    }
    (void)hipEventRecord(ev, mine);
    (void)hipStreamWaitEvent(orig, ev, 0);
    (void)hipStreamSynchronize(mine);
    ctx->stream = orig;
    (void)hipStreamDestroy(mine);
    (void)hipEventDestroy(ev);
  }
};

// Ends a stream capture left open by an exception.
struct CaptureGuard {
  hipStream_t s;
  bool open = false;
  ~CaptureGuard() {
    if (open) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
    }
  }
};

// -------------------------------------------------------------------------
// device helpers (wave64)
// -------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (uint64_t(1) << lane_id()) - 1;  // lane 0 -> 0
}

template <class T> __device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// XCD-contiguous block numbering: workgroups are dealt to the 8 XCDs round
// robin (block b runs on XCD b % 8), so logical block b % 8 * per + b / 8
// gives each XCD (and its L2) one contiguous range of the work; launch
// xcd_grid(count) blocks and skip logical blocks >= count.
__device__ __forceinline__ int xcd_block(int count) {
  const int per = (count + 7) / 8;
  return int(blockIdx.x % 8) * per + int(blockIdx.x / 8);
}
inline unsigned xcd_grid(int count) { return unsigned((count + 7) / 8 * 8); }

__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
  return uint32_t(__builtin_amdgcn_readfirstlane(int(v)));
}
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
  const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v))));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(uint64_t(v) >> 32))));
  return int64_t((uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ double uniform_f64(double v) {
  return __longlong_as_double(uniform_i64(__double_as_longlong(v)));
}

// -------------------------------------------------------------------------
// device-wide primitives (prims.hip)
// -------------------------------------------------------------------------

// out[i] = sum_{k<i} in[k]; *total_dev (device, may be null) = sum of all.
// in and out may alias.
void scan_exclusive(abh_context *ctx, const uint32_t *in, uint32_t *out, int64_t n,
                    uint32_t *total_dev);
void scan_exclusive(abh_context *ctx, const int64_t *in, int64_t *out, int64_t n,
                    int64_t *total_dev);

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the
// key.  Results are left in keys/vals; tmp buffers are taken from ctx.
void radix_sort_pairs(abh_context *ctx, uint32_t *keys, int32_t *vals, int64_t n, int bits);

void gather_bytes(abh_context *ctx, const int32_t *order, int64_t n, const void *src, void *dst,
                  size_t elem_bytes);

}  // namespace abh
