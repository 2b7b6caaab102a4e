// Radius search and sparse kernel operator over a CellListOrdered.
//
// Work decomposition (MI355X): queries are grouped by home cell; one
// wavefront takes up to 64 queries of one home cell, one query per lane.  The
// candidate buckets are the (2w+1)^D stencil around the home cell (w = r /
// side rounded up), each mapped to a bucket and a periodic image.  Stencil
// slots are visited in ascending bucket order, so candidates stream through
// the wave in ascending container index with wave-uniform addresses (scalar
// loads, one broadcast per candidate) and every lane tests its own query.
//
// Exactness: for every (query, slot) the lane evaluates the predicate "the
// reference's search iterator visits this bucket under this image, and how
// many times" -- the quadrant starts, clamping and box-distance pruning of
// lattice_iterator_within_distance (src/NeighbourSearchBase.h:1860-2008)
// evaluated with the same FP64 expressions -- and accepts a candidate iff
// that multiplicity is non-zero and sum_d dx_d^2 <= r^2 (src/Search.h:438-465).
// Each lane also checks that the buckets the reference would visit all lie in
// the stencil; if any does not (possible only within rounding of a cell face)
// the host widens the stencil and repeats.  The neighbour multiset of every
// query therefore equals the reference's.
#include <cmath>
#include <vector>

#include "cells.hpp"
#include "kernel_fn.hpp"
#include "solvers.hpp"

using abh::DevBuf;
using abh::GridParams;
using abh::KernelDesc;
using abh::kb_add;
using abh::phi32;
using abh::phi64;

namespace {

constexpr int kThreads = 256;
constexpr int kMaxSlots = 343;
constexpr int kGroup = 64;  // queries per work item = lanes per wave

struct WorkItem {
  int32_t home[3];
  int32_t start;  // first query in sorted query order
  int32_t count;  // <= 64
  int32_t pad;
};

struct Slot {
  int32_t b;    // collapsed bucket index, -1 = no bucket
  int32_t beg;  // bucket_begin
  int32_t len;  // bucket_end - bucket_begin
  int32_t img;  // image digits (I_d + 1) packed two bits per dimension
};

// Virtual home cell of a centre in dimension d: floor((q - lo) * inv),
// clamped to [-1, size] for a non-periodic dimension and to [-size, 2*size)
// for a periodic one (centres up to one domain length outside still reach
// the domain through the +-1 images).  NaN maps to the lowest value.
__device__ __forceinline__ int home_lo(const GridParams &g, int d) {
  return g.periodic[d] ? -g.size[d] : -1;
}
__device__ __forceinline__ int home_span(const GridParams &g, int d) {
  return g.periodic[d] ? 3 * g.size[d] : g.size[d] + 2;
}
template <int D> __device__ __forceinline__ int home_of(const GridParams &g, const double *q, int d) {
  double t = floor((q[d] - g.lo[d]) * g.inv[d]);
  const double lo = double(home_lo(g, d));
  const double hi = lo + double(home_span(g, d) - 1);
  if (!(t >= lo)) t = lo;
  if (t > hi) t = hi;
  return int(t);
}

// Sort key of a centre: its virtual home cell, then (levels > 0) the Morton
// code of its sub-cell on a 2^levels lattice inside the home cell, so that
// consecutive centres of a crowded cell are close together (the radius
// test culls candidates against the bounding box of a work item).
template <int D>
__global__ __launch_bounds__(kThreads) void k_home_keys(const double *__restrict__ q, int64_t n,
                                                        GridParams g, int levels,
                                                        uint32_t *__restrict__ keys,
                                                        int32_t *__restrict__ vals) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  double p[D];
#pragma unroll
  for (int d = 0; d < D; ++d) p[d] = q[i * D + d];
  uint32_t key = 0, code = 0;
  int h[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    h[d] = home_of<D>(g, p, d);
    key = key * uint32_t(home_span(g, d)) + uint32_t(h[d] - home_lo(g, d));
  }
  if (levels > 0) {
    const int side = 1 << levels;
    int sub[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double f = ((p[d] - g.lo[d]) * g.inv[d] - double(h[d])) * double(side);
      int v = f >= 0.0 ? int(f) : 0;  // NaN and rounding below the cell -> 0
      sub[d] = v < side ? v : side - 1;
    }
    for (int l = levels - 1; l >= 0; --l)
#pragma unroll
      for (int d = 0; d < D; ++d) code = (code << 1) | uint32_t((sub[d] >> l) & 1);
    key = (key << (levels * D)) | code;
  }
  keys[i] = key;
  if (vals) vals[i] = int32_t(i);
}

// runs of equal home cell (the key without its sub-cell bits)
__global__ __launch_bounds__(kThreads) void k_run_flags(const uint32_t *__restrict__ keys,
                                                        int64_t n, int shift,
                                                        uint32_t *__restrict__ flags) {
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n) return;
  flags[k] = (k == 0 || (keys[k] >> shift) != (keys[k - 1] >> shift)) ? 1u : 0u;
}

__global__ __launch_bounds__(kThreads) void k_run_starts(const uint32_t *__restrict__ flags,
                                                         const uint32_t *__restrict__ ridx,
                                                         int64_t n,
                                                         uint32_t *__restrict__ run_start) {
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n || !flags[k]) return;
  run_start[ridx[k]] = uint32_t(k);
}

__global__ __launch_bounds__(kThreads) void k_run_chunks(const uint32_t *__restrict__ run_start,
                                                         const uint32_t *__restrict__ nruns_dev,
                                                         int64_t n, uint32_t *__restrict__ chunks) {
  const int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (r >= n) return;
  const int64_t nruns = int64_t(*nruns_dev);
  if (r >= nruns) {
    chunks[r] = 0;
    return;
  }
  const int64_t s = run_start[r];
  const int64_t e = (r + 1 < nruns) ? int64_t(run_start[r + 1]) : n;
  chunks[r] = uint32_t((e - s + kGroup - 1) / kGroup);
}

template <int D>
__global__ __launch_bounds__(kThreads) void k_fill_work(
    const uint32_t *__restrict__ run_start, const uint32_t *__restrict__ nruns_dev, int64_t n,
    const uint32_t *__restrict__ choff, const uint32_t *__restrict__ chunks,
    const double *__restrict__ q_sorted, GridParams g, WorkItem *__restrict__ work) {
  const int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t nruns = int64_t(*nruns_dev);
  if (r >= nruns) return;
  const int64_t s = run_start[r];
  const int64_t e = (r + 1 < nruns) ? int64_t(run_start[r + 1]) : n;
  double p[D];
#pragma unroll
  for (int d = 0; d < D; ++d) p[d] = q_sorted[s * D + d];
  WorkItem w{};
#pragma unroll
  for (int d = 0; d < 3; ++d) w.home[d] = d < D ? home_of<D>(g, p, d) : 0;
  const uint32_t nc = chunks[r];
  for (uint32_t c = 0; c < nc; ++c) {
    w.start = int32_t(s + int64_t(c) * kGroup);
    const int64_t left = e - int64_t(w.start);
    w.count = int32_t(left < kGroup ? left : kGroup);
    work[choff[r] + c] = w;
  }
}

// ---------------------------------------------------------------------------
// reference predicates
// ---------------------------------------------------------------------------

// (x + 0.5)*side + lo as find_bucket_centre, then max(|c - q| - half, 0)^2
__device__ __forceinline__ double box_term(const GridParams &g, int d, int x, double q) {
  const double centre = (double(x) + 0.5) * g.side[d] + g.lo[d];
  double t = fabs(centre - q) - g.half[d];
  t = t < 0.0 ? 0.0 : t;  // std::max(t, 0.0) keeps t unless t < 0
  return t * t;
}

// Per-lane quadrant state for one image point (reset_min_and_index,
// src/NeighbourSearchBase.h:1860-1932).
template <int D> struct QuadState {
  int start[1 << D][D];
  uint32_t valid;  // bit Q set = quadrant Q non-empty
};

template <int D>
__device__ __forceinline__ bool outside_domain(const GridParams &g, const double *q, double r2) {
  double acc = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double mid = 0.5 * (g.lo[d] + g.hi[d]);
    const double halfL = 0.5 * g.L[d];
    double t = fabs(mid - q[d]) - halfL;
    t = t < 0.0 ? 0.0 : t;
    acc = acc + t * t;
  }
  return acc > r2;
}

template <int D>
__device__ __forceinline__ void quad_state(const GridParams &g, const double *q, double r2,
                                           QuadState<D> &qs) {
  int mu[D], md[D];
  double gu[D], gd[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    mu[d] = int(floor((q[d] + 0.5 * g.side[d] - g.lo[d]) * g.inv[d]));
    md[d] = int(floor((q[d] + (-0.5) * g.side[d] - g.lo[d]) * g.inv[d]));
    gu[d] = box_term(g, d, mu[d], q[d]);
    gd[d] = box_term(g, d, md[d], q[d]);
  }
  qs.valid = 0;
#pragma unroll
  for (int Q = 0; Q < (1 << D); ++Q) {
    double acc = 0.0;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const bool up = (Q >> d) & 1;
      acc = acc + (up ? gu[d] : gd[d]);
    }
    ok = !(acc > r2);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const bool up = (Q >> d) & 1;
      int s;
      if (up) {
        s = mu[d] < 0 ? 0 : mu[d];
        if (mu[d] > g.end_bucket[d]) ok = false;
      } else {
        s = md[d] > g.end_bucket[d] ? g.end_bucket[d] : md[d];
        if (md[d] < 0) ok = false;
      }
      qs.start[Q][d] = s;
    }
    if (ok) qs.valid |= 1u << Q;
  }
}

// Number of times the reference visits bucket b for this image point.
template <int D>
__device__ __forceinline__ int visit_count(const GridParams &g, const double *q, double r2,
                                           const QuadState<D> &qs, const int *b) {
  if (!qs.valid) return 0;
  double dist = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) dist = dist + box_term(g, d, b[d], q[d]);
  const bool near = !(dist > r2);
  int m = 0;
#pragma unroll
  for (int Q = 0; Q < (1 << D); ++Q) {
    if (!((qs.valid >> Q) & 1u)) continue;
    bool in = true, is_start = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const bool up = (Q >> d) & 1;
      const int s = qs.start[Q][d];
      in = in && (up ? b[d] >= s : b[d] <= s);
      is_start = is_start && (b[d] == s);
    }
    if (in && (is_start || near)) ++m;
  }
  return m;
}

// True if the buckets visited for this image point leave the stencil
// [c_d - w_d, c_d + w_d] (bucket coordinates of the image).
template <int D>
__device__ __forceinline__ bool escapes_stencil(const GridParams &g, const double *q, double r2,
                                                const QuadState<D> &qs, const int *c,
                                                const int *w) {
#pragma unroll
  for (int Q = 0; Q < (1 << D); ++Q) {
    if (!((qs.valid >> Q) & 1u)) continue;
    int s[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      s[d] = qs.start[Q][d];
      if (s[d] < c[d] - w[d] || s[d] > c[d] + w[d]) return true;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const bool up = (Q >> d) & 1;
      const int e = up ? c[d] + w[d] + 1 : c[d] - w[d] - 1;
      if (e < 0 || e > g.end_bucket[d]) continue;
      if (up ? e < s[d] : e > s[d]) continue;
      double dist = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) dist = dist + box_term(g, k, k == d ? e : s[k], q[k]);
      if (!(dist > r2)) return true;
    }
  }
  return false;
}

// ---------------------------------------------------------------------------
// kernel profiles
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// the search kernel
// ---------------------------------------------------------------------------

enum Mode {
  kCount = 0,
  kFill = 1,
  kMatvec64 = 2,
  kMatvec32 = 3,
  kAssemble = 4,
  kCapacity = 5,   // upper bound of every row: candidates of its visited buckets
  kFillPadded = 6  // fill rows at capacity offsets and record their counts
};
constexpr int kMaxRhs = 8;

struct SearchArgs {
  const WorkItem *work;
  const uint32_t *nwork;
  const double *qpos;   // sorted query positions
  const int32_t *qperm; // sorted query -> original row (null = identity)
  const uint32_t *bbeg, *bend;
  const double *pos;    // particle positions in container order
  double r2;
  int w[3];
  int nslots;
  // count / fill
  int64_t *counts;
  const int64_t *row_ptr;
  int32_t *col;
  double *val;   // assemble: kernel value of every stored pair
  float *val32;  // assemble (optional): the same value as the FP32 matvec forms it
  uint32_t *escape_flag;
  uint32_t *next;  // dynamic work counter, zero at launch
  int ring;        // fill: stage each row's entries in LDS and store whole
                   // 64-byte segments (col is 64-byte aligned)
  // matvec
  KernelDesc k;
  int nrhs;
  const double *x;
  int64_t ldx;
  double *y;
  int64_t ldy;
};

#ifndef ABH_STAGE_UNROLL
#define ABH_STAGE_UNROLL 4
#endif
// staged candidates tested per step of the count / list passes
constexpr int kStageUnroll = ABH_STAGE_UNROLL;

// A staged candidate: coordinates (unused ones 0) and container index, read
// back as two 16-byte LDS loads.
struct __align__(16) StageRec {
  double x, y, z;
  int32_t j;
  int32_t pad;
};

__device__ __forceinline__ void load_stage(const StageRec &r, double (&e)[3], int &j) {
  const double2 xy = *reinterpret_cast<const double2 *>(&r.x);
  const double2 zj = *reinterpret_cast<const double2 *>(&r.z);
  e[0] = xy.x;
  e[1] = xy.y;
  e[2] = zj.x;
  j = int(uint32_t(uint64_t(__double_as_longlong(zj.y))));
}

// Fill output ring: the last 32 entries of each lane's row (two 16-entry,
// 64-byte segments), slot-major ([slot][lane]) so that the lanes' appends
// and segment reads fall in distinct LDS banks whatever their positions.
constexpr int kRing = 32;
constexpr int kSeg = 16;



// Writes v to this lane's word of the ring slot whose number k8 holds in
// bits 8..12 (slot-major ring, 256 bytes per slot).
static_assert(kRing == 32 && kGroup == 64, "ring_put assumes 32 slots of 64 lanes");
__device__ __forceinline__ void ring_put(uint32_t *ring, uint32_t k8, int lane, uint32_t v) {
  *reinterpret_cast<uint32_t *>(reinterpret_cast<unsigned char *>(ring) +
                                ((k8 & 0x1f00u) | (uint32_t(lane) << 2))) = v;
}

// a wave-uniform double moved to scalar registers
__device__ __forceinline__ double uniform_f64(double v) {
  const uint64_t b = uint64_t(__double_as_longlong(v));
  const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(b))));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(b >> 32))));
  return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

__device__ __forceinline__ double wave_min_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// q + I*L for the image digit I + 1 (the product is exact: I is -1, 0, +1)
__device__ __forceinline__ double image_of(int digit, double q, double L) {
  return digit == 1 ? q : q + double(digit - 1) * L;
}

__device__ __forceinline__ double pick3(int digit, double m, double z, double p) {
  return digit == 0 ? m : (digit == 1 ? z : p);
}

// Action of the count and list passes on one tested pair (candidate j of a
// lane's centre, squared distance acc): count it m times, or append j m
// times (with the kernel value when assembling).  single: no lane of the
// wave visits the bucket more than once, so a hit appends one entry.
template <int MODE>
__device__ __forceinline__ void staged_hit(const SearchArgs &a, bool single, bool hit, double acc,
                                           int j, int m, double wrow, int64_t out, int &cnt) {
  if (MODE == kCount) {
    if (single) cnt += hit ? 1 : 0;
    else cnt += hit ? m : 0;
  } else if (MODE == kFill || MODE == kFillPadded) {
    if (single) {
      if (hit) a.col[out + cnt] = j;
      cnt += hit ? 1 : 0;
    } else if (hit) {
      for (int rep = 0; rep < m; ++rep) a.col[out + cnt + rep] = j;
      cnt += m;
    }
  } else if (MODE == kAssemble) {
    if (hit) {
      double w = wrow;
      if (a.k.col_s) w = w + a.k.wc * a.k.col_s[j];
      const double f = phi64(a.k, acc) * w;
      const float f32 = phi32(a.k, float(acc)) * float(w);
      for (int rep = 0; rep < m; ++rep) {
        a.col[out + cnt + rep] = j;
        if (a.val) a.val[out + cnt + rep] = f;
        if (a.val32) a.val32[out + cnt + rep] = f32;
      }
      cnt += m;
    }
  }
}

template <int D, int MODE>
__global__ __launch_bounds__(kGroup) void k_search(SearchArgs a, GridParams g) {
  extern __shared__ __align__(16) unsigned char lds[];
  constexpr bool kMV = (MODE == kMatvec64 || MODE == kMatvec32);
  const int S = a.nslots;
  // fill passes: the output ring, [32][64] words in static LDS (a lane's
  // ring address is a constant plus (slot << 8) | (lane << 2))
  constexpr bool kFillMode = (MODE == kFill || MODE == kFillPadded);
  const bool use_ring = kFillMode && a.ring;
  uint32_t *ring = nullptr;
  if constexpr (kFillMode) {
    __shared__ __align__(16) uint32_t ring_lds[kRing * kGroup];
    ring = ring_lds;
  }
  Slot *slots = reinterpret_cast<Slot *>(lds);
  int32_t *sorder = reinterpret_cast<int32_t *>(slots + S);
  uint8_t *mult = reinterpret_cast<uint8_t *>(sorder + S);  // [S][64]
  // count / list passes: staged candidates (coordinates [D][64], index [64])
  StageRec *st = reinterpret_cast<StageRec *>(
      lds + ((size_t(S) * (sizeof(Slot) + 4 + kGroup) + 4 * size_t(S + 1) + 15) & ~size_t(15)));
  double *xs = reinterpret_cast<double *>(st + kGroup);  // [64][kMaxRhs], matvec only
  const int lane = threadIdx.x;
  const int nwork = int(*a.nwork);
  int wd[3] = {a.w[0], a.w[1], a.w[2]};
  int span[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) span[d] = 2 * wd[d] + 1;

  // work items are taken dynamically: their cost varies by orders of
  // magnitude between sparse and dense cells
  for (;;) {
    int wi = 0;
    if (lane == 0) wi = int(atomicAdd(a.next, 1u));
    wi = __shfl(wi, 0, 64);
    if (wi >= nwork) break;
    const WorkItem it = a.work[wi];
    const int qk = it.start + (lane < it.count ? lane : 0);
    double q[D];
    bool finite = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      q[d] = a.qpos[int64_t(qk) * D + d];
      finite = finite && isfinite(q[d]);
    }
    // a centre with a non-finite coordinate has no neighbours
    const bool active = lane < it.count && finite;
    const bool has_row = lane < it.count;
    // box of the work item's active centres (empty when none is active)
    double qlo[3] = {0.0, 0.0, 0.0}, qhi[3] = {0.0, 0.0, 0.0};
    if (MODE != kCapacity) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        qlo[d] = uniform_f64(wave_min_f64(active ? q[d] : INFINITY));
        qhi[d] = uniform_f64(wave_max_f64(active ? q[d] : -INFINITY));
      }
    }
    const int64_t row = a.qperm ? int64_t(a.qperm[qk]) : int64_t(qk);
    // image points q + I*L for I = -1, 0, +1 (search_iterator,
    // src/Search.h:188-190), formed where needed by image_of

    // stencil slots: virtual cell v = home + o -> bucket b under image I
    for (int s = lane; s < S; s += kGroup) {
      int o[3] = {0, 0, 0};
      int rem = s;
#pragma unroll
      for (int d = D - 1; d >= 0; --d) {
        o[d] = rem % span[d] - wd[d];
        rem /= span[d];
      }
      bool ok = true;
      int bv[3] = {0, 0, 0};
      int imgc = 0, p3 = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int v = it.home[d] + o[d];
        const int n = g.size[d];
        int I = 0, b = v;
        if (v < 0) { b = v + n; I = 1; }
        else if (v >= n) { b = v - n; I = -1; }
        if (I != 0 && !g.periodic[d]) ok = false;
        if (b < 0 || b >= n) ok = false;  // would need an image beyond +-1
        bv[d] = b;
        imgc |= (I + 1) << (2 * d);
        (void)p3;
      }
      if (ok) {
        uint32_t index = 0, m = 1;
#pragma unroll
        for (int i = D - 1; i >= 0; --i) {
          if (i != D - 1) m *= uint32_t(g.size[i + 1]);
          index += m * uint32_t(bv[i]);
        }
        const int bidx = int(index);
        const int beg = int(a.bbeg[bidx]);
        const int end = int(a.bend[bidx]);
        slots[s] = Slot{bidx, beg, end - beg, imgc};
      } else {
        slots[s] = Slot{-1, 0, 0, imgc};
      }
    }
    __syncthreads();
    // ascending bucket order (ties by slot index)
    for (int s = lane; s < S; s += kGroup) {
      const int bs = slots[s].b;
      int rank = 0;
      for (int t = 0; t < S; ++t) {
        const int bt = slots[t].b;
        rank += (bt < bs || (bt == bs && t < s)) ? 1 : 0;
      }
      sorder[rank] = s;
    }
    // per-lane visit multiplicities, one image at a time
    bool escape = false;
    for (int s = 0; s < S; ++s) mult[s * kGroup + lane] = 0;
    {
      int nimg = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) nimg *= 3;
      for (int ii = 0; ii < nimg; ++ii) {
        int I[D];
        bool img_ok = true;
        int t = ii;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          I[d] = t % 3 - 1;
          t /= 3;
          if (I[d] != 0 && !g.periodic[d]) img_ok = false;
        }
        if (!img_ok) continue;  // uniform
        int pk = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) pk |= (I[d] + 1) << (2 * d);
        double qi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) qi[d] = image_of(I[d] + 1, q[d], g.L[d]);
        const bool in_range = active && !outside_domain<D>(g, qi, a.r2);
        if (__ballot(in_range) == 0) continue;
        QuadState<D> qs;
        if (in_range) {
          quad_state<D>(g, qi, a.r2, qs);
        } else {
          qs.valid = 0;
        }
        if ((MODE == kCount || MODE == kCapacity) && in_range) {
          int c[D];
#pragma unroll
          for (int d = 0; d < D; ++d) c[d] = it.home[d] + I[d] * g.size[d];
          if (escapes_stencil<D>(g, qi, a.r2, qs, c, wd)) escape = true;
        }
        for (int s = 0; s < S; ++s) {
          const Slot sl = slots[s];
          if (sl.img != pk || sl.b < 0) continue;
          int b[D];
          int rem = sl.b;
#pragma unroll
          for (int d = D - 1; d >= 0; --d) {
            b[d] = rem % g.size[d];
            rem /= g.size[d];
          }
          const int m = in_range ? visit_count<D>(g, qi, a.r2, qs, b) : 0;
          mult[s * kGroup + lane] = uint8_t(m);
        }
      }
    }
    if ((MODE == kCount || MODE == kCapacity) && escape) atomicOr(a.escape_flag, 1u);
    __syncthreads();
    if (MODE == kCapacity) {
      // candidates of every visited bucket, with multiplicity: no distance
      // test can accept more
      int64_t cap = 0;
      for (int s = 0; s < S; ++s) cap += int64_t(mult[s * kGroup + lane]) * slots[s].len;
      // rounded up to 16 entries: every padded row then starts on a 64-byte
      // boundary, where the fill's ring stores whole segments from the first
      if (has_row) a.counts[row] = active ? ((cap + 15) & ~int64_t(15)) : 0;
      __syncthreads();
      continue;
    }
    int cnt = 0;
    int64_t out = 0;
    constexpr bool kList = (MODE == kFill || MODE == kAssemble || MODE == kFillPadded);
    if (kList && active) out = a.row_ptr[row];
    // ring state: entries [fpos, out + cnt) of this lane's row are staged
    int64_t fpos = out;
    uint32_t *rl = ring + lane;  // slot k of this lane: rl[k * kGroup]
    // store the segment holding fpos once it is complete (64-byte vector
    // stores when it starts at its segment boundary, the row's first,
    // partial segment entry by entry)
    auto ring_segment = [&]() {
      const int64_t pos = out + cnt;
      const int64_t seg = fpos & ~int64_t(kSeg - 1);
      if (pos >= seg + kSeg) {
        if (fpos == seg) {
          const uint32_t *src = rl + int(seg & (kRing - 1)) * kGroup;
          uint32_t v[kSeg];
#pragma unroll
          for (int k = 0; k < kSeg; ++k) v[k] = src[k * kGroup];
          uint4 *dst = reinterpret_cast<uint4 *>(a.col + seg);
#pragma unroll
          for (int k = 0; k < kSeg / 4; ++k)
            dst[k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
        } else {
          for (int64_t e = fpos; e < seg + kSeg; ++e)
            a.col[e] = int32_t(rl[int(e & (kRing - 1)) * kGroup]);
        }
        fpos = seg + kSeg;
      }
    };
    // store everything staged (row end, or before a bucket written directly)
    auto ring_drain = [&]() {
      const int64_t pos = out + cnt;
      for (int64_t e = fpos; e < pos; ++e) a.col[e] = int32_t(rl[int(e & (kRing - 1)) * kGroup]);
      fpos = pos;
    };
    double acc64[kMaxRhs];
    float acc32[kMaxRhs], cmp32[kMaxRhs];
    if (kMV) {
#pragma unroll
      for (int r = 0; r < kMaxRhs; ++r) {
        acc64[r] = (has_row && r < a.nrhs) ? a.y[int64_t(r) * a.ldy + row] : 0.0;
        acc32[r] = 0.0f;
        cmp32[r] = 0.0f;
      }
    }
    double wrow = 1.0;
    if ((kMV || MODE == kAssemble) && (a.k.row_s || a.k.col_s)) {
      wrow = a.k.w0;
      if (a.k.row_s && active) wrow = wrow + a.k.wr * a.k.row_s[row];
    }

    for (int k = 0; k < S; ++k) {
      const int s = abh::uniform_i32(sorder[k]);
      const int len = abh::uniform_i32(slots[s].len);
      if (len <= 0) continue;
      const int m = active ? int(mult[s * kGroup + lane]) : 0;
      if (__ballot(m != 0) == 0) continue;  // no lane visits this bucket
      const int beg = abh::uniform_i32(slots[s].beg);
      const int img = abh::uniform_i32(slots[s].img);
      double qi[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int d = 0; d < D; ++d) qi[d] = image_of((img >> (2 * d)) & 3, q[d], g.L[d]);
      const int end = beg + len;

      if (MODE == kCount || kList) {
        // Candidates of this bucket in blocks of 64: each lane loads one and
        // keeps it if it lies within r of the work item's query box (the
        // same FP64 operations as the radius test, monotone in the query, so
        // no accepted pair is dropped); the kept ones are compacted into LDS
        // in index order and tested against every lane's centre.  acc <= r2
        // equals the reference's !(acc > r2) for every finite centre (acc is
        // never NaN there); lanes with m == 0 compare against -1.
        double blo[3] = {0.0, 0.0, 0.0}, bhi[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const int dg = (img >> (2 * d)) & 3;
          blo[d] = uniform_f64(
              pick3(dg, qlo[d] + double(-1) * g.L[d], qlo[d], qlo[d] + double(1) * g.L[d]));
          bhi[d] = uniform_f64(
              pick3(dg, qhi[d] + double(-1) * g.L[d], qhi[d], qhi[d] + double(1) * g.L[d]));
        }
        const double r2m = m != 0 ? a.r2 : -1.0;
        const bool single = __ballot(m > 1) == 0;
        for (int c0 = beg; c0 < end; c0 += kGroup) {
          const int jc = c0 + lane;
          bool keep = jc < end;
          double x[3] = {0.0, 0.0, 0.0};
          if (keep) {
#pragma unroll
            for (int d = 0; d < D; ++d) x[d] = a.pos[int64_t(jc) * D + d];
          }
          double dd = 0.0;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            double e = 0.0;
            if (x[d] < blo[d]) e = blo[d] - x[d];
            else if (x[d] > bhi[d]) e = x[d] - bhi[d];
            dd = dd + e * e;
          }
          keep = keep && dd <= a.r2;
          const uint64_t km = __ballot(keep);
          if (km == 0) continue;
          const int nk = __popcll(km);
          if (keep) {
            const int sl = int(__builtin_amdgcn_mbcnt_hi(uint32_t(km >> 32),
                                                         __builtin_amdgcn_mbcnt_lo(uint32_t(km), 0u)));
            st[sl] = StageRec{x[0], x[1], x[2], jc, 0};
          }
          __syncthreads();  // single-wave group: the staged block is complete
          if (use_ring && single) {
            // at most 16 appends per lane between segment stores: a block
            // starts with the next entry in the segment of the first
            // unstored one, so fewer than 32 entries are ever unstored and
            // the slot of the next entry is free.  Every candidate is
            // written there and the position advances on a hit only.
            for (int t0 = 0; t0 < nk; t0 += kSeg) {
              const int t1 = nk - t0 < kSeg ? nk : t0 + kSeg;
              // byte offset of the next entry's ring slot in bits 8..12
              const uint32_t k8s = uint32_t(out + cnt) << 8;
              uint32_t k8 = k8s;
              int t = t0;
              for (; t + kStageUnroll <= t1; t += kStageUnroll) {
                double e[kStageUnroll][3];
                int jj[kStageUnroll];
#pragma unroll
                for (int u = 0; u < kStageUnroll; ++u) load_stage(st[t + u], e[u], jj[u]);
                bool h[kStageUnroll];
#pragma unroll
                for (int u = 0; u < kStageUnroll; ++u) {
                  double acc = 0.0;
#pragma unroll
                  for (int d = 0; d < D; ++d) {
                    const double dx = e[u][d] - qi[d];
                    acc = acc + dx * dx;
                  }
                  h[u] = acc <= r2m;
                }
#pragma unroll
                for (int u = 0; u < kStageUnroll; ++u) {
                  ring_put(ring, k8, lane, uint32_t(jj[u]));
                  k8 += h[u] ? 256u : 0u;
                }
              }
              for (; t < t1; ++t) {
                double e1[3];
                int j1;
                load_stage(st[t], e1, j1);
                double acc = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                  const double dx = e1[d] - qi[d];
                  acc = acc + dx * dx;
                }
                ring_put(ring, k8, lane, uint32_t(j1));
                k8 += acc <= r2m ? 256u : 0u;
              }
              cnt += int((k8 - k8s) >> 8);
              ring_segment();
            }
            __syncthreads();  // the next block overwrites the stage
            continue;
          }
          if (MODE == kCount) {
            // a hit adds m (1 whenever no lane visits the bucket twice)
            int t = 0;
            for (; t + kStageUnroll <= nk; t += kStageUnroll) {
              double e[kStageUnroll][3];
              int jj[kStageUnroll];
#pragma unroll
              for (int u = 0; u < kStageUnroll; ++u) load_stage(st[t + u], e[u], jj[u]);
#pragma unroll
              for (int u = 0; u < kStageUnroll; ++u) {
                double acc = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                  const double dx = e[u][d] - qi[d];
                  acc = acc + dx * dx;
                }
                cnt += acc <= r2m ? m : 0;
              }
            }
            for (; t < nk; ++t) {
              double e1[3];
              int j1;
              load_stage(st[t], e1, j1);
              double acc = 0.0;
#pragma unroll
              for (int d = 0; d < D; ++d) {
                const double dx = e1[d] - qi[d];
                acc = acc + dx * dx;
              }
              cnt += acc <= r2m ? m : 0;
            }
            __syncthreads();  // the next block overwrites the stage
            continue;
          }
          if (use_ring) ring_drain();  // several visits of one bucket: direct stores
          int t = 0;
          for (; t + kStageUnroll <= nk; t += kStageUnroll) {
            double e[kStageUnroll][3];
            int jj[kStageUnroll];
#pragma unroll
            for (int u = 0; u < kStageUnroll; ++u) load_stage(st[t + u], e[u], jj[u]);
            double acc[kStageUnroll];
#pragma unroll
            for (int u = 0; u < kStageUnroll; ++u) {
              acc[u] = 0.0;
#pragma unroll
              for (int d = 0; d < D; ++d) {
                const double dx = e[u][d] - qi[d];
                acc[u] = acc[u] + dx * dx;
              }
            }
#pragma unroll
            for (int u = 0; u < kStageUnroll; ++u)
              staged_hit<MODE>(a, single, acc[u] <= r2m, acc[u], jj[u], m, wrow, out, cnt);
          }
          for (; t < nk; ++t) {
            double e1[3];
            int j1;
            load_stage(st[t], e1, j1);
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
              const double dx = e1[d] - qi[d];
              acc = acc + dx * dx;
            }
            staged_hit<MODE>(a, single, acc <= r2m, acc, j1, m, wrow, out, cnt);
          }
          if (use_ring) fpos = out + cnt;
          __syncthreads();  // the next block overwrites the stage
        }
      } else {
        // matvec: the candidates are staged and culled as in the list
        // passes, with their right-hand sides; every lane then evaluates the
        // kernel for the staged candidates within its radius, in candidate
        // order (y += f x for each of the m visits)
        double blo[3] = {0.0, 0.0, 0.0}, bhi[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const int dg = (img >> (2 * d)) & 3;
          blo[d] = uniform_f64(
              pick3(dg, qlo[d] + double(-1) * g.L[d], qlo[d], qlo[d] + double(1) * g.L[d]));
          bhi[d] = uniform_f64(
              pick3(dg, qhi[d] + double(-1) * g.L[d], qhi[d], qhi[d] + double(1) * g.L[d]));
        }
        const double r2m = m != 0 ? a.r2 : -1.0;
        for (int c0 = beg; c0 < end; c0 += kGroup) {
          const int jc = c0 + lane;
          bool keep = jc < end;
          double x[3] = {0.0, 0.0, 0.0};
          if (keep) {
#pragma unroll
            for (int d = 0; d < D; ++d) x[d] = a.pos[int64_t(jc) * D + d];
          }
          double dd = 0.0;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            double e = 0.0;
            if (x[d] < blo[d]) e = blo[d] - x[d];
            else if (x[d] > bhi[d]) e = x[d] - bhi[d];
            dd = dd + e * e;
          }
          keep = keep && dd <= a.r2;
          const uint64_t km = __ballot(keep);
          if (km == 0) continue;
          const int nk = __popcll(km);
          if (keep) {
            const int sl = int(__builtin_amdgcn_mbcnt_hi(uint32_t(km >> 32),
                                                         __builtin_amdgcn_mbcnt_lo(uint32_t(km), 0u)));
            st[sl] = StageRec{x[0], x[1], x[2], jc, 0};
            if (MODE == kMatvec32) {
              // the FP32 form stages its right-hand sides in FP32 (rounded
              // once here instead of at every use)
              float *xf = reinterpret_cast<float *>(xs);
#pragma unroll
              for (int r = 0; r < kMaxRhs; ++r)
                xf[sl * kMaxRhs + r] = r < a.nrhs ? float(a.x[int64_t(r) * a.ldx + jc]) : 0.0f;
            } else {
#pragma unroll
              for (int r = 0; r < kMaxRhs; ++r)
                xs[sl * kMaxRhs + r] = r < a.nrhs ? a.x[int64_t(r) * a.ldx + jc] : 0.0;
            }
          }
          __syncthreads();  // single-wave group: the staged block is complete
          // radius tests of the whole block first, this lane's hits as bits
          // (candidate t at bit nk - 1 - t), then the kernel for the hits
          // only, in candidate order: a round evaluates one hit per lane,
          // so lanes do not wait on the kernels of candidates they miss
          uint64_t hm = 0;
          {
            int t = 0;
            for (; t + kStageUnroll <= nk; t += kStageUnroll) {
              double e[kStageUnroll][3];
              int jj[kStageUnroll];
#pragma unroll
              for (int u = 0; u < kStageUnroll; ++u) load_stage(st[t + u], e[u], jj[u]);
#pragma unroll
              for (int u = 0; u < kStageUnroll; ++u) {
                double acc = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                  const double dx = e[u][d] - qi[d];
                  acc = acc + dx * dx;
                }
                hm = (hm << 1) + (acc <= r2m ? 1ull : 0ull);
              }
            }
            for (; t < nk; ++t) {
              double e1[3];
              int j1;
              load_stage(st[t], e1, j1);
              double acc = 0.0;
#pragma unroll
              for (int d = 0; d < D; ++d) {
                const double dx = e1[d] - qi[d];
                acc = acc + dx * dx;
              }
              hm = (hm << 1) + (acc <= r2m ? 1ull : 0ull);
            }
          }
          // (a divergent loop: the wave runs as many rounds as its lane
          // with the most hits)
          while (hm != 0) {
            const int p = 63 - __builtin_clzll(hm);
            hm ^= 1ull << p;
            const int t = nk - 1 - p;
            double e1[3];
            int j1;
            load_stage(st[t], e1, j1);
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
              const double dx = e1[d] - qi[d];
              acc = acc + dx * dx;
            }
            double w = wrow;
            if (a.k.col_s) w = w + a.k.wc * a.k.col_s[j1];
            const double *xr = xs + t * kMaxRhs;
            if (MODE == kMatvec64) {
              const double f = phi64(a.k, acc) * w;
              for (int rep = 0; rep < m; ++rep) {
#pragma unroll
                for (int r = 0; r < kMaxRhs; ++r)
                  if (r < a.nrhs) acc64[r] += f * xr[r];
              }
            } else {
              const float f = phi32(a.k, float(acc)) * float(w);
              const float *xf = reinterpret_cast<const float *>(xs) + t * kMaxRhs;
              for (int rep = 0; rep < m; ++rep) {
#pragma unroll
                for (int r = 0; r < kMaxRhs; ++r)
                  if (r < a.nrhs) kb_add(acc32[r], cmp32[r], f * xf[r]);
              }
            }
          }
          __syncthreads();  // the next block overwrites the stage
        }
      }
    }
    if (use_ring) ring_drain();
    if (has_row) {
      if (MODE == kCount || MODE == kFillPadded) a.counts[row] = cnt;
      if (MODE == kMatvec64) {
#pragma unroll
        for (int r = 0; r < kMaxRhs; ++r)
          if (r < a.nrhs) a.y[int64_t(r) * a.ldy + row] = acc64[r];
      }
      if (MODE == kMatvec32) {
#pragma unroll
        for (int r = 0; r < kMaxRhs; ++r)
          if (r < a.nrhs)
            a.y[int64_t(r) * a.ldy + row] = acc64[r] + (double(acc32[r]) + double(cmp32[r]));
      }
    }
    __syncthreads();  // slots / mult reused by the next work item
  }
}

// xr[j][r] = x[r][j] (r < nrhs, zero beyond): the right-hand sides of one
// column in one 64-byte segment for the CSR product's gathers
template <typename T>
__global__ __launch_bounds__(256) void k_rhs_rows(const double *__restrict__ x, int64_t ldx,
                                                  int nrhs, int64_t n, T *__restrict__ xr) {
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (j >= n) return;
  T v[kMaxRhs];
#pragma unroll
  for (int r = 0; r < kMaxRhs; ++r) v[r] = r < nrhs ? T(x[int64_t(r) * ldx + j]) : T(0);
#pragma unroll
  for (int r = 0; r < kMaxRhs; ++r) xr[j * kMaxRhs + r] = v[r];
}

// y += K x from the cached CSR form of the operator: one wavefront per row.
// Four lanes share an entry and each takes two of its eight right-hand
// sides, so that one 16-byte load per lane reads an entry's whole 64-byte
// row of xr (one cache line for four lanes); sixteen entries per step, the
// partial sums then combined over the lanes holding the same pair.  FP64
// reads val; the FP32 form reads val32 (the values the FP32 matvec forms)
// and sums each lane's terms with compensation.  xr: the right-hand sides by
// column (k_rhs_rows), in FP64, or in FP32 for the FP32 form (which rounds
// every x to FP32 anyway: a 32-byte row per entry).
template <bool FP32>
__global__ __launch_bounds__(256) void k_csr_spmv(int64_t nrows, const int64_t *__restrict__ rp,
                                                  const int32_t *__restrict__ col,
                                                  const double *__restrict__ val,
                                                  const float *__restrict__ val32, int nrhs,
                                                  const void *__restrict__ xrv,
                                                  double *__restrict__ y, int64_t ldy) {
  const double *xr = static_cast<const double *>(xrv);
  const float *xr32 = static_cast<const float *>(xrv);
  const int blk = abh::xcd_block(int((nrows + 3) / 4));
  const int64_t row = int64_t(blk) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nrows) return;
  const int g = lane >> 2, q = lane & 3;  // entry slot, right-hand-side pair
  const int64_t e0 = rp[row], e1 = rp[row + 1];
  double s64[2] = {0.0, 0.0};
  float s32[2] = {0.0f, 0.0f}, c32[2] = {0.0f, 0.0f};
  // the row's entries are read 64 at a time (one contiguous load of col and
  // of val per block, two blocks ahead in flight), then handed to the lane
  // groups by shuffles
  int64_t b = e0;
  int32_t jn = 0, jn2 = 0;
  double fn = 0.0, fn2 = 0.0;
  if (b + lane < e1) {
    jn = col[b + lane];
    fn = FP32 ? double(val32[b + lane]) : val[b + lane];
  }
  if (b + 64 + lane < e1) {
    jn2 = col[b + 64 + lane];
    fn2 = FP32 ? double(val32[b + 64 + lane]) : val[b + 64 + lane];
  }
  for (; b < e1; b += 64) {
    const int32_t jl = jn;
    const double fl = fn;
    jn = jn2;
    fn = fn2;
    const int nb = e1 - b < 64 ? int(e1 - b) : 64;
    // the four gathers of x for the block's 64 entries are issued together,
    // before the next block's col / val loads (slots past the row's end read
    // a column of the row, or column 0, and add nothing), then accumulated
    // in entry order
    int32_t j[4];
    double f[4];
    double2 xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      j[u] = __shfl(jl, 16 * u + g, 64);
      f[u] = __shfl(fl, 16 * u + g, 64);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (FP32) {
        const float2 v = *reinterpret_cast<const float2 *>(xr32 + int64_t(j[u]) * kMaxRhs + 2 * q);
        xv[u] = make_double2(double(v.x), double(v.y));
      } else {
        xv[u] = *reinterpret_cast<const double2 *>(xr + int64_t(j[u]) * kMaxRhs + 2 * q);
      }
    }
    const int64_t bn = b + 128 + lane;
    if (bn < e1) {
      jn2 = col[bn];
      fn2 = FP32 ? double(val32[bn]) : val[bn];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool ok = 16 * u + g < nb;
      if (FP32) {
        const float f32 = float(f[u]);  // exact: f holds an FP32 value
        float sa = s32[0], ca = c32[0], sb = s32[1], cb = c32[1];
        kb_add(sa, ca, f32 * float(xv[u].x));
        kb_add(sb, cb, f32 * float(xv[u].y));
        s32[0] = ok ? sa : s32[0];
        c32[0] = ok ? ca : c32[0];
        s32[1] = ok ? sb : s32[1];
        c32[1] = ok ? cb : c32[1];
      } else {
        const double ta = s64[0] + f[u] * xv[u].x;
        const double tb = s64[1] + f[u] * xv[u].y;
        s64[0] = ok ? ta : s64[0];
        s64[1] = ok ? tb : s64[1];
      }
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double t = FP32 ? double(s32[h]) + double(c32[h]) : s64[h];
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) t += __shfl_xor(t, o, 64);
    const int r = 2 * q + h;
    if (g == 0 && r < nrhs) y[int64_t(r) * ldy + row] += t;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// plan (grouping of queries into work items), shared by search and operator
// ---------------------------------------------------------------------------

namespace abh {

struct QueryPlan {
  abh_context *ctx = nullptr;
  const abh_cells *cells = nullptr;
  int D = 3;
  int64_t nq = 0;
  bool self = false;      // centres are the cell list's own positions
  bool permuted = false;  // centres were sorted (perm / qbuf hold the order)
  int sub_levels = 0;     // Morton sub-cell levels inside a home cell
  double radius = 0, r2 = 0;
  int w[3] = {0, 0, 0};
  int nslots = 1;
  const double *q_sorted = nullptr;
  DevBuf qbuf, perm, keys, flags, ridx, run_start, chunks, choff, work, counters, escape, sched;

  const int32_t *qperm() const { return permuted ? perm.as<int32_t>() : nullptr; }
};

static void set_stencil(QueryPlan &p, int extra) {
  const abh_cells *c = p.cells;
  p.nslots = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < p.D) {
      // w = ceil(r / side), shaved by 1e-9 so r == side gives 1; rounding
      // at the boundary is caught by the escape check.
      int w = int(std::ceil(p.radius * c->inv[d] * (1.0 - 1e-9)));
      if (w < 0) w = 0;
      w += extra;
      // virtual cells needed lie within [-size, 2*size): a wider stencil
      // only adds slots that need images beyond +-1
      if (w > 2 * int(c->size[d]) + 1) w = 2 * int(c->size[d]) + 1;
      p.w[d] = w;
      p.nslots *= 2 * w + 1;
    } else {
      p.w[d] = 0;
    }
  }
  ABH_REQUIRE(p.nslots <= kMaxSlots, ABH_ERR_UNSUPPORTED,
              "search radius spans too many buckets (stencil > 343 cells); "
              "use fewer particles per bucket or a smaller radius");
}

template <int D> static void build_plan_d(QueryPlan &p, const double *q) {
  abh_context *ctx = p.ctx;
  hipStream_t s = ctx->stream;
  const GridParams g = p.cells->params();
  const int64_t n = p.nq;
  const unsigned blocks = unsigned(ceil_div(n, kThreads));
  p.keys.ensure(size_t(n + 1) * 4);
  uint32_t *keys = p.keys.as<uint32_t>();
  uint64_t maxkey = 1;
  for (int d = 0; d < D; ++d)
    maxkey *= uint64_t(g.periodic[d] ? 3 * g.size[d] : g.size[d] + 2);
  int bits = 0;
  while (bits < 64 && ((maxkey - 1) >> bits) != 0) ++bits;
  // sub-cell levels that fit the 32-bit key (at most 2: 4^D sub-cells)
  int levels = 0;
  while (levels < 2 && bits + (levels + 1) * D <= 32) ++levels;
  p.sub_levels = levels;
  p.permuted = !p.self || levels > 0;
  if (!p.permuted) {
    p.q_sorted = q;
    k_home_keys<D><<<blocks, kThreads, 0, s>>>(q, n, g, 0, keys, nullptr);
    check_launch();
  } else {
    ABH_REQUIRE(bits <= 32, ABH_ERR_UNSUPPORTED,
                "too many buckets to group arbitrary centres; search from the particle "
                "positions themselves");
    p.perm.ensure(size_t(n + 1) * 4);
    k_home_keys<D><<<blocks, kThreads, 0, s>>>(q, n, g, levels, keys, p.perm.as<int32_t>());
    check_launch();
    radix_sort_pairs(ctx, keys, p.perm.as<int32_t>(), n, bits + levels * D);
    p.qbuf.ensure(size_t(n + 1) * D * 8);
    gather_bytes(ctx, p.perm.as<int32_t>(), n, q, p.qbuf.p, size_t(D) * 8);
    p.q_sorted = p.qbuf.as<double>();
  }
  p.flags.ensure(size_t(n + 1) * 4);
  p.ridx.ensure(size_t(n + 1) * 4);
  p.run_start.ensure(size_t(n + 1) * 4);
  p.chunks.ensure(size_t(n + 1) * 4);
  p.choff.ensure(size_t(n + 1) * 4);
  p.work.ensure(size_t(n + 1) * sizeof(WorkItem));
  p.counters.ensure(16);
  p.sched.ensure(16);
  uint32_t *ctr = p.counters.as<uint32_t>();  // [0] = runs, [1] = work items
  k_run_flags<<<blocks, kThreads, 0, s>>>(keys, n, p.sub_levels * D, p.flags.as<uint32_t>());
  check_launch();
  scan_exclusive(ctx, p.flags.as<uint32_t>(), p.ridx.as<uint32_t>(), n, ctr + 0);
  k_run_starts<<<blocks, kThreads, 0, s>>>(p.flags.as<uint32_t>(), p.ridx.as<uint32_t>(), n,
                                           p.run_start.as<uint32_t>());
  check_launch();
  k_run_chunks<<<blocks, kThreads, 0, s>>>(p.run_start.as<uint32_t>(), ctr + 0, n,
                                           p.chunks.as<uint32_t>());
  check_launch();
  scan_exclusive(ctx, p.chunks.as<uint32_t>(), p.choff.as<uint32_t>(), n, ctr + 1);
  k_fill_work<D><<<blocks, kThreads, 0, s>>>(p.run_start.as<uint32_t>(), ctr + 0, n,
                                             p.choff.as<uint32_t>(), p.chunks.as<uint32_t>(),
                                             p.q_sorted, g, p.work.as<WorkItem>());
  check_launch();
}

static void build_plan(QueryPlan &p, abh_context *ctx, const abh_cells *cells, int64_t nq,
                       const double *q, double radius) {
  ABH_REQUIRE(cells && cells->ctx == ctx, ABH_ERR_INVALID_ARG, "cells belong to another context");
  ABH_REQUIRE(nq >= 0 && nq < (int64_t(1) << 31), ABH_ERR_INVALID_ARG, "nq out of range");
  ABH_REQUIRE(nq == 0 || q, ABH_ERR_INVALID_ARG, "query positions are null");
  ABH_REQUIRE(radius >= 0 && std::isfinite(radius), ABH_ERR_INVALID_ARG, "radius must be finite");
  p.ctx = ctx;
  p.cells = cells;
  p.D = cells->dim;
  p.nq = nq;
  p.radius = radius;
  p.r2 = radius * radius;  // distance_helper<2>::get_value_to_accumulate
  p.self = (q == cells->pos.as<double>() && nq == cells->n_alive);
  set_stencil(p, 0);
  if (nq == 0) return;
  if (p.D == 1) build_plan_d<1>(p, q);
  else if (p.D == 2) build_plan_d<2>(p, q);
  else build_plan_d<3>(p, q);
}

static SearchArgs base_args(const QueryPlan &p) {
  SearchArgs a{};
  a.work = p.work.as<WorkItem>();
  a.nwork = p.counters.as<uint32_t>() + 1;
  a.qpos = p.q_sorted;
  a.qperm = p.qperm();
  a.bbeg = p.cells->begin.as<uint32_t>();
  a.bend = p.cells->end.as<uint32_t>();
  a.pos = p.cells->pos.as<double>();
  a.r2 = p.r2;
  for (int d = 0; d < 3; ++d) a.w[d] = p.w[d];
  a.nslots = p.nslots;
  return a;
}

// dynamic LDS of k_search (the fill passes add their static output ring)
static size_t lds_bytes(int nslots, bool matvec) {
  const size_t slots = size_t(nslots) * (sizeof(Slot) + 4 + kGroup) + 4 * size_t(nslots + 1);
  return ((slots + 15) & ~size_t(15)) + size_t(kGroup) * sizeof(StageRec) +
         (matvec ? size_t(kGroup) * kMaxRhs * sizeof(double) : 0);
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

template <int MODE> static void launch_search(const QueryPlan &p, const SearchArgs &a0) {
  // persistent waves per CU: the list passes run best at 16 (measured on
  // the 4M-particle mixture: 9.4 / 14.3 ms at 16 against 10.0 / 15.6 ms at
  // 32; 8 leaves the scalar-load latency exposed)
  const int wpc = (MODE == kMatvec64 || MODE == kMatvec32) ? 32 : 16;
  const unsigned grid = unsigned(std::min<int64_t>(ceil_div(p.nq, kGroup), int64_t(p.ctx->n_cu) * wpc));
  if (grid == 0) return;
  SearchArgs a = a0;
  a.next = p.sched.as<uint32_t>();
  ABH_REQUIRE(a.next, ABH_ERR_INVALID_ARG, "search plan not built");
  k_zero_u32<<<1, 1, 0, p.ctx->stream>>>(a.next);
  check_launch();
  const GridParams g = p.cells->params();
  const size_t lds = lds_bytes(a.nslots, MODE == kMatvec64 || MODE == kMatvec32);
  if (p.D == 1) k_search<1, MODE><<<grid, kGroup, lds, p.ctx->stream>>>(a, g);
  else if (p.D == 2) k_search<2, MODE><<<grid, kGroup, lds, p.ctx->stream>>>(a, g);
  else k_search<3, MODE><<<grid, kGroup, lds, p.ctx->stream>>>(a, g);
  check_launch();
}

// Counts neighbours into counts[nq] (device), widening the stencil until no
// query's visited buckets escape it.
// col[row_ptr[i] + e] = padded[pad_ptr[i] + e], one wavefront per row
// (the row's body is stored 16 bytes per lane from its first 16-byte
// aligned address, head and tail element by element)
__global__ __launch_bounds__(256) void k_compact_rows(const int64_t *__restrict__ pad_ptr,
                                                      const int64_t *__restrict__ row_ptr,
                                                      int64_t nrows,
                                                      const int32_t *__restrict__ padded,
                                                      int32_t *__restrict__ col) {
  const int64_t i = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (i >= nrows) return;
  const int lane = threadIdx.x & 63;
  const int64_t src = pad_ptr[i], dst = row_ptr[i], n = row_ptr[i + 1] - dst;
  const uintptr_t a0 = reinterpret_cast<uintptr_t>(col + dst);
  const int64_t head = std::min<int64_t>(n, int64_t(((16 - (a0 & 15)) & 15) >> 2));
  if (lane < head) col[dst + lane] = padded[src + lane];
  const int64_t d0 = dst + head, s0 = src + head;
  const int64_t nb4 = (n - head) >> 2;
  for (int64_t c = lane; c < nb4; c += 64) {
    const int32_t *sp = padded + s0 + 4 * c;
    *reinterpret_cast<int4 *>(col + d0 + 4 * c) = make_int4(sp[0], sp[1], sp[2], sp[3]);
  }
  const int64_t rem = (n - head) - 4 * nb4;
  if (lane < rem) col[d0 + 4 * nb4 + lane] = padded[s0 + 4 * nb4 + lane];
}

// Like count_with_escape, with the capacity pass (no distance tests).
static void capacity_with_escape(QueryPlan &p, int64_t *caps) {
  p.escape.ensure(4);
  for (int extra = 0;; ++extra) {
    if (extra > 0) set_stencil(p, extra);
    ABH_HIP(hipMemsetAsync(p.escape.p, 0, 4, p.ctx->stream));
    SearchArgs a = base_args(p);
    a.counts = caps;
    a.escape_flag = p.escape.as<uint32_t>();
    launch_search<kCapacity>(p, a);
    uint32_t esc = 0;
    ABH_HIP(hipMemcpyAsync(&esc, p.escape.p, 4, hipMemcpyDeviceToHost, p.ctx->stream));
    ABH_HIP(hipStreamSynchronize(p.ctx->stream));
    if (!esc) return;
    ABH_REQUIRE(extra < 4, ABH_ERR_UNSUPPORTED, "search stencil failed to cover visited buckets");
  }
}

static void count_with_escape(QueryPlan &p, int64_t *counts) {
  p.escape.ensure(4);
  for (int extra = 0;; ++extra) {
    if (extra > 0) set_stencil(p, extra);
    ABH_HIP(hipMemsetAsync(p.escape.p, 0, 4, p.ctx->stream));
    SearchArgs a = base_args(p);
    a.counts = counts;
    a.escape_flag = p.escape.as<uint32_t>();
    launch_search<kCount>(p, a);
    uint32_t esc = 0;
    ABH_HIP(hipMemcpyAsync(&esc, p.escape.p, 4, hipMemcpyDeviceToHost, p.ctx->stream));
    ABH_HIP(hipStreamSynchronize(p.ctx->stream));
    if (!esc) return;
    ABH_REQUIRE(extra < 4, ABH_ERR_UNSUPPORTED, "search stencil failed to cover visited buckets");
  }
}

}  // namespace abh

struct abh_search {
  abh::QueryPlan plan;
  abh::DevBuf row_ptr, total_buf;
  int64_t total = 0;
  abh::DevBuf pad_ptr, padded;  // single-traversal search: rows at capacity offsets
};

struct abh_sparse_op {
  abh::QueryPlan plan;
  KernelDesc k{};
  int64_t nnz = 0;
  int64_t ncols = 0;
  // cached CSR form (abh_sparse_op_cache): row pointers, columns and the
  // kernel values of the FP64 and / or FP32 forms
  abh::DevBuf c_rp, c_col, c_val, c_val32, c_xr;
  bool cached64 = false, cached32 = false;
};

namespace abh {

// y += K x for nrhs <= kMaxRhs right-hand sides, from the cached CSR form
// when it holds this precision, else evaluating the kernel on the fly
static void apply_group(abh_sparse_op *op, int nrhs, const double *x, int64_t ldx, double *y,
                        int64_t ldy, bool fp32) {
  const int64_t nq = op->plan.nq;
  if (nq == 0 || nrhs == 0) return;
  if (fp32 ? op->cached32 : op->cached64) {
    hipStream_t st = op->plan.ctx->stream;
    const int64_t nc = op->ncols;
    op->c_xr.ensure(size_t(nc) * kMaxRhs * 8 + 64);
    void *xr = op->c_xr.p;
    if (nc > 0) {
      if (fp32)
        k_rhs_rows<float><<<unsigned(ceil_div(nc, 256)), 256, 0, st>>>(x, ldx, nrhs, nc,
                                                                         op->c_xr.as<float>());
      else
        k_rhs_rows<double><<<unsigned(ceil_div(nc, 256)), 256, 0, st>>>(x, ldx, nrhs, nc,
                                                                          op->c_xr.as<double>());
      check_launch();
    }
    const unsigned blocks = xcd_grid(int(ceil_div(nq, 4)));
    if (fp32)
      k_csr_spmv<true><<<blocks, 256, 0, st>>>(nq, op->c_rp.as<int64_t>(), op->c_col.as<int32_t>(),
                                               nullptr, op->c_val32.as<float>(), nrhs, xr, y,
                                               ldy);
    else
      k_csr_spmv<false><<<blocks, 256, 0, st>>>(nq, op->c_rp.as<int64_t>(),
                                                op->c_col.as<int32_t>(), op->c_val.as<double>(),
                                                nullptr, nrhs, xr, y, ldy);
    check_launch();
    return;
  }
  SearchArgs a = base_args(op->plan);
  a.k = op->k;
  a.nrhs = nrhs;
  a.x = x;
  a.ldx = ldx;
  a.y = y;
  a.ldy = ldy;
  if (fp32) launch_search<kMatvec32>(op->plan, a);
  else launch_search<kMatvec64>(op->plan, a);
}

void sparse_apply(abh_sparse_op *op, const double *x, double *y, bool fp32) {
  apply_group(op, 1, x, op->ncols, y, op->plan.nq, fp32);
}

int64_t sparse_rows(const abh_sparse_op *op) { return op->plan.nq; }
int64_t sparse_cols(const abh_sparse_op *op) { return op->ncols; }
abh_context *sparse_context(const abh_sparse_op *op) { return op->plan.ctx; }

}  // namespace abh

extern "C" {

abh_status abh_search_update(abh_search *s, const abh_cells *cells, int64_t nq,
                             const double *q, double radius) {
  if (!s) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = s->plan.ctx;
  return abh::guard(ctx, [&] {
    abh::build_plan(s->plan, ctx, cells, nq, q, radius);
    s->row_ptr.ensure(size_t(nq + 1) * 8);
    s->total_buf.ensure(8);
    int64_t *rp = s->row_ptr.as<int64_t>();
    ABH_HIP(hipMemsetAsync(rp + nq, 0, 8, ctx->stream));
    if (nq > 0) abh::count_with_escape(s->plan, rp);
    int64_t *total_dev = s->total_buf.as<int64_t>();
    abh::scan_exclusive(ctx, rp, rp, nq + 1, total_dev);
    int64_t total = 0;
    ABH_HIP(hipMemcpyAsync(&total, total_dev, 8, hipMemcpyDeviceToHost, ctx->stream));
    ABH_HIP(hipStreamSynchronize(ctx->stream));
    s->total = total;
  });
}

abh_status abh_search_create(abh_context *ctx, const abh_cells *cells, int64_t nq,
                             const double *q, double radius, abh_search **out) {
  if (!ctx || !out) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_search *s = new (std::nothrow) abh_search();
  if (!s) return ABH_ERR_OUT_OF_MEMORY;
  if (ctx->closed) {
    delete s;
    return ABH_ERR_INVALID_ARG;
  }
  s->plan.ctx = ctx;
  abh_status st = abh_search_update(s, cells, nq, q, radius);
  if (st != ABH_OK) {
    delete s;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = s;
  return ABH_OK;
}

abh_status abh_search_destroy(abh_search *s) {
  if (!s) return ABH_OK;
  abh_context *ctx = s->plan.ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete s;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_search_total(const abh_search *s, int64_t *total) {
  if (!s || !total) return ABH_ERR_INVALID_ARG;
  *total = s->total;
  return ABH_OK;
}

const int64_t *abh_search_row_ptr(const abh_search *s) {
  return s ? s->row_ptr.as<int64_t>() : nullptr;
}

abh_status abh_search_fill(abh_search *s, int32_t *col, int64_t capacity) {
  if (!s) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = s->plan.ctx;
  return abh::guard(ctx, [&] {
    ABH_REQUIRE(capacity >= s->total, ABH_ERR_CAPACITY, "neighbour index buffer too small");
    ABH_REQUIRE(s->total == 0 || col, ABH_ERR_INVALID_ARG, "col is null");
    if (s->plan.nq == 0 || s->total == 0) return;
    SearchArgs a = abh::base_args(s->plan);
    a.row_ptr = s->row_ptr.as<int64_t>();
    a.col = col;
    a.ring = (reinterpret_cast<uintptr_t>(col) & 63) == 0 ? 1 : 0;
    abh::launch_search<kFill>(s->plan, a);
  });
}

abh_status abh_search_run_fused(abh_search *s, const abh_cells *cells, int64_t nq,
                                const double *q, double radius, int32_t *col, int64_t capacity,
                                int64_t *total_pairs) {
  if (!s) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = s->plan.ctx;
  return abh::guard(ctx, [&] {
    abh::build_plan(s->plan, ctx, cells, nq, q, radius);
    hipStream_t st = ctx->stream;
    s->row_ptr.ensure(size_t(nq + 1) * 8);
    s->pad_ptr.ensure(size_t(nq + 1) * 8);
    s->total_buf.ensure(16);
    int64_t *pp = s->pad_ptr.as<int64_t>();
    int64_t *rp = s->row_ptr.as<int64_t>();
    ABH_HIP(hipMemsetAsync(pp + nq, 0, 8, st));
    ABH_HIP(hipMemsetAsync(rp + nq, 0, 8, st));
    int64_t padded_total = 0;
    if (nq > 0) {
      abh::capacity_with_escape(s->plan, pp);
      abh::scan_exclusive(ctx, pp, pp, nq + 1, s->total_buf.as<int64_t>());
      ABH_HIP(hipMemcpyAsync(&padded_total, s->total_buf.p, 8, hipMemcpyDeviceToHost, st));
      ABH_HIP(hipStreamSynchronize(st));
    }
    s->padded.ensure(size_t(padded_total) * 4 + 64);
    if (nq > 0 && padded_total > 0) {
      SearchArgs a = abh::base_args(s->plan);
      a.row_ptr = pp;
      a.col = s->padded.as<int32_t>();
      a.ring = (reinterpret_cast<uintptr_t>(a.col) & 63) == 0 ? 1 : 0;
      a.counts = rp;
      abh::launch_search<kFillPadded>(s->plan, a);
    } else if (nq > 0) {
      ABH_HIP(hipMemsetAsync(rp, 0, size_t(nq) * 8, st));
    }
    int64_t total = 0;
    if (nq > 0) {
      abh::scan_exclusive(ctx, rp, rp, nq + 1, s->total_buf.as<int64_t>());
      ABH_HIP(hipMemcpyAsync(&total, s->total_buf.p, 8, hipMemcpyDeviceToHost, st));
      ABH_HIP(hipStreamSynchronize(st));
    }
    s->total = total;
    if (total_pairs) *total_pairs = total;
    ABH_REQUIRE(capacity >= total, ABH_ERR_CAPACITY, "neighbour index buffer too small");
    ABH_REQUIRE(total == 0 || col, ABH_ERR_INVALID_ARG, "col is null");
    if (total > 0) {
      abh::k_compact_rows<<<unsigned(abh::ceil_div(nq, 4)), 256, 0, st>>>(
          pp, rp, nq, s->padded.as<int32_t>(), col);
      abh::check_launch();
    }
  });
}

abh_status abh_sparse_op_create(abh_context *ctx, const abh_cells *cols, int64_t nrows,
                                const double *row_pos, double radius, const abh_kernel *kernel,
                                const double *row_scalar, const double *col_scalar,
                                abh_sparse_op **out) {
  if (!ctx || !out || !kernel) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_sparse_op *op = new (std::nothrow) abh_sparse_op();
  if (!op) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    ABH_REQUIRE(kernel->type >= ABH_KERNEL_CONSTANT && kernel->type <= ABH_KERNEL_SMOOTH_INV_R,
                ABH_ERR_UNSUPPORTED, "unknown kernel type");
    abh::build_plan(op->plan, ctx, cols, nrows, row_pos, radius);
    op->ncols = cols->n_alive;
    op->k.type = kernel->type;
    op->k.param = kernel->param;
    op->k.param2 = kernel->param * kernel->param;
    abh::set_inverses(op->k);
    op->k.w0 = kernel->w0;
    op->k.wr = kernel->w_row;
    op->k.wc = kernel->w_col;
    op->k.row_s = row_scalar;
    op->k.col_s = col_scalar;
    if (nrows > 0) {
      // the count pass fixes the stencil (escape check) and gives nnz
      abh::DevBuf counts;
      counts.ensure(size_t(nrows + 1) * 8);
      int64_t *c = counts.as<int64_t>();
      ABH_HIP(hipMemsetAsync(c + nrows, 0, 8, ctx->stream));
      abh::count_with_escape(op->plan, c);
      abh::DevBuf total_buf;
      total_buf.ensure(8);
      int64_t *total_dev = total_buf.as<int64_t>();
      abh::scan_exclusive(ctx, c, c, nrows + 1, total_dev);
      ABH_HIP(hipMemcpyAsync(&op->nnz, total_dev, 8, hipMemcpyDeviceToHost, ctx->stream));
      ABH_HIP(hipStreamSynchronize(ctx->stream));
    }
  });
  if (st != ABH_OK) {
    delete op;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = op;
  return ABH_OK;
}

abh_status abh_sparse_op_destroy(abh_sparse_op *op) {
  if (!op) return ABH_OK;
  abh_context *ctx = op->plan.ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete op;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_sparse_op_assemble(abh_sparse_op *op, int64_t *row_ptr, int32_t *col,
                                  double *val, int64_t capacity) {
  if (!op) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = op->plan.ctx;
  return abh::guard(ctx, [&] {
    const int64_t nq = op->plan.nq;
    ABH_REQUIRE(row_ptr, ABH_ERR_INVALID_ARG, "row_ptr is null");
    ABH_REQUIRE(capacity >= op->nnz, ABH_ERR_CAPACITY, "col / val capacity below nnz");
    ABH_REQUIRE(op->nnz == 0 || (col && val), ABH_ERR_INVALID_ARG, "col or val is null");
    ABH_HIP(hipMemsetAsync(row_ptr + nq, 0, 8, ctx->stream));
    if (nq == 0) return;
    abh::count_with_escape(op->plan, row_ptr);
    abh::DevBuf total_buf;
    total_buf.ensure(8);
    abh::scan_exclusive(ctx, row_ptr, row_ptr, nq + 1, total_buf.as<int64_t>());
    SearchArgs a = abh::base_args(op->plan);
    a.k = op->k;
    a.row_ptr = row_ptr;
    a.col = col;
    a.val = val;
    abh::launch_search<kAssemble>(op->plan, a);
  });
}

abh_status abh_sparse_op_nnz(const abh_sparse_op *op, int64_t *nnz) {
  if (!op || !nnz) return ABH_ERR_INVALID_ARG;
  *nnz = op->nnz;
  return ABH_OK;
}

abh_status abh_sparse_op_apply(abh_sparse_op *op, int nrhs, const double *x, int64_t ldx,
                               double *y, int64_t ldy, abh_precision precision) {
  if (!op) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = op->plan.ctx;
  return abh::guard(ctx, [&] {
    ABH_REQUIRE(nrhs >= 0, ABH_ERR_INVALID_ARG, "nrhs must be >= 0");
    ABH_REQUIRE(precision == ABH_FP64 || precision == ABH_FP32_COMPENSATED, ABH_ERR_UNSUPPORTED,
                "unknown precision");
    if (nrhs == 0 || op->plan.nq == 0) return;
    ABH_REQUIRE(x && y, ABH_ERR_INVALID_ARG, "x or y is null");
    ABH_REQUIRE(ldx >= op->ncols && ldy >= op->plan.nq, ABH_ERR_INVALID_ARG,
                "leading dimension smaller than the operator");
    for (int r0 = 0; r0 < nrhs; r0 += kMaxRhs)
      abh::apply_group(op, std::min(kMaxRhs, nrhs - r0), x + int64_t(r0) * ldx, ldx,
                       y + int64_t(r0) * ldy, ldy, precision == ABH_FP32_COMPENSATED);
  });
}

abh_status abh_sparse_op_cache(abh_sparse_op *op, abh_precision precision, int enable) {
  if (!op) return ABH_ERR_INVALID_ARG;
  abh_context *ctx = op->plan.ctx;
  return abh::guard(ctx, [&] {
    ABH_REQUIRE(precision == ABH_FP64 || precision == ABH_FP32_COMPENSATED, ABH_ERR_UNSUPPORTED,
                "unknown precision");
    const bool fp32 = precision == ABH_FP32_COMPENSATED;
    if (!enable) {
      (fp32 ? op->c_val32 : op->c_val).release();
      (fp32 ? op->cached32 : op->cached64) = false;
      if (!op->cached32 && !op->cached64) {
        op->c_col.release();
        op->c_rp.release();
      }
      return;
    }
    if (fp32 ? op->cached32 : op->cached64) return;
    const int64_t nq = op->plan.nq;
    const bool have = op->cached32 || op->cached64;
    const size_t need = size_t(op->nnz) * (fp32 ? 4 : 8) +
                        (have ? 0 : size_t(op->nnz) * 4 + size_t(nq + 1) * 8);
    size_t free_b = 0, total_b = 0;
    ABH_HIP(hipMemGetInfo(&free_b, &total_b));
    ABH_REQUIRE(double(need) <= 0.8 * double(free_b), ABH_ERR_OUT_OF_MEMORY,
                "not enough device memory to cache the operator");
    op->c_rp.ensure(size_t(nq + 1) * 8);
    op->c_col.ensure(size_t(op->nnz) * 4 + 64);
    (fp32 ? op->c_val32 : op->c_val).ensure(size_t(op->nnz) * (fp32 ? 4 : 8) + 64);
    if (nq == 0) {
      (fp32 ? op->cached32 : op->cached64) = true;
      return;
    }
    int64_t *rp = op->c_rp.as<int64_t>();
    ABH_HIP(hipMemsetAsync(rp + nq, 0, 8, ctx->stream));
    abh::count_with_escape(op->plan, rp);
    abh::DevBuf total_buf;
    total_buf.ensure(8);
    abh::scan_exclusive(ctx, rp, rp, nq + 1, total_buf.as<int64_t>());
    SearchArgs a = abh::base_args(op->plan);
    a.k = op->k;
    a.row_ptr = rp;
    a.col = op->c_col.as<int32_t>();
    a.val = fp32 ? nullptr : op->c_val.as<double>();
    a.val32 = fp32 ? op->c_val32.as<float>() : nullptr;
    // values of the other form already cached are rewritten identically
    if (!fp32 && op->cached32) a.val32 = op->c_val32.as<float>();
    if (fp32 && op->cached64) a.val = op->c_val.as<double>();
    abh::launch_search<kAssemble>(op->plan, a);
    ABH_HIP(hipStreamSynchronize(ctx->stream));
    (fp32 ? op->cached32 : op->cached64) = true;
  });
}

}  // extern "C"
