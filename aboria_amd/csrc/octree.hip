// The HyperOctree on the device (src/OctTree.h:102-911).
//
// Build (octree_update), one stream:
//   1. enforce the domain (shared with the cell list), exclusive scan of the
//      kept flags;
//   2. k_oct_tag: point_to_tag (src/detail/SpatialUtil.h:255-286) of every
//      kept particle, written at its compacted slot with its input index;
//   3. stable LSD radix sort of (tag, index) on the dim * max_level tag bits,
//      gather of the positions into tree order;
//   4. level by level (build_tree, src/OctTree.h:328-475): k_oct_children
//      tags the 2^dim children of every active node, finds their particle
//      range by binary search inside the parent's range (the same lower /
//      upper bound as over the whole tag array: the parent's range holds
//      every tag of the child's interval) and classifies them EMPTY / LEAF
//      (count <= int(n_leaf), or the last level) / NODE; one exclusive scan
//      of the packed (leaf, node) flags numbers the leaves and the next
//      level's nodes; k_oct_assign writes the child codes, the leaves and
//      the next active list.  The host reads one count pair per level.
//
// Traversal (octree_lists): the dual traversal of the FMM / H2 matrix
// evaluated level by level instead of by recursion.  The strong (too close
// for M2L) connections of every node are derived from its parent's, in the
// parent's order, children in child order; a count pass, a scan and a fill
// pass per level.  Then per non-empty leaf the strong connections expanded
// depth first to their non-empty leaves give the P2P list; the lists equal
// the recursive traversal's entry for entry.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cells.hpp"
#include "octree.hpp"

using abh::DevBuf;

namespace {

constexpr int kThreads = 256;
constexpr int32_t kEmpty = int32_t(0xffffffffu);

struct OctBounds {
  double lo[3], hi[3];
};

// active node of the level being built: its tag, node index (-1: the root)
// and particle range
struct Act {
  uint32_t tag;
  int32_t node;
  int32_t lb, ub;
};

template <int D> constexpr int max_level_of() { return 32 / D - 2; }

// point_to_tag (src/detail/SpatialUtil.h:255-286): per level one bit per
// dimension, first dimension most significant.
template <int D>
__global__ __launch_bounds__(kThreads) void k_oct_tag(const double *__restrict__ pos,
                                                      const uint32_t *__restrict__ flags,
                                                      const uint32_t *__restrict__ offs,
                                                      int64_t n, OctBounds b,
                                                      uint32_t *__restrict__ keys,
                                                      int32_t *__restrict__ order) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n || !flags[i]) return;
  constexpr int ML = max_level_of<D>();
  double p[D], lo[D], hi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    p[d] = pos[i * D + d];
    lo[d] = b.lo[d];
    hi[d] = b.hi[d];
  }
  uint32_t result = 0;
  for (int level = 1; level <= ML; ++level) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double mid = __dmul_rn(0.5, __dadd_rn(lo[d], hi[d]));
      const bool high = !(p[d] < mid);
      result = (result | uint32_t(high)) << 1;
      if (high) lo[d] = mid;
      else hi[d] = mid;
    }
  }
  result >>= 1;
  const uint32_t k = offs[i];
  keys[k] = result;
  order[k] = int32_t(i);
}

// children of the active nodes: tag, parent, particle range, and the
// packed flag (1 for a leaf, 1 << 32 for a node) the scan numbers
template <int D>
__global__ __launch_bounds__(kThreads) void k_oct_children(
    const Act *__restrict__ act, int64_t nch, const uint32_t *__restrict__ keys, int level,
    int threshold, int64_t cb, uint32_t *__restrict__ ntag, int32_t *__restrict__ parent,
    int2 *__restrict__ range, int64_t *__restrict__ cflags) {
  const int64_t c = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (c >= nch) return;
  constexpr int ML = max_level_of<D>();
  const Act a = act[c >> D];
  const int shift = (ML - level) * D;
  const uint32_t tag = a.tag | (uint32_t(c & ((1 << D) - 1)) << shift);
  const uint32_t last = tag + ((1u << shift) - 1u);
  int lo = a.lb, hi = a.ub;
  while (lo < hi) {  // lower_bound(tag)
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < tag) lo = mid + 1;
    else hi = mid;
  }
  const int lb = lo;
  hi = a.ub;
  while (lo < hi) {  // upper_bound(tag + length)
    const int mid = (lo + hi) >> 1;
    if (keys[mid] <= last) lo = mid + 1;
    else hi = mid;
  }
  const int ub = lo;
  ntag[cb + c] = tag;
  parent[cb + c] = a.node;
  range[c] = make_int2(lb, ub);
  const int count = ub - lb;
  int64_t f = 0;
  if (count > 0) f = (level == ML || count <= threshold) ? int64_t(1) : (int64_t(1) << 32);
  cflags[c] = f;
}

// write_nodes / make_leaf / copy_if of build_tree (src/OctTree.h:398-471)
template <int D>
__global__ __launch_bounds__(kThreads) void k_oct_assign(
    int64_t nch, int64_t cb, int64_t leaf_base, int64_t int_base, const uint32_t *__restrict__ ntag,
    const int2 *__restrict__ range, const int64_t *__restrict__ cflags,
    const int64_t *__restrict__ scanned, int32_t *__restrict__ nodes, int2 *__restrict__ leaves,
    int32_t *__restrict__ leaf_node, int32_t *__restrict__ internal, Act *__restrict__ next) {
  const int64_t c = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (c >= nch) return;
  const int64_t f = cflags[c];
  const int64_t sc = scanned[c];
  const int64_t k = cb + c;
  if (f == 0) {
    nodes[k] = kEmpty;
  } else if (f == 1) {
    const int64_t li = leaf_base + (sc & 0xffffffff);
    nodes[k] = int32_t(0x80000000u | uint32_t(li));
    leaves[li] = range[c];
    leaf_node[li] = int32_t(k);
  } else {
    const int64_t ni = sc >> 32;
    nodes[k] = int32_t(cb + nch + (int64_t(1) << D) * ni);
    internal[int_base + ni] = int32_t(k);
    const int2 r = range[c];
    next[ni] = Act{ntag[k], int32_t(k), r.x, r.y};
  }
}

// box and integer coordinates of the nodes of one level from their tags
template <int D>
__global__ __launch_bounds__(kThreads) void k_oct_geometry(int64_t b, int64_t cnt, int level,
                                                           const uint32_t *__restrict__ ntag,
                                                           OctBounds root,
                                                           double *__restrict__ box,
                                                           int32_t *__restrict__ coord) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= cnt) return;
  constexpr int ML = max_level_of<D>();
  const int64_t k = b + i;
  const uint32_t tag = ntag[k];
  double lo[D], hi[D];
  int32_t cr[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = root.lo[d];
    hi[d] = root.hi[d];
    cr[d] = 0;
  }
  for (int lv = 1; lv <= level; ++lv) {
    const uint32_t which = (tag >> ((ML - lv) * D)) & ((1u << D) - 1u);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const bool high = (which >> (D - 1 - d)) & 1u;
      const double mid = __dmul_rn(0.5, __dadd_rn(hi[d], lo[d]));
      if (high) lo[d] = mid;
      else hi[d] = mid;
      cr[d] = 2 * cr[d] + int32_t(high);
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    box[k * 2 * D + d] = lo[d];
    box[k * 2 * D + D + d] = hi[d];
    coord[k * D + d] = cr[d];
  }
}

// theta_condition::check (src/detail/FastMultipoleMethod.h:1249-1279): the
// source box is too close for M2L when the largest side of either box
// exceeds their distance.  t, s: lo[D] then hi[D].
template <int D>
__device__ __forceinline__ bool theta_close(const double *t, const double *s) {
  double md = 2.2250738585072014e-308;  // numeric_limits<double>::min()
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double w = __dsub_rn(t[D + i], t[i]);
    if (w > md) md = w;
  }
  double sq = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double dist = 0.0;
    if (t[D + i] < s[i]) dist = __dsub_rn(s[i], t[D + i]);
    else if (s[D + i] < t[i]) dist = __dsub_rn(t[i], s[D + i]);
    const double w = __dsub_rn(s[D + i], s[i]);
    if (w > md) md = w;
    sq = __dadd_rn(sq, __dmul_rn(dist, dist));
  }
  return __dmul_rn(md, md) > sq;
}

// Strong and weak connections of the nodes of one level.  Count pass:
// strong and weak counts per node (into ptr, scanned afterwards).  Fill
// pass: entries at base + the scanned offset; ptr made absolute.
template <int D, bool FILL>
__global__ __launch_bounds__(kThreads) void k_oct_conn(
    int level, int64_t b, int64_t cnt, const int32_t *__restrict__ nodes,
    const int32_t *__restrict__ parent, const double *__restrict__ box,
    int32_t *__restrict__ s_cnt, uint32_t *__restrict__ s_ptr, int32_t *__restrict__ s_idx,
    uint32_t s_base, uint32_t *__restrict__ w_ptr, int32_t *__restrict__ w_idx, uint32_t w_base,
    unsigned long long *__restrict__ m2l) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= cnt) return;
  const int64_t k = b + i;
  constexpr int NCH = 1 << D;
  double tb[2 * D];
#pragma unroll
  for (int j = 0; j < 2 * D; ++j) tb[j] = box[k * 2 * D + j];
  uint32_t ns = 0, nw = 0;
  int32_t *so = nullptr, *wo = nullptr;
  if (FILL) {
    const uint32_t rs = s_ptr[k];
    so = s_idx + s_base + rs;
    s_ptr[k] = s_base + rs;
    if (w_ptr) {
      const uint32_t rw = w_ptr[k];
      wo = w_idx + w_base + rw;
      w_ptr[k] = w_base + rw;
    }
  }
  auto test = [&](int32_t src) {
    if (theta_close<D>(tb, box + int64_t(src) * 2 * D)) {
      if (FILL) so[ns] = src;
      ++ns;
    } else {
      if (FILL && wo) wo[nw] = src;
      ++nw;
    }
  };
  if (level == 1) {
    for (int c = 0; c < NCH; ++c) test(c);
  } else {
    const int32_t p = parent[k];
    const uint32_t st = s_ptr[p];
    const int32_t ct = s_cnt[p];
    for (int32_t j = 0; j < ct; ++j) {
      const int32_t src = s_idx[st + j];
      const int32_t code = nodes[src];
      if (code < 0) {  // a leaf or an empty box stays connected as it is
        if (FILL) so[ns] = src;
        ++ns;
      } else {
        for (int c = 0; c < NCH; ++c) test(code + c);
      }
    }
  }
  if (!FILL) {
    s_cnt[k] = int32_t(ns);
    s_ptr[k] = ns;
    if (w_ptr) w_ptr[k] = nw;
    if (nw) atomicAdd(m2l, (unsigned long long)nw);
  }
}

// non-empty leaves under every node of one level (children first)
template <int D>
__global__ __launch_bounds__(kThreads) void k_oct_sub(int64_t b, int64_t cnt,
                                                      const int32_t *__restrict__ nodes,
                                                      int32_t *__restrict__ sub) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= cnt) return;
  const int64_t k = b + i;
  const int32_t code = nodes[k];
  int32_t v;
  if (code == kEmpty) v = 0;
  else if (code < 0) v = 1;
  else {
    v = 0;
    for (int c = 0; c < (1 << D); ++c) v += sub[code + c];
  }
  sub[k] = v;
}

// P2P lists: the strong connections of every non-empty leaf expanded depth
// first (children in order) to their non-empty leaves
template <int D, bool FILL>
__global__ __launch_bounds__(kThreads) void k_oct_p2p(
    int64_t nleaf, const int32_t *__restrict__ leaf_node, const int2 *__restrict__ leaves,
    const int32_t *__restrict__ nodes, const int32_t *__restrict__ sub,
    const int32_t *__restrict__ s_cnt, const uint32_t *__restrict__ s_ptr,
    const int32_t *__restrict__ s_idx, uint32_t *__restrict__ p_ptr, int32_t *__restrict__ p_src,
    unsigned long long *__restrict__ inter) {
  const int64_t li = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (li >= nleaf) return;
  const int32_t t = leaf_node[li];
  const uint32_t st = s_ptr[t];
  const int32_t ct = s_cnt[t];
  if (!FILL) {
    uint32_t c = 0;
    for (int32_t j = 0; j < ct; ++j) c += uint32_t(sub[s_idx[st + j]]);
    p_ptr[li] = c;
    return;
  }
  const int2 tr = leaves[li];
  int32_t *out = p_src + p_ptr[li];
  uint32_t w = 0;
  int64_t acc = 0;
  auto emit = [&](int32_t code) {
    if (code == kEmpty) return;
    const int32_t sl = int32_t(0x80000000u ^ uint32_t(code));
    out[w++] = sl;
    const int2 sr = leaves[sl];
    acc += int64_t(tr.y - tr.x) * (sr.y - sr.x);
  };
  constexpr int ML = max_level_of<D>();
  int32_t base[ML + 1], nxt[ML + 1];
  for (int32_t j = 0; j < ct; ++j) {
    const int32_t code0 = nodes[s_idx[st + j]];
    if (code0 < 0) {
      emit(code0);
      continue;
    }
    int top = 0;
    base[0] = code0;
    nxt[0] = 0;
    while (top >= 0) {
      if (nxt[top] == (1 << D)) {
        --top;
        continue;
      }
      const int32_t code = nodes[base[top] + nxt[top]++];
      if (code < 0) {
        emit(code);
      } else {
        ++top;  // the source is on level 1 or deeper: top stays below ML
        base[top] = code;
        nxt[top] = 0;
      }
    }
  }
  atomicAdd(inter, (unsigned long long)acc);
}

template <int D> void build_levels(abh_context *ctx, abh::Octree &t) {
  constexpr int NCH = 1 << D;
  constexpr int ML = max_level_of<D>();
  hipStream_t s = ctx->stream;
  const int64_t n = t.n_alive;
  const int threshold = int(t.n_leaf);  // classify_node's int threshold
  auto grow = [&](DevBuf &buf, size_t used, size_t want) {
    if (buf.p && want <= buf.bytes) return;
    DevBuf nb;
    nb.ensure(std::max(want, 2 * buf.bytes));
    if (used && buf.p) ABH_HIP(hipMemcpyAsync(nb.p, buf.p, used, hipMemcpyDeviceToDevice, s));
    ABH_HIP(hipStreamSynchronize(s));
    std::swap(buf.p, nb.p);
    std::swap(buf.bytes, nb.bytes);
  };
  t.act_a.ensure(sizeof(Act));
  const Act root{0u, -1, 0, int32_t(n)};
  ABH_HIP(hipMemcpyAsync(t.act_a.p, &root, sizeof(Act), hipMemcpyHostToDevice, s));
  t.total.ensure(16);
  t.level_begin.assign(1, 0);
  t.leaf_begin.assign(1, 0);
  t.par_off.assign(2, 0);  // index 0 unused, level 1 starts at 0
  int64_t na = 1, cb = 0, leaf_base = 0, int_base = 0;
  int level = 0;
  while (na > 0) {
    ++level;
    ABH_REQUIRE(level <= ML, ABH_ERR_INVALID_ARG, "octree exceeded its maximum depth");
    const int64_t nch = na * NCH;
    ABH_REQUIRE(cb + nch < (int64_t(1) << 31), ABH_ERR_CAPACITY, "octree node count too large");
    grow(t.nodes, size_t(cb) * 4, size_t(cb + nch) * 4);
    grow(t.ntag, size_t(cb) * 4, size_t(cb + nch) * 4);
    grow(t.parent, size_t(cb) * 4, size_t(cb + nch) * 4);
    grow(t.leaves, size_t(leaf_base) * 8, size_t(leaf_base + nch) * 8);
    grow(t.leaf_node, size_t(leaf_base) * 4, size_t(leaf_base + nch) * 4);
    grow(t.internal, size_t(int_base) * 4, size_t(int_base + nch) * 4);
    t.range.ensure(size_t(nch) * 8);
    t.cflags.ensure(size_t(nch) * 16);
    int64_t *cf = t.cflags.as<int64_t>();
    int64_t *sc = cf + nch;
    const unsigned g = unsigned(abh::ceil_div(nch, kThreads));
    k_oct_children<D><<<g, kThreads, 0, s>>>(t.act_a.as<Act>(), nch, t.keys.as<uint32_t>(), level,
                                             threshold, cb, t.ntag.as<uint32_t>(),
                                             t.parent.as<int32_t>(), t.range.as<int2>(), cf);
    abh::check_launch();
    abh::scan_exclusive(ctx, cf, sc, nch, t.total.as<int64_t>());
    int64_t tot = 0;
    ABH_HIP(hipMemcpyAsync(&tot, t.total.p, 8, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    const int64_t nl = tot & 0xffffffff, nn = tot >> 32;
    t.act_b.ensure(size_t(nn + 1) * sizeof(Act));
    k_oct_assign<D><<<g, kThreads, 0, s>>>(nch, cb, leaf_base, int_base, t.ntag.as<uint32_t>(),
                                           t.range.as<int2>(), cf, sc, t.nodes.as<int32_t>(),
                                           t.leaves.as<int2>(), t.leaf_node.as<int32_t>(),
                                           t.internal.as<int32_t>(), t.act_b.as<Act>());
    abh::check_launch();
    cb += nch;
    leaf_base += nl;
    int_base += nn;
    t.level_begin.push_back(cb);
    t.leaf_begin.push_back(leaf_base);
    t.par_off.push_back(int_base);  // par_off[level + 1]
    std::swap(t.act_a.p, t.act_b.p);
    std::swap(t.act_a.bytes, t.act_b.bytes);
    na = nn;
  }
  t.levels = level;
  t.n_nodes = cb;
  t.n_leaves = leaf_base;
  ABH_HIP(hipStreamSynchronize(s));
}

template <int D> void update_impl(abh_context *ctx, abh::Octree &t, int64_t n, double *pos,
                                  uint8_t *alive) {
  hipStream_t s = ctx->stream;
  abh::GridParams g{};
  for (int d = 0; d < 3; ++d) {
    const bool on = d < D;
    g.lo[d] = on ? t.lo[d] : 0.0;
    g.hi[d] = on ? t.hi[d] : 1.0;
    g.L[d] = g.hi[d] - g.lo[d];
    g.periodic[d] = on && t.periodic[d];
  }
  t.counters.ensure(sizeof(abh::BuildCounters));
  abh::BuildCounters *ctr = t.counters.as<abh::BuildCounters>();
  ABH_HIP(hipMemsetAsync(ctr, 0, sizeof(abh::BuildCounters), s));
  t.flags.ensure(size_t(n + 1) * 4);
  t.offs.ensure(size_t(n + 1) * 4);
  uint32_t *flags = t.flags.as<uint32_t>();
  uint32_t *offs = t.offs.as<uint32_t>();
  abh::enforce_domain(ctx, D, g, n, pos, alive, flags, ctr);
  abh::scan_exclusive(ctx, flags, offs, n, &ctr->n_alive);
  abh::BuildCounters h{};
  ABH_HIP(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
  ABH_HIP(hipStreamSynchronize(s));
  const int64_t na = int64_t(h.n_alive);
  t.n_alive = na;
  t.keys.ensure(size_t(na + 1) * 4);
  t.order.ensure(size_t(na + 1) * 4);
  t.pos.ensure(size_t(na + 1) * D * 8);
  if (na > 0) {
    OctBounds b{};
    for (int d = 0; d < D; ++d) {
      b.lo[d] = t.lo[d];
      b.hi[d] = t.hi[d];
    }
    k_oct_tag<D><<<unsigned(abh::ceil_div(n, kThreads)), kThreads, 0, s>>>(
        pos, flags, offs, n, b, t.keys.as<uint32_t>(), t.order.as<int32_t>());
    abh::check_launch();
    abh::radix_sort_pairs(ctx, t.keys.as<uint32_t>(), t.order.as<int32_t>(), na,
                          D * max_level_of<D>());
    abh::gather_bytes(ctx, t.order.as<int32_t>(), na, pos, t.pos.p, size_t(D) * 8);
  }
  build_levels<D>(ctx, t);
}

template <int D> void geometry_impl(abh_context *ctx, const abh::Octree &t, double *box,
                                    int32_t *coord) {
  OctBounds root{};
  for (int d = 0; d < D; ++d) {
    root.lo[d] = t.lo[d];
    root.hi[d] = t.hi[d];
  }
  for (int l = 1; l <= t.levels; ++l) {
    const int64_t b = t.level_begin[size_t(l) - 1], cnt = t.level_begin[size_t(l)] - b;
    if (cnt == 0) continue;
    k_oct_geometry<D><<<unsigned(abh::ceil_div(cnt, kThreads)), kThreads, 0, ctx->stream>>>(
        b, cnt, l, t.ntag.as<uint32_t>(), root, box, coord);
    abh::check_launch();
  }
}

template <int D>
void lists_impl(abh_context *ctx, const abh::Octree &t, const double *box, abh::OctreeLists &L,
                bool weak_lists) {
  hipStream_t s = ctx->stream;
  const int64_t nn = t.n_nodes, nleaf = t.n_leaves;
  DevBuf s_cnt, s_ptr, s_idx, sub, ctr;
  s_cnt.ensure(size_t(nn + 1) * 4);
  s_ptr.ensure(size_t(nn + 1) * 4);
  sub.ensure(size_t(nn + 1) * 4);
  ctr.ensure(32);
  L.weak_ptr.ensure(size_t(nn + 1) * 4);
  uint32_t *wptr = weak_lists ? L.weak_ptr.as<uint32_t>() : nullptr;
  unsigned long long *m2l = ctr.as<unsigned long long>();
  unsigned long long *inter = m2l + 1;
  uint32_t *tot = reinterpret_cast<uint32_t *>(m2l + 2);
  ABH_HIP(hipMemsetAsync(ctr.p, 0, 32, s));
  // strong / weak connections, level by level
  uint64_t s_total = 0, w_total = 0;
  uint32_t wt = 0;  // host source of an asynchronous copy: lives to the final synchronise
  std::vector<uint32_t> s_base(size_t(t.levels) + 1), w_base(size_t(t.levels) + 1);
  for (int l = 1; l <= t.levels; ++l) {
    const int64_t b = t.level_begin[size_t(l) - 1], cnt = t.level_begin[size_t(l)] - b;
    const unsigned g = unsigned(abh::ceil_div(cnt, kThreads));
    k_oct_conn<D, false><<<g, kThreads, 0, s>>>(l, b, cnt, t.nodes.as<int32_t>(),
                                             t.parent.as<int32_t>(), box, s_cnt.as<int32_t>(),
                                             s_ptr.as<uint32_t>(), s_idx.as<int32_t>(), 0, wptr,
                                             nullptr, 0, m2l);
    abh::check_launch();
    abh::scan_exclusive(ctx, s_ptr.as<uint32_t>() + b, s_ptr.as<uint32_t>() + b, cnt, tot);
    if (wptr) abh::scan_exclusive(ctx, wptr + b, wptr + b, cnt, tot + 1);
    uint32_t h[2] = {0, 0};
    ABH_HIP(hipMemcpyAsync(h, tot, 8, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    ABH_REQUIRE(s_total + h[0] < (uint64_t(1) << 31) && w_total + h[1] < (uint64_t(1) << 31),
                ABH_ERR_CAPACITY, "interaction lists too large");
    s_base[size_t(l)] = uint32_t(s_total);
    w_base[size_t(l)] = uint32_t(w_total);
    // grow the entry arrays, keeping the previous levels' entries
    auto grow = [&](DevBuf &buf, uint64_t used, uint64_t want) {
      if (buf.p && want * 4 <= buf.bytes) return;
      DevBuf nb;
      nb.ensure(std::max(size_t(want) * 4, 2 * buf.bytes));
      if (used && buf.p)
        ABH_HIP(hipMemcpyAsync(nb.p, buf.p, size_t(used) * 4, hipMemcpyDeviceToDevice, s));
      ABH_HIP(hipStreamSynchronize(s));
      std::swap(buf.p, nb.p);
      std::swap(buf.bytes, nb.bytes);
    };
    grow(s_idx, s_total, s_total + h[0] + 1);
    if (wptr) grow(L.weak_src, w_total, w_total + h[1] + 1);
    k_oct_conn<D, true><<<g, kThreads, 0, s>>>(l, b, cnt, t.nodes.as<int32_t>(),
                                            t.parent.as<int32_t>(), box, s_cnt.as<int32_t>(),
                                            s_ptr.as<uint32_t>(), s_idx.as<int32_t>(),
                                            uint32_t(s_total), wptr, L.weak_src.as<int32_t>(),
                                            uint32_t(w_total), m2l);
    abh::check_launch();
    s_total += h[0];
    w_total += h[1];
  }
  if (wptr) {
    wt = uint32_t(w_total);
    ABH_HIP(hipMemcpyAsync(wptr + nn, &wt, 4, hipMemcpyHostToDevice, s));
  }
  // non-empty leaves per subtree, deepest level first
  for (int l = t.levels; l >= 1; --l) {
    const int64_t b = t.level_begin[size_t(l) - 1], cnt = t.level_begin[size_t(l)] - b;
    k_oct_sub<D><<<unsigned(abh::ceil_div(cnt, kThreads)), kThreads, 0, s>>>(
        b, cnt, t.nodes.as<int32_t>(), sub.as<int32_t>());
    abh::check_launch();
  }
  // P2P lists
  L.p2p_ptr.ensure(size_t(nleaf + 1) * 4);
  uint32_t *pptr = L.p2p_ptr.as<uint32_t>();
  uint32_t ptot = 0;
  if (nleaf > 0) {
    const unsigned g = unsigned(abh::ceil_div(nleaf, kThreads));
    k_oct_p2p<D, false><<<g, kThreads, 0, s>>>(nleaf, t.leaf_node.as<int32_t>(), t.leaves.as<int2>(),
                                            t.nodes.as<int32_t>(), sub.as<int32_t>(),
                                            s_cnt.as<int32_t>(), s_ptr.as<uint32_t>(),
                                            s_idx.as<int32_t>(), pptr, nullptr, inter);
    abh::check_launch();
    abh::scan_exclusive(ctx, pptr, pptr, nleaf, pptr + nleaf);
    ABH_HIP(hipMemcpyAsync(&ptot, pptr + nleaf, 4, hipMemcpyDeviceToHost, s));
    ABH_HIP(hipStreamSynchronize(s));
    ABH_REQUIRE(ptot < (uint32_t(1) << 31), ABH_ERR_CAPACITY, "P2P list too large");
    L.p2p_src.ensure(size_t(ptot + 1) * 4);
    k_oct_p2p<D, true><<<g, kThreads, 0, s>>>(nleaf, t.leaf_node.as<int32_t>(), t.leaves.as<int2>(),
                                           t.nodes.as<int32_t>(), sub.as<int32_t>(),
                                           s_cnt.as<int32_t>(), s_ptr.as<uint32_t>(),
                                           s_idx.as<int32_t>(), pptr, L.p2p_src.as<int32_t>(),
                                           inter);
    abh::check_launch();
  } else {
    ABH_HIP(hipMemsetAsync(pptr, 0, 4, s));
    L.p2p_src.ensure(4);
  }
  unsigned long long hc[2] = {0, 0};
  ABH_HIP(hipMemcpyAsync(hc, ctr.p, 16, hipMemcpyDeviceToHost, s));
  ABH_HIP(hipStreamSynchronize(s));  // the temporaries above go out of scope
  L.m2l = int64_t(hc[0]);
  L.interactions = int64_t(hc[1]);
  L.p2p_pairs = int64_t(ptot);
}

}  // namespace

namespace abh {

void octree_update(abh_context *ctx, Octree &t, int64_t n, double *pos, uint8_t *alive) {
  ABH_REQUIRE(n >= 0 && n < (int64_t(1) << 31), ABH_ERR_INVALID_ARG, "n out of range");
  ABH_REQUIRE(n == 0 || pos, ABH_ERR_INVALID_ARG, "pos is null");
  if (t.dim == 1) update_impl<1>(ctx, t, n, pos, alive);
  else if (t.dim == 2) update_impl<2>(ctx, t, n, pos, alive);
  else update_impl<3>(ctx, t, n, pos, alive);
}

void octree_geometry(abh_context *ctx, const Octree &t, double *box, int32_t *coord) {
  if (t.dim == 1) geometry_impl<1>(ctx, t, box, coord);
  else if (t.dim == 2) geometry_impl<2>(ctx, t, box, coord);
  else geometry_impl<3>(ctx, t, box, coord);
}

void octree_lists(abh_context *ctx, const Octree &t, const double *box, OctreeLists &L,
                  bool weak_lists) {
  if (t.dim == 1) lists_impl<1>(ctx, t, box, L, weak_lists);
  else if (t.dim == 2) lists_impl<2>(ctx, t, box, L, weak_lists);
  else lists_impl<3>(ctx, t, box, L, weak_lists);
}

}  // namespace abh

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

struct abh_octree {
  abh_context *ctx = nullptr;
  abh::Octree t;
};

extern "C" {

abh_status abh_octree_create(abh_context *ctx, int dim, const double *low, const double *high,
                             const uint8_t *periodic, double n_particles_in_leaf,
                             abh_octree **out) {
  if (!ctx || !out || !low || !high) return ABH_ERR_INVALID_ARG;
  *out = nullptr;
  abh_octree *o = new (std::nothrow) abh_octree();
  if (!o) return ABH_ERR_OUT_OF_MEMORY;
  abh_status st = abh::guard(ctx, [&] {
    ABH_REQUIRE(!ctx->closed, ABH_ERR_INVALID_ARG, "context already destroyed");
    ABH_REQUIRE(dim >= 1 && dim <= 3, ABH_ERR_UNSUPPORTED, "dim must be 1, 2 or 3");
    ABH_REQUIRE(n_particles_in_leaf > 0.0 && n_particles_in_leaf < 2147483648.0,
                ABH_ERR_INVALID_ARG, "n_particles_in_leaf must be in (0, 2^31)");
    o->ctx = ctx;
    o->t.dim = dim;
    o->t.max_level = 32 / dim - 2;
    o->t.n_leaf = n_particles_in_leaf;
    for (int d = 0; d < dim; ++d) {
      ABH_REQUIRE(std::isfinite(low[d]) && std::isfinite(high[d]) && high[d] > low[d],
                  ABH_ERR_INVALID_ARG, "domain high must exceed low");
      o->t.lo[d] = low[d];
      o->t.hi[d] = high[d];
      o->t.periodic[d] = periodic && periodic[d];
    }
    // an empty tree: one level of 2^dim empty boxes (src/OctTree.h:133-144)
    abh::octree_update(ctx, o->t, 0, nullptr, nullptr);
  });
  if (st != ABH_OK) {
    delete o;
    return st;
  }
  abh::ctx_retain(ctx);
  *out = o;
  return ABH_OK;
}

abh_status abh_octree_destroy(abh_octree *o) {
  if (!o) return ABH_OK;
  abh_context *ctx = o->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete o;
  abh::ctx_release(ctx);
  return ABH_OK;
}

abh_status abh_octree_update(abh_octree *o, int64_t n, double *pos, uint8_t *alive,
                             int64_t *n_alive_out) {
  if (!o) return ABH_ERR_INVALID_ARG;
  return abh::guard(o->ctx, [&] {
    abh::octree_update(o->ctx, o->t, n, pos, alive);
    if (n_alive_out) *n_alive_out = o->t.n_alive;
  });
}

abh_status abh_octree_get_info(const abh_octree *o, abh_octree_info *info) {
  if (!o || !info) return ABH_ERR_INVALID_ARG;
  std::memset(info, 0, sizeof(*info));
  info->dim = o->t.dim;
  info->levels = o->t.levels;
  info->n_alive = o->t.n_alive;
  info->n_nodes = o->t.n_nodes;
  info->n_leaves = o->t.n_leaves;
  return ABH_OK;
}

const int32_t *abh_octree_order(const abh_octree *o) { return o ? o->t.order.as<int32_t>() : nullptr; }
const double *abh_octree_positions(const abh_octree *o) { return o ? o->t.pos.as<double>() : nullptr; }
const uint32_t *abh_octree_tags(const abh_octree *o) { return o ? o->t.keys.as<uint32_t>() : nullptr; }
const int32_t *abh_octree_nodes(const abh_octree *o) { return o ? o->t.nodes.as<int32_t>() : nullptr; }
const int32_t *abh_octree_leaves(const abh_octree *o) { return o ? o->t.leaves.as<int32_t>() : nullptr; }

abh_status abh_octree_query_view(const abh_octree *o, abh_octree_view *v) {
  if (!o || !v) return ABH_ERR_INVALID_ARG;
  std::memset(v, 0, sizeof(*v));
  const abh::Octree &t = o->t;
  v->dim = t.dim;
  v->levels = t.levels;
  for (int d = 0; d < 3; ++d) {
    const bool on = d < t.dim;
    v->periodic[d] = on && t.periodic[d];
    v->low[d] = on ? t.lo[d] : 0.0;
    v->high[d] = on ? t.hi[d] : 1.0;
  }
  v->nodes = t.nodes.as<int32_t>();
  v->leaves = t.leaves.as<int32_t>();
  v->pos = t.pos.as<double>();
  v->order = t.order.as<int32_t>();
  v->n = t.n_alive;
  v->n_nodes = t.n_nodes;
  v->n_leaves = t.n_leaves;
  return ABH_OK;
}

abh_status abh_octree_interaction_lists(const abh_octree *o, int32_t *p2p_ptr, int32_t *p2p_src,
                                        int32_t *m2l_ptr, int32_t *m2l_src, int64_t *counts) {
  if (!o || !counts) return ABH_ERR_INVALID_ARG;
  return abh::guard(o->ctx, [&] {
    const abh::Octree &t = o->t;
    hipStream_t s = o->ctx->stream;
    DevBuf box, coord;
    box.ensure(size_t(t.n_nodes + 1) * 16 * t.dim);
    coord.ensure(size_t(t.n_nodes + 1) * 4 * t.dim);
    abh::octree_geometry(o->ctx, t, box.as<double>(), coord.as<int32_t>());
    abh::OctreeLists L;
    const bool weak = m2l_ptr || m2l_src;
    abh::octree_lists(o->ctx, t, box.as<double>(), L, weak);
    auto copy = [&](void *dst, const DevBuf &src, int64_t elems) {
      if (dst && elems > 0)
        ABH_HIP(hipMemcpyAsync(dst, src.p, size_t(elems) * 4, hipMemcpyDeviceToDevice, s));
    };
    copy(p2p_ptr, L.p2p_ptr, t.n_leaves + 1);
    copy(p2p_src, L.p2p_src, L.p2p_pairs);
    if (weak) {
      copy(m2l_ptr, L.weak_ptr, t.n_nodes + 1);
      copy(m2l_src, L.weak_src, L.m2l);
    }
    ABH_HIP(hipStreamSynchronize(s));
    counts[0] = L.p2p_pairs;
    counts[1] = L.m2l;
    counts[2] = L.interactions;
  });
}

}  // extern "C"
