/*
 * aboria_hip_device.hpp -- device-side neighbour queries for HIP kernels
 * written against libaboria_hip.so's cell lists and octrees.
 *
 * abh_query_view (aboria_hip.h, filled by abh_cells_query_view) is a plain
 * struct: pass it to a kernel by value, as the reference copies its
 * CellListOrderedQuery into device lambdas (tests/parallel.h:220-228), and
 * call abh::for_each_neighbour from device code.  The visit is the one
 * euclidean_search(query, centre, radius) performs (src/Search.h:64-496,
 * :839-845): every periodic image of the centre in lattice order over
 * {-1,0,1} per periodic dimension, for each image the buckets
 * lattice_iterator_within_distance visits (src/NeighbourSearchBase.h:
 * 1720-2009: the 2^D quadrant walks from the bucket nearest the point, a
 * bucket visited while its box lies within the radius), and in each bucket
 * the particles with sum_d (x_j - image)_d^2 <= radius^2 in FP64.  A bucket
 * reached through two images is visited twice, as the reference does; a
 * non-finite centre has no neighbours.
 *
 * Every FP64 operation that decides a bucket or a neighbour is written with
 * explicitly rounded intrinsics (__dadd_rn, __dsub_rn, __dmul_rn), so the
 * result is bit-exact with the library's own search and the reference's
 * x86-64 build whatever contraction flags the including code is compiled
 * with.
 *
 *   __global__ void k(abh_query_view q, const double *centres, double r, int *count) {
 *     const int i = blockIdx.x * blockDim.x + threadIdx.x;
 *     int c = 0;
 *     abh::for_each_neighbour(q, centres + 3 * i, r,
 *                             [&](uint32_t j, const double *dx, double d2) { ++c; });
 *     count[i] = c;
 *   }
 *
 * abh_octree_view (abh_octree_query_view) is the HyperOctreeQuery
 * counterpart: abh::octree_get_bucket, abh::octree_visit_buckets (the
 * tree_query_iterator walk) and abh::octree_for_each_neighbour (the same
 * euclidean_search visit over the tree's leaf buckets), with the same
 * explicitly rounded FP64 operations.
 */
#ifndef ABORIA_HIP_DEVICE_HPP_
#define ABORIA_HIP_DEVICE_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aboria_hip.h"

namespace abh {

// box distance^2 from point p to bucket v
// (get_min_distance_to_bucket with find_bucket_centre, src/NeighbourSearchBase.h:1849-1858)
template <int D>
__device__ __forceinline__ double view_bucket_dist2(const abh_query_view &q, const double *p,
                                                    const int *v) {
  double acc = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double centre = __dadd_rn(__dmul_rn(__dadd_rn(double(v[d]), 0.5), q.side[d]), q.low[d]);
    const double half = __dmul_rn(0.5, q.side[d]);
    double x = __dsub_rn(fabs(__dsub_rn(centre, p[d])), half);
    x = x > 0.0 ? x : 0.0;
    acc = __dadd_rn(acc, __dmul_rn(x, x));
  }
  return acc;
}

// lattice_iterator_within_distance::outside_domain (src/NeighbourSearchBase.h:1953-1963)
template <int D>
__device__ __forceinline__ bool view_outside(const abh_query_view &q, const double *p, double r2) {
  double acc = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double mid = __dmul_rn(0.5, __dadd_rn(q.low[d], q.high[d]));
    const double half = __dmul_rn(0.5, __dsub_rn(q.high[d], q.low[d]));
    double x = __dsub_rn(fabs(__dsub_rn(mid, p[d])), half);
    x = x > 0.0 ? x : 0.0;
    acc = __dadd_rn(acc, __dmul_rn(x, x));
  }
  return acc > r2;
}

// The buckets visited for one image point p, in the reference's order; fn(b)
// gets the collapsed bucket index (last dimension fastest).
template <int D, class F>
__device__ void view_visit_buckets(const abh_query_view &q, const double *p, double r2, F &&fn) {
  if (view_outside<D>(q, p, r2)) return;
  for (int quad = 0; quad < (1 << D); ++quad) {
    int mn[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const bool up = (quad >> d) & 1;
      const double shifted = __dadd_rn(p[d], __dmul_rn(up ? 0.5 : -0.5, q.side[d]));
      mn[d] = int(floor(__dmul_rn(__dsub_rn(shifted, q.low[d]), q.inv_side[d])));
    }
    bool empty = view_bucket_dist2<D>(q, p, mn) > r2;
    if (!empty) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const bool up = (quad >> d) & 1;
        if (up) {
          if (mn[d] < 0) mn[d] = 0;
          else if (mn[d] > q.end_bucket[d]) { empty = true; mn[d] = q.end_bucket[d]; }
        } else {
          if (mn[d] < 0) { empty = true; mn[d] = 0; }
          else if (mn[d] > q.end_bucket[d]) mn[d] = q.end_bucket[d];
        }
      }
    }
    if (empty) continue;
    int idx[D];
#pragma unroll
    for (int d = 0; d < D; ++d) idx[d] = mn[d];
    for (;;) {
      uint32_t b = 0, mult = 1;
      for (int d = D - 1; d >= 0; --d) {
        b += mult * uint32_t(idx[d]);
        mult *= uint32_t(q.size[d]);
      }
      fn(b);
      int i = D - 1;
      for (; i >= 0; --i) {
        const bool up = (quad >> i) & 1;
        bool ok;
        if (up) { ++idx[i]; ok = idx[i] <= q.end_bucket[i]; }
        else { --idx[i]; ok = idx[i] >= 0; }
        if (ok) ok = view_bucket_dist2<D>(q, p, idx) <= r2;
        if (ok) break;
        idx[i] = mn[i];
      }
      if (i < 0) break;
    }
  }
}

// fn(j, dx, d2) for every neighbour j (index into the view's positions) of
// centre c within radius r, dx[d] = x_j - image, d2 = sum_d dx_d^2.
template <int D, class F>
__device__ void for_each_neighbour_d(const abh_query_view &q, const double *c, double r, F &&fn) {
  const double r2 = __dmul_rn(r, r);
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (!isfinite(c[d])) return;
  int lo[D], hi[D], img[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = q.periodic[d] ? -1 : 0;
    hi[d] = q.periodic[d] ? 2 : 1;
    img[d] = lo[d];
  }
  for (;;) {
    double p[D];
#pragma unroll
    for (int d = 0; d < D; ++d)
      p[d] = __dadd_rn(c[d], __dmul_rn(double(img[d]), __dsub_rn(q.high[d], q.low[d])));
    view_visit_buckets<D>(q, p, r2, [&](uint32_t b) {
      for (uint32_t j = q.bucket_begin[b]; j < q.bucket_end[b]; ++j) {
        double dx[D];
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          dx[d] = __dsub_rn(q.pos[size_t(j) * D + d], p[d]);
          acc = __dadd_rn(acc, __dmul_rn(dx[d], dx[d]));
        }
        if (!(acc > r2)) fn(j, dx, acc);
      }
    });
    int d = D - 1;
    for (; d >= 0; --d) {
      if (++img[d] < hi[d]) break;
      img[d] = lo[d];
    }
    if (d < 0) break;
  }
}

// for_each_neighbour_d for the view's dimension (1, 2 or 3)
template <class F>
__device__ void for_each_neighbour(const abh_query_view &q, const double *c, double r, F &&fn) {
  if (q.dim == 1) for_each_neighbour_d<1>(q, c, r, fn);
  else if (q.dim == 2) for_each_neighbour_d<2>(q, c, r, fn);
  else for_each_neighbour_d<3>(q, c, r, fn);
}

// ---------------------------------------------------------------------------
// HyperOctree queries on an abh_octree_view (abh_octree_query_view)
// ---------------------------------------------------------------------------

// detail::is_leaf / is_empty / get_leaf_offset (src/detail/SpatialUtil.h:
// 228-240): a negative child code is a leaf node (is_leaf_node), 0xffffffff
// an empty one
__device__ __forceinline__ bool octree_is_leaf_node(int32_t code) { return code < 0; }
__device__ __forceinline__ bool octree_is_empty(int32_t code) {
  return code == int32_t(0xffffffffu);
}
__device__ __forceinline__ int32_t octree_leaf_index(int32_t code) {
  return int32_t(0x80000000u ^ uint32_t(code));
}

// octree_child_iterator::get_bounds (src/OctTree.h:633-644): child c halves
// its parent's box, bit D-1-d of c selecting the high half of dimension d
template <int D>
__device__ __forceinline__ void octree_child_box(const double *plo, const double *phi, int c,
                                                 double *lo, double *hi) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double mid = __dmul_rn(0.5, __dadd_rn(phi[d], plo[d]));
    const bool high = (c >> (D - 1 - d)) & 1;
    lo[d] = high ? mid : plo[d];
    hi[d] = high ? phi[d] : mid;
  }
}

// HyperOctreeQuery::get_bucket (src/OctTree.h:810-822, go_to :598-612):
// the node index of the leaf (possibly empty) whose box holds p, which must
// lie inside the domain
template <int D>
__device__ int64_t octree_get_bucket(const abh_octree_view &q, const double *p) {
  double lo[D], hi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = q.low[d];
    hi[d] = q.high[d];
  }
  int64_t base = 0;
  for (;;) {
    int which = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double mid = __dmul_rn(0.5, __dadd_rn(hi[d], lo[d]));
      const bool high = !(p[d] < mid);
      which = (which << 1) | int(high);
      if (high) lo[d] = mid;
      else hi[d] = mid;
    }
    const int64_t k = base + which;
    const int32_t code = q.nodes[k];
    if (code < 0) return k;
    base = code;
  }
}

// tree_query_iterator (src/NeighbourSearchBase.h:1439-1716) from
// get_buckets_near_point (src/OctTree.h:836-848): depth first from the
// root's children, children in order, entering a box whose squared
// distance to p (get_dist_to_bucket: per dimension max(|centre - p| - half
// side, 0)) is below r2, and calling fn(node, code) for every leaf node
// reached, empty ones included
template <int D, class F>
__device__ void octree_visit_buckets(const abh_octree_view &q, const double *p, double r2,
                                     F &&fn) {
  constexpr int kDepth = 32 / D - 2;
  int64_t base[kDepth];
  int cur[kDepth];
  double blo[kDepth][D], bhi[kDepth][D];
  int top = 0;
  base[0] = 0;
  cur[0] = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    blo[0][d] = q.low[d];
    bhi[0][d] = q.high[d];
  }
  while (top >= 0) {
    if (cur[top] == (1 << D)) {
      --top;
      continue;
    }
    const int c = cur[top]++;
    const int64_t k = base[top] + c;
    double lo[D], hi[D];
    octree_child_box<D>(blo[top], bhi[top], c, lo, hi);
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double half = __dmul_rn(0.5, __dsub_rn(hi[d], lo[d]));
      const double centre = __dmul_rn(0.5, __dadd_rn(lo[d], hi[d]));
      double x = __dsub_rn(fabs(__dsub_rn(centre, p[d])), half);
      x = x > 0.0 ? x : 0.0;
      acc = __dadd_rn(acc, __dmul_rn(x, x));
    }
    if (!(acc < r2)) continue;
    const int32_t code = q.nodes[k];
    if (code < 0) {
      fn(k, code);
      continue;
    }
    ++top;  // a node below the deepest level has no children: top < kDepth
    base[top] = code;
    cur[top] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      blo[top][d] = lo[d];
      bhi[top][d] = hi[d];
    }
  }
}

// euclidean_search over a HyperOctree (src/Search.h:64-496, :839-845):
// every periodic image of the centre in lattice order, for each the leaf
// buckets octree_visit_buckets reaches, and in each the particles with
// sum_d (x_j - image)_d^2 <= r^2.  fn(j, dx, d2) with j the tree-order
// index (view.order[j] is the input index).  A non-finite centre has no
// neighbours.
template <int D, class F>
__device__ void octree_for_each_neighbour_d(const abh_octree_view &q, const double *c, double r,
                                            F &&fn) {
  const double r2 = __dmul_rn(r, r);
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (!isfinite(c[d])) return;
  int lo[D], hi[D], img[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = q.periodic[d] ? -1 : 0;
    hi[d] = q.periodic[d] ? 2 : 1;
    img[d] = lo[d];
  }
  for (;;) {
    double p[D];
#pragma unroll
    for (int d = 0; d < D; ++d)
      p[d] = __dadd_rn(c[d], __dmul_rn(double(img[d]), __dsub_rn(q.high[d], q.low[d])));
    octree_visit_buckets<D>(q, p, r2, [&](int64_t, int32_t code) {
      if (octree_is_empty(code)) return;
      const int32_t li = octree_leaf_index(code);
      const int32_t b = q.leaves[2 * li], e = q.leaves[2 * li + 1];
      for (int32_t j = b; j < e; ++j) {
        double dx[D];
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          dx[d] = __dsub_rn(q.pos[size_t(j) * D + d], p[d]);
          acc = __dadd_rn(acc, __dmul_rn(dx[d], dx[d]));
        }
        if (!(acc > r2)) fn(uint32_t(j), dx, acc);
      }
    });
    int d = D - 1;
    for (; d >= 0; --d) {
      if (++img[d] < hi[d]) break;
      img[d] = lo[d];
    }
    if (d < 0) break;
  }
}

// octree_for_each_neighbour_d for the view's dimension (1, 2 or 3)
template <class F>
__device__ void octree_for_each_neighbour(const abh_octree_view &q, const double *c, double r,
                                          F &&fn) {
  if (q.dim == 1) octree_for_each_neighbour_d<1>(q, c, r, fn);
  else if (q.dim == 2) octree_for_each_neighbour_d<2>(q, c, r, fn);
  else octree_for_each_neighbour_d<3>(q, c, r, fn);
}

}  // namespace abh

#endif  // ABORIA_HIP_DEVICE_HPP_
