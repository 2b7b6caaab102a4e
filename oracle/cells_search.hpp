// CPU oracle: CellListOrdered build and euclidean_search restated from the
// reference's CPU path.  TEST INFRASTRUCTURE ONLY -- linked by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
// product library.
//
// Compiled with -ffp-contract=off so every FP64 expression rounds exactly as
// written (the reference's x86-64 build has no FMA contraction).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace orc {

// Grid state of a CellListOrdered (src/CellListOrdered.h:269-277 and the
// point_to_bucket_index it holds, src/detail/SpatialUtil.h:98-156).
template <int D> struct Grid {
  double lo[D], hi[D];
  bool periodic[D];
  double n_leaf;
  unsigned size[D];
  double side[D], inv[D];
  int end_bucket[D];
  size_t sized_with_n = std::numeric_limits<size_t>::max();
  size_t nbuckets() const {
    size_t p = 1;
    for (int d = 0; d < D; ++d) p *= size[d];
    return p;
  }
};

// CellListOrdered::set_domain_impl (src/CellListOrdered.h:132-186): bucket
// count from the mean number of particles per bucket; only recomputed when
// the particle count moved by a factor of two since the last sizing.
template <int D> bool size_grid(Grid<D> &g, size_t n) {
  if (!(n < 0.5 * g.sized_with_n || n > 2 * g.sized_with_n)) return false;
  g.sized_with_n = n;
  if (g.n_leaf > n) {
    for (int d = 0; d < D; ++d) g.size[d] = 1;
  } else {
    double volume = 1;
    for (int d = 0; d < D; ++d) volume *= (g.hi[d] - g.lo[d]);
    const double box_volume = g.n_leaf / double(n) * volume;
    const double box_side = std::pow(box_volume, 1.0 / D);
    for (int d = 0; d < D; ++d) {
      g.size[d] = static_cast<unsigned>(std::floor((g.hi[d] - g.lo[d]) / box_side));
      if (g.size[d] == 0) g.size[d] = 1;
    }
  }
  for (int d = 0; d < D; ++d) {
    g.side[d] = (g.hi[d] - g.lo[d]) / g.size[d];
    g.inv[d] = 1.0 / g.side[d];
    g.end_bucket[d] = int(g.size[d]) - 1;
  }
  return true;
}

// bucket_index::collapse_index_vector (src/detail/SpatialUtil.h:49-60):
// row-major, last dimension fastest, evaluated modulo 2^32.
template <int D> inline uint32_t collapse(const Grid<D> &g, const int *v) {
  uint32_t index = 0, mult = 1;
  for (int i = D - 1; i >= 0; --i) {
    if (i != D - 1) mult *= g.size[i + 1];
    index += mult * static_cast<uint32_t>(v[i]);
  }
  return index;
}

// point_to_bucket_index::find_bucket_index (src/detail/SpatialUtil.h:122-138)
template <int D> inline uint32_t bucket_of(const Grid<D> &g, const double *p) {
  int v[D];
  for (int d = 0; d < D; ++d)
    v[d] = static_cast<int>(std::floor((p[d] - g.lo[d]) * g.inv[d]));
  return collapse<D>(g, v);
}

// State after an update: m_alive_indices sorted by bucket, the sorted bucket
// keys, bucket_begin/end, and the particle positions in that order
// (what Particles::reorder leaves in the container).
template <int D> struct Cells {
  Grid<D> g;
  std::vector<int32_t> order;
  std::vector<uint32_t> keys;
  std::vector<uint32_t> begin, end;
  std::vector<double> pos;  // n_alive x D
};

// neighbour_search_base::update_positions (src/NeighbourSearchBase.h:350-495)
// with enforce_domain_lambda (:185-238), CellListOrdered::update_positions_impl
// (src/CellListOrdered.h:190-259) and the position gather of Particles::reorder
// (src/Particles.h:694-732).  The sort is stable by (bucket, alive index):
// the reference's CPU std::sort leaves within-bucket order unspecified, its
// Thrust sort_by_key is stable; stable is the order pinned here.
template <int D>
size_t update(Cells<D> &c, size_t n, double *pos, uint8_t *alive) {
  Grid<D> &g = c.g;
  for (size_t i = 0; i < n; ++i) {
    double *r = pos + i * D;
    for (int d = 0; d < D; ++d) {
      if (!std::isfinite(r[d])) {
        alive[i] = 0;
      } else if (g.periodic[d]) {
        const double L = g.hi[d] - g.lo[d];
        while (r[d] < g.lo[d]) r[d] += L;
        while (r[d] >= g.hi[d]) r[d] -= L;
      } else if (r[d] < g.lo[d] || r[d] >= g.hi[d]) {
        alive[i] = 0;
      }
    }
  }
  c.order.clear();
  for (size_t i = 0; i < n; ++i)
    if (alive[i]) c.order.push_back(int32_t(i));
  const size_t na = c.order.size();
  size_grid<D>(g, na);
  std::vector<uint32_t> key(na);
  for (size_t k = 0; k < na; ++k) key[k] = bucket_of<D>(g, pos + size_t(c.order[k]) * D);
  std::vector<size_t> idx(na);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t a, size_t b) { return key[a] < key[b]; });
  std::vector<int32_t> order2(na);
  c.keys.resize(na);
  for (size_t k = 0; k < na; ++k) {
    order2[k] = c.order[idx[k]];
    c.keys[k] = key[idx[k]];
  }
  c.order.swap(order2);
  const size_t nb = g.nbuckets();
  c.begin.resize(nb);
  c.end.resize(nb);
  for (size_t b = 0; b < nb; ++b) {
    c.begin[b] = uint32_t(std::lower_bound(c.keys.begin(), c.keys.end(), uint32_t(b)) - c.keys.begin());
    c.end[b] = uint32_t(std::upper_bound(c.keys.begin(), c.keys.end(), uint32_t(b)) - c.keys.begin());
  }
  c.pos.resize(na * D);
  for (size_t k = 0; k < na; ++k)
    for (int d = 0; d < D; ++d) c.pos[k * D + d] = pos[size_t(c.order[k]) * D + d];
  return na;
}

// Box distance^2 from point q to bucket v, as computed by
// lattice_iterator_within_distance::get_min_distance_to_bucket
// (src/NeighbourSearchBase.h:1849-1858) with find_bucket_centre
// (src/detail/SpatialUtil.h:144-147) and the L2 norm of
// distance_helper<2> (src/detail/Distance.h:46-139).
template <int D>
inline double bucket_dist2(const Grid<D> &g, const double *q, const int *v) {
  double acc = 0;
  for (int d = 0; d < D; ++d) {
    const double centre = (double(v[d]) + 0.5) * g.side[d] + g.lo[d];
    const double half = 0.5 * g.side[d];
    double x = std::abs(centre - q[d]) - half;
    x = std::max(x, 0.0);
    acc = acc + x * x;
  }
  return acc;
}

// lattice_iterator_within_distance::outside_domain
// (src/NeighbourSearchBase.h:1953-1963)
template <int D> inline bool outside_domain(const Grid<D> &g, const double *q, double r2) {
  double acc = 0;
  for (int d = 0; d < D; ++d) {
    const double mid = 0.5 * (g.lo[d] + g.hi[d]);
    const double half = 0.5 * (g.hi[d] - g.lo[d]);
    double x = std::abs(mid - q[d]) - half;
    x = std::max(x, 0.0);
    acc = acc + x * x;
  }
  return acc > r2;
}

// The buckets visited for one (image) point, in the reference's order: the
// 2^D quadrants in turn (quadrant bit d set = walk upwards in dimension d),
// each walked as an odometer from its nearest bucket with the last dimension
// fastest, a digit advancing while the bucket stays within r and inside the
// grid and resetting otherwise.  Restates reset_min_and_index and increment
// of lattice_iterator_within_distance (src/NeighbourSearchBase.h:1860-2008)
// with get_min_index_by_quadrant (src/detail/SpatialUtil.h:149-156).
template <int D, class Visit>
void visit_buckets(const Grid<D> &g, const double *q, double r2, Visit &&visit) {
  if (outside_domain<D>(g, q, r2)) return;
  for (int quad = 0; quad < (1 << D); ++quad) {
    int mn[D];
    for (int d = 0; d < D; ++d) {
      const bool up = (quad >> d) & 1;
      mn[d] = int(std::floor((q[d] + (up ? 0.5 : -0.5) * g.side[d] - g.lo[d]) * g.inv[d]));
    }
    bool empty = bucket_dist2<D>(g, q, mn) > r2;
    if (!empty) {
      for (int d = 0; d < D; ++d) {
        const bool up = (quad >> d) & 1;
        if (up) {
          if (mn[d] < 0) mn[d] = 0;
          else if (mn[d] > g.end_bucket[d]) { empty = true; mn[d] = g.end_bucket[d]; }
        } else {
          if (mn[d] < 0) { empty = true; mn[d] = 0; }
          else if (mn[d] > g.end_bucket[d]) mn[d] = g.end_bucket[d];
        }
      }
    }
    if (empty) continue;
    int idx[D];
    for (int d = 0; d < D; ++d) idx[d] = mn[d];
    for (;;) {
      visit(idx);
      int i = D - 1;
      for (; i >= 0; --i) {
        const bool up = (quad >> i) & 1;
        bool ok;
        if (up) { ++idx[i]; ok = idx[i] <= g.end_bucket[i]; }
        else { --idx[i]; ok = idx[i] >= 0; }
        if (ok) ok = bucket_dist2<D>(g, q, idx) <= r2;
        if (ok) break;
        idx[i] = mn[i];
      }
      if (i < 0) break;
    }
  }
}

// euclidean_search(query, centre, radius) (src/Search.h:64-496, 839-845):
// periodic images in lattice order over {-1,0,1} per periodic dimension
// (get_periodic_range, :152-159), image point = centre + image*(hi-lo),
// candidate kept iff sum_d (p_j - image point)_d^2 <= radius^2
// (check_candidate, :438-465).  visit(j, dx) is called in iteration order.
template <int D, class Visit>
void euclidean_search(const Cells<D> &c, const double *centre, double radius, Visit &&visit) {
  const Grid<D> &g = c.g;
  const double r2 = radius * radius;
  // A non-finite centre has no neighbours (the reference converts
  // floor(NaN) to int in get_min_index_by_quadrant, which is undefined).
  for (int d = 0; d < D; ++d)
    if (!std::isfinite(centre[d])) return;
  int lo_i[D], hi_i[D], img[D];
  for (int d = 0; d < D; ++d) {
    lo_i[d] = g.periodic[d] ? -1 : 0;
    hi_i[d] = g.periodic[d] ? 2 : 1;
    img[d] = lo_i[d];
  }
  for (;;) {
    double q[D];
    for (int d = 0; d < D; ++d) q[d] = centre[d] + double(img[d]) * (g.hi[d] - g.lo[d]);
    visit_buckets<D>(g, q, r2, [&](const int *v) {
      const uint32_t b = collapse<D>(g, v);
      for (uint32_t j = c.begin[b]; j < c.end[b]; ++j) {
        double dx[D];
        double acc = 0;
        for (int d = 0; d < D; ++d) {
          dx[d] = c.pos[size_t(j) * D + d] - q[d];
          acc = acc + dx[d] * dx[d];
        }
        if (!(acc > r2)) visit(j, dx, acc);
      }
    });
    int d = D - 1;
    for (; d >= 0; --d) {
      if (++img[d] < hi_i[d]) break;
      img[d] = lo_i[d];
    }
    if (d < 0) break;
  }
}

}  // namespace orc

namespace orc {

// lattice_iterator<D> over [start, end) (src/LatticeIterator.h:48-295): an
// odometer with the last dimension fastest; invalid once it runs past the
// last point (or when built empty).
template <int D> struct Lattice {
  int start[D], end[D], cur[D];
  bool valid = false;
  Lattice() = default;
  Lattice(const int *s, const int *e) {
    valid = true;
    for (int d = 0; d < D; ++d) {
      start[d] = s[d];
      end[d] = e[d];
      cur[d] = s[d];
      if (e[d] <= s[d]) valid = false;
    }
  }
  void set(const int *p) {
    for (int d = 0; d < D; ++d) cur[d] = p[d];
  }
  void inc() {
    for (int d = D - 1; d >= 0; --d) {
      if (++cur[d] < end[d]) return;
      cur[d] = start[d];
    }
    valid = false;
  }
  bool has_next() const {  // !(it + 1 == false)
    Lattice t = *this;
    t.inc();
    return t.valid;
  }
};

template <int D> struct BucketPair {
  int i[D], j[D], q[D];
};

// get_neighbouring_buckets(query, bucket) (src/Search.h:688-724): the 3^D
// block around `bucket` clipped to the grid, empty when it misses the grid.
template <int D> Lattice<D> neighbouring_buckets(const Grid<D> &g, const int *b) {
  int s[D], e[D];
  bool none = false;
  for (int d = 0; d < D; ++d) {
    s[d] = b[d] - 1;
    e[d] = b[d] + 1;
    if (s[d] < 0) s[d] = 0;
    else if (s[d] > g.end_bucket[d]) { none = true; s[d] = g.end_bucket[d]; }
    if (e[d] < 0) { none = true; e[d] = 0; }
    else if (e[d] > g.end_bucket[d]) e[d] = g.end_bucket[d];
    e[d] += 1;
  }
  if (none) return Lattice<D>();
  return Lattice<D>(s, e);
}

// get_regular_buckets(query, quadrant) (src/Search.h:739-770): every bucket
// for quadrant 0; the face a periodic quadrant's ghosts touch otherwise.
template <int D> Lattice<D> regular_buckets(const Grid<D> &g, const int *q) {
  int s[D], e[D];
  for (int d = 0; d < D; ++d) {
    s[d] = 0;
    e[d] = g.end_bucket[d];
    if (g.periodic[d]) {
      if (q[d] > 0) { s[d] = 0; e[d] = 0; }
      else if (q[d] < 0) { s[d] = g.end_bucket[d]; e[d] = g.end_bucket[d]; }
    }
    e[d] += 1;
  }
  return Lattice<D>(s, e);
}

// The sequence bucket_pair_iterator<Query> yields (src/Search.h:498-680,
// get_neighbouring_buckets(query) :857-860): the constructor and increment
// state machine restated over the lattice iterators above.  Each item is
// (bucket i, bucket j, periodic quadrant q); the position offset the
// reference yields is q * (high - low).
template <int D> void bucket_pairs(const Grid<D> &g, std::vector<BucketPair<D>> &out) {
  out.clear();
  int ps[D], pe[D], zero[D];
  for (int d = 0; d < D; ++d) {
    ps[d] = g.periodic[d] ? -1 : 0;
    pe[d] = g.periodic[d] ? 2 : 1;
    zero[d] = 0;
  }
  Lattice<D> per(ps, pe);
  per.set(zero);
  bool dd = true;
  Lattice<D> mi = regular_buckets<D>(g, per.cur), mj;
  auto ghost_block = [&]() {
    int b[D];
    for (int d = 0; d < D; ++d) b[d] = mi.cur[d] + per.cur[d] * (g.end_bucket[d] + 1);
    return neighbouring_buckets<D>(g, b);
  };
  if (!mi.has_next()) {
    dd = false;
    per.inc();
    if (!per.valid) return;
    mi = regular_buckets<D>(g, per.cur);
    mj = ghost_block();
  } else {
    mj = neighbouring_buckets<D>(g, mi.cur);
    mj.set(mi.cur);
    mj.inc();
  }
  for (;;) {
    BucketPair<D> p;
    for (int d = 0; d < D; ++d) {
      p.i[d] = mi.cur[d];
      p.j[d] = mj.cur[d];
      p.q[d] = per.cur[d];
    }
    out.push_back(p);
    // increment
    mj.inc();
    if (!mj.valid) {
      mi.inc();
      if (dd ? !mi.has_next() : !mi.valid) {
        dd = false;
        per.inc();
        if (!per.valid) return;
        mi = regular_buckets<D>(g, per.cur);
        mj = ghost_block();
      } else {
        mj = ghost_block();
        if (dd) {
          mj.set(mi.cur);
          mj.inc();
        }
      }
    }
  }
}

// Neighbour counts by bucket-pair traversal, as the reference's
// aboria_fast_bucketsearch_check_neighbour + _check_self functors do
// (tests/neighbours.h:892-951): for every pair (i, j, q) and particles
// a in i, b in j, |(x_a + q (high - low)) - x_b|^2 < r^2 adds one to both;
// within a bucket every particle counts itself and each later particle of
// the bucket within r adds one to both.  Sums in dimension order.
template <int D> void pair_counts(const Cells<D> &c, double r, int32_t *counts) {
  const Grid<D> &g = c.g;
  const double r2 = r * r;
  const size_t n = c.order.size();
  for (size_t k = 0; k < n; ++k) counts[k] = 0;
  std::vector<BucketPair<D>> pairs;
  bucket_pairs<D>(g, pairs);
  for (const BucketPair<D> &p : pairs) {
    const uint32_t bi = collapse<D>(g, p.i), bj = collapse<D>(g, p.j);
    double off[D];
    for (int d = 0; d < D; ++d) off[d] = double(p.q[d]) * (g.hi[d] - g.lo[d]);
    for (uint32_t a = c.begin[bi]; a < c.end[bi]; ++a) {
      double pa[D];
      for (int d = 0; d < D; ++d) pa[d] = c.pos[size_t(a) * D + d] + off[d];
      for (uint32_t b = c.begin[bj]; b < c.end[bj]; ++b) {
        double s = 0;
        for (int d = 0; d < D; ++d) {
          const double dx = pa[d] - c.pos[size_t(b) * D + d];
          s = s + dx * dx;
        }
        if (s < r2) {
          ++counts[a];
          ++counts[b];
        }
      }
    }
  }
  const size_t nb = g.nbuckets();
  for (size_t b = 0; b < nb; ++b)
    for (uint32_t a = c.begin[b]; a < c.end[b]; ++a) {
      ++counts[a];
      for (uint32_t e = a + 1; e < c.end[b]; ++e) {
        double s = 0;
        for (int d = 0; d < D; ++d) {
          const double dx = c.pos[size_t(a) * D + d] - c.pos[size_t(e) * D + d];
          s = s + dx * dx;
        }
        if (s < r2) {
          ++counts[a];
          ++counts[e];
        }
      }
    }
}

}  // namespace orc
