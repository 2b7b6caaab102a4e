// C entry points of the CPU oracle (liboracle.so), loaded by tests/ and by
// bench.py's cpu_baseline leg through ctypes.  TEST INFRASTRUCTURE ONLY.
#include <omp.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "cells_search.hpp"
#include "fmm.hpp"
#include "kernels.hpp"
#include "schwarz.hpp"

using namespace orc;

namespace {

struct AnyCells {
  int dim = 0;
  Cells<1> c1;
  Cells<2> c2;
  Cells<3> c3;
};

template <int D> Cells<D> &get(AnyCells *h);
template <> Cells<1> &get<1>(AnyCells *h) { return h->c1; }
template <> Cells<2> &get<2>(AnyCells *h) { return h->c2; }
template <> Cells<3> &get<3>(AnyCells *h) { return h->c3; }

template <int D>
void init(AnyCells *h, const double *lo, const double *hi, const uint8_t *per, double n_leaf) {
  Grid<D> &g = get<D>(h).g;
  for (int d = 0; d < D; ++d) {
    g.lo[d] = lo[d];
    g.hi[d] = hi[d];
    g.periodic[d] = per[d] != 0;
  }
  g.n_leaf = n_leaf;
  g.sized_with_n = std::numeric_limits<size_t>::max();
  // set_domain at init_neighbour_search time sizes with the (empty) alive set
  size_grid<D>(g, 0);
}

// 64-bit finaliser of SplitMix64 (Steele, Lea, Flood 2014) applied to an
// index: the per-entry term of the neighbour-list digest's fourth column.
inline uint64_t digest_mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <int D, int N> void fmm_n(const Octree<D> &t, const Kernel *k, const double *q, double *y) {
  FMM<D, N> f(t, *k);
  f.matvec(q, y);
}

template <class F> auto dispatch(int dim, F &&f) {
  if (dim == 1) return f(std::integral_constant<int, 1>());
  if (dim == 2) return f(std::integral_constant<int, 2>());
  return f(std::integral_constant<int, 3>());
}

}  // namespace

// one tree of dimension 1, 2 or 3 behind an opaque handle
struct OctH {
  int dim = 3;
  Octree<1> t1;
  Octree<2> t2;
  Octree<3> t3;
  std::vector<std::vector<int>> weak, p2p;  // interaction lists, filled on demand
  bool have_lists = false;
  template <class F> void with(F &&f) {
    if (dim == 1) f(t1);
    else if (dim == 2) f(t2);
    else f(t3);
  }
};

extern "C" {

void *orc_cells_create(int dim, const double *lo, const double *hi, const uint8_t *periodic,
                       double n_leaf) {
  if (dim < 1 || dim > 3) return nullptr;
  auto *h = new AnyCells();
  h->dim = dim;
  dispatch(dim, [&](auto DD) {
    init<decltype(DD)::value>(h, lo, hi, periodic, n_leaf);
    return 0;
  });
  return h;
}

void orc_cells_destroy(void *p) { delete static_cast<AnyCells *>(p); }

// pos (n x dim) and alive (n) are updated in place, as the reference does.
int64_t orc_cells_update(void *p, int64_t n, double *pos, uint8_t *alive) {
  auto *h = static_cast<AnyCells *>(p);
  return dispatch(h->dim, [&](auto DD) {
    return int64_t(update<decltype(DD)::value>(get<decltype(DD)::value>(h), size_t(n), pos, alive));
  });
}

// grid[dim], side[dim]; returns number of buckets
int64_t orc_cells_grid(void *p, uint32_t *grid, double *side) {
  auto *h = static_cast<AnyCells *>(p);
  return dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    const Grid<D> &g = get<D>(h).g;
    for (int d = 0; d < D; ++d) {
      grid[d] = g.size[d];
      side[d] = g.side[d];
    }
    return int64_t(g.nbuckets());
  });
}

void orc_cells_get(void *p, int32_t *order, uint32_t *keys, uint32_t *begin, uint32_t *end,
                   double *pos) {
  auto *h = static_cast<AnyCells *>(p);
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    const Cells<D> &c = get<D>(h);
    if (order) std::memcpy(order, c.order.data(), c.order.size() * 4);
    if (keys) std::memcpy(keys, c.keys.data(), c.keys.size() * 4);
    if (begin) std::memcpy(begin, c.begin.data(), c.begin.size() * 4);
    if (end) std::memcpy(end, c.end.data(), c.end.size() * 4);
    if (pos) std::memcpy(pos, c.pos.data(), c.pos.size() * 8);
    return 0;
  });
}

// counts[i] = number of neighbours of q[i]; returns the total.
int64_t orc_search_count(void *p, int64_t nq, const double *q, double r, int64_t *counts,
                         int nthreads) {
  auto *h = static_cast<AnyCells *>(p);
  return dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    const Cells<D> &c = get<D>(h);
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total) num_threads(nthreads)
    for (int64_t i = 0; i < nq; ++i) {
      int64_t cnt = 0;
      euclidean_search<D>(c, q + i * D, r, [&](uint32_t, const double *, double) { ++cnt; });
      counts[i] = cnt;
      total += cnt;
    }
    return total;
  });
}

// Fills col[row_ptr[i]..] with the neighbours of q[i] in the reference's
// iteration order.
void orc_search_fill(void *p, int64_t nq, const double *q, double r, const int64_t *row_ptr,
                     int32_t *col, int nthreads) {
  auto *h = static_cast<AnyCells *>(p);
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    const Cells<D> &c = get<D>(h);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int64_t i = 0; i < nq; ++i) {
      int64_t k = row_ptr[i];
      euclidean_search<D>(c, q + i * D, r,
                          [&](uint32_t j, const double *, double) { col[k++] = int32_t(j); });
    }
    return 0;
  });
}

// The bucket_pair_iterator sequence: returns the number of pairs; when bi is
// non-null writes collapsed bucket indices bi[p], bj[p] and the periodic
// quadrant q[p * 3 + d] (zero-padded to 3 entries) of each.
int64_t orc_bucket_pairs(void *p, int32_t *bi, int32_t *bj, int32_t *q) {
  auto *h = static_cast<AnyCells *>(p);
  return dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    const Cells<D> &c = get<D>(h);
    std::vector<BucketPair<D>> pairs;
    bucket_pairs<D>(c.g, pairs);
    if (bi)
      for (size_t k = 0; k < pairs.size(); ++k) {
        bi[k] = int32_t(collapse<D>(c.g, pairs[k].i));
        bj[k] = int32_t(collapse<D>(c.g, pairs[k].j));
        for (int d = 0; d < 3; ++d) q[k * 3 + d] = d < D ? pairs[k].q[d] : 0;
      }
    return int64_t(pairs.size());
  });
}

// neighbour counts by bucket-pair traversal (strict r), container order
void orc_pair_counts(void *p, double r, int32_t *counts) {
  auto *h = static_cast<AnyCells *>(p);
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    pair_counts<D>(get<D>(h), r, counts);
    return 0;
  });
}

// Order-independent digest of every neighbour list, so that lists of
// billions of pairs can be compared without storing them: per centre the
// count, the sum of the neighbour indices and the sum of their squares, and
// a sum of a 64-bit mix of each index (modulo 2^64), all in int64.
// digest: nq x 4 int64, row-major.
void orc_search_digest(void *p, int64_t nq, const double *q, double r, int64_t *digest,
                       int nthreads) {
  auto *h = static_cast<AnyCells *>(p);
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    const Cells<D> &c = get<D>(h);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int64_t i = 0; i < nq; ++i) {
      int64_t cnt = 0, s1 = 0, s2 = 0;
      uint64_t sh = 0;
      euclidean_search<D>(c, q + i * D, r, [&](uint32_t j, const double *, double) {
        ++cnt;
        s1 += int64_t(j);
        s2 += int64_t(j) * int64_t(j);
        sh += digest_mix(j);
      });
      digest[4 * i] = cnt;
      digest[4 * i + 1] = s1;
      digest[4 * i + 2] = s2;
      digest[4 * i + 3] = int64_t(sh);
    }
    return 0;
  });
}

// KernelSparse::evaluate (src/Kernels.h:720-751): y[i] += f(dx,a_i,b_j)*x[j]
// over the euclidean_search neighbours of row position q[i], in iteration
// order, one right-hand side at a time (the reference's GEMV-only product,
// src/detail/Operators.h:206-231).  x: column-major ncols x nrhs (ldx),
// y: nrows x nrhs (ldy).
void orc_sparse_matvec(void *p, int64_t nq, const double *q, double r, const Kernel *k,
                       const double *row_s, const double *col_s, int nrhs, const double *x,
                       int64_t ldx, double *y, int64_t ldy, int nthreads) {
  auto *h = static_cast<AnyCells *>(p);
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    const Cells<D> &c = get<D>(h);
    for (int rhs = 0; rhs < nrhs; ++rhs) {
      const double *xr = x + rhs * ldx;
      double *yr = y + rhs * ldy;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
      for (int64_t i = 0; i < nq; ++i) {
        double acc = yr[i];
        euclidean_search<D>(c, q + i * D, r, [&](uint32_t j, const double *, double s) {
          const double f = phi(*k, s) * weight(*k, row_s, col_s, size_t(i), j);
          acc += f * xr[j];
        });
        yr[i] = acc;
      }
    }
    return 0;
  });
}

// KernelDense::evaluate (src/Kernels.h:212-267): y[i] += f(a_i,b_j)*x[j] over
// all j in order.  Positions a (na x dim) and b (nb x dim).
void orc_dense_matvec(int dim, int64_t na, const double *a, int64_t nb, const double *b,
                      const Kernel *k, int nrhs, const double *x, int64_t ldx, double *y,
                      int64_t ldy, int nthreads) {
  for (int rhs = 0; rhs < nrhs; ++rhs) {
    const double *xr = x + rhs * ldx;
    double *yr = y + rhs * ldy;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < na; ++i) {
      double acc = yr[i];
      for (int64_t j = 0; j < nb; ++j) {
        double s = 0;
        for (int d = 0; d < dim; ++d) {
          const double dx = b[j * dim + d] - a[i * dim + d];
          s = s + dx * dx;
        }
        acc += phi(*k, s) * xr[j];
      }
      yr[i] = acc;
    }
  }
}

int orc_max_threads() { return omp_get_max_threads(); }

// ---------------------------------------------------------------------------
// HyperOctree + black-box Chebyshev FMM (3D, one RHS per call)
// ---------------------------------------------------------------------------

void *orc_octree_create(int dim, const double *lo, const double *hi, int64_t n,
                        const double *pos, double n_leaf) {
  auto *h = new OctH();
  h->dim = dim;
  h->with([&](auto &t) {
    constexpr int D = sizeof(t.bounds.lo) / sizeof(double);
    for (int d = 0; d < D; ++d) {
      t.bounds.lo[d] = lo[d];
      t.bounds.hi[d] = hi[d];
    }
    t.build(pos, size_t(n), n_leaf);
  });
  return h;
}
void orc_octree_destroy(void *p) { delete static_cast<OctH *>(p); }
// counts[0] = nodes, [1] = leaves, [2] = levels
void orc_octree_info(void *p, int64_t *counts) {
  static_cast<OctH *>(p)->with([&](auto &t) {
    counts[0] = int64_t(t.nodes.size());
    counts[1] = int64_t(t.leaves.size());
    counts[2] = t.levels;
  });
}
void orc_octree_get(void *p, int32_t *order, int32_t *tags, int32_t *nodes, int32_t *leaves) {
  static_cast<OctH *>(p)->with([&](auto &t) {
    if (order) std::memcpy(order, t.order.data(), t.order.size() * 4);
    if (tags) std::memcpy(tags, t.tags.data(), t.tags.size() * 4);
    if (nodes) std::memcpy(nodes, t.nodes.data(), t.nodes.size() * 4);
    if (leaves)
      for (size_t i = 0; i < t.leaves.size(); ++i) {
        leaves[2 * i] = t.leaves[i].first;
        leaves[2 * i + 1] = t.leaves[i].second;
      }
  });
}

// Interaction lists of the tree.  counts: [0] P2P entries, [1] M2L
// entries, [2] particle pairs of the P2P lists; arrays (may be NULL):
// p2p_ptr (leaves + 1), p2p_src, m2l_ptr (nodes + 1), m2l_src.
int orc_octree_lists(void *p, int64_t *counts, int32_t *p2p_ptr, int32_t *p2p_src,
                     int32_t *m2l_ptr, int32_t *m2l_src) {
  auto *h = static_cast<OctH *>(p);
  if (!h->have_lists) {
    h->with([&](auto &t) {
      constexpr int D = sizeof(t.bounds.lo) / sizeof(double);
      Lists<D> L;
      interaction_lists<D>(t, L);
      h->weak = std::move(L.weak);
      h->p2p = std::move(L.p2p);
    });
    h->have_lists = true;
  }
  int64_t np = 0, nm = 0, inter = 0;
  h->with([&](auto &t) {
    for (size_t l = 0; l < h->p2p.size(); ++l) {
      if (p2p_ptr) p2p_ptr[l] = int32_t(np);
      for (int s : h->p2p[l]) {
        if (p2p_src) p2p_src[np] = s;
        ++np;
        inter += int64_t(t.leaves[l].second - t.leaves[l].first) *
                 (t.leaves[size_t(s)].second - t.leaves[size_t(s)].first);
      }
    }
    if (p2p_ptr) p2p_ptr[h->p2p.size()] = int32_t(np);
  });
  for (size_t k = 0; k < h->weak.size(); ++k) {
    if (m2l_ptr) m2l_ptr[k] = int32_t(nm);
    for (int s : h->weak[k]) {
      if (m2l_src) m2l_src[nm] = s;
      ++nm;
    }
  }
  if (m2l_ptr) m2l_ptr[h->weak.size()] = int32_t(nm);
  counts[0] = np;
  counts[1] = nm;
  counts[2] = inter;
  return 0;
}

// y += K q with Chebyshev order n; q, y in tree (sorted) order.
int orc_fmm_matvec(void *p, int order, const Kernel *k, const double *q, double *y) {
  auto *h = static_cast<OctH *>(p);
  int rc = 0;
  h->with([&](auto &t) {
    constexpr int D = sizeof(t.bounds.lo) / sizeof(double);
    switch (order) {
      case 2: fmm_n<D, 2>(t, k, q, y); break;
      case 3: fmm_n<D, 3>(t, k, q, y); break;
      case 4: fmm_n<D, 4>(t, k, q, y); break;
      case 5: fmm_n<D, 5>(t, k, q, y); break;
      case 6: fmm_n<D, 6>(t, k, q, y); break;
      case 7: fmm_n<D, 7>(t, k, q, y); break;
      case 8: fmm_n<D, 8>(t, k, q, y); break;
      default: rc = -1;
    }
  });
  return rc;
}

// Chebyshev helpers for the closed-form checks
double orc_cheb_node(unsigned i, unsigned n) { return cheb_node(i, n); }
double orc_cheb_Sn(double x, unsigned i, unsigned n) { return cheb_Sn(x, i, n); }

// ---- Schwarz preconditioner, k nearest neighbours, least-squares rows ----

struct SchwarzH {
  Schwarz s;
};

// buffer: the reference's m_neighbourhood_buffer (default 1e5 * DBL_EPSILON)
void *orc_schwarz_create(void *p, const Kernel *k, double radius, double shift, double buffer) {
  auto *h = static_cast<AnyCells *>(p);
  auto *sh = new SchwarzH();
  const double r2 = radius > 0.0 ? radius * radius : -1.0;
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    schwarz_analyse<D>(get<D>(h), buffer, sh->s);
    schwarz_factorise<D>(get<D>(h), *k, r2, shift, sh->s);
    return 0;
  });
  return sh;
}

// with the coarse-grid domain of set_coarse_grid_n(n) (seeded sampler)
void *orc_schwarz_create_coarse(void *p, const Kernel *k, double radius, double shift,
                                double buffer, int n, uint64_t seed) {
  auto *h = static_cast<AnyCells *>(p);
  auto *sh = new SchwarzH();
  const double r2 = radius > 0.0 ? radius * radius : -1.0;
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    schwarz_analyse<D>(get<D>(h), buffer, sh->s);
    if (n > 0) schwarz_coarse<D>(get<D>(h), n, seed, sh->s);
    schwarz_factorise<D>(get<D>(h), *k, r2, shift, sh->s);
    return 0;
  });
  return sh;
}

// analyzePattern only (src/Preconditioners.h:1169-1316): the domain lists,
// no factorisation (orc_schwarz_apply must not be called on the handle)
void *orc_schwarz_analyse(void *p, double buffer) {
  auto *h = static_cast<AnyCells *>(p);
  auto *sh = new SchwarzH();
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    schwarz_analyse<D>(get<D>(h), buffer, sh->s);
    return 0;
  });
  return sh;
}

void orc_schwarz_destroy(void *p) { delete static_cast<SchwarzH *>(p); }

// counts = {n_domains, total indices, failed}
void orc_schwarz_info(void *p, int64_t *counts) {
  const Schwarz &s = static_cast<SchwarzH *>(p)->s;
  counts[0] = int64_t(s.nint.size());
  counts[1] = int64_t(s.idx.size());
  counts[2] = s.failed;
}

void orc_schwarz_domains(void *p, int64_t *off, int32_t *idx, int32_t *nint) {
  const Schwarz &s = static_cast<SchwarzH *>(p)->s;
  std::copy(s.off.begin(), s.off.end(), off);
  std::copy(s.idx.begin(), s.idx.end(), idx);
  std::copy(s.nint.begin(), s.nint.end(), nint);
}

void orc_schwarz_apply(void *p, const double *b, double *x) {
  schwarz_apply(static_cast<SchwarzH *>(p)->s, b, x);
}

void orc_knn(void *p, int k, int32_t *nbr, double *dist2) {
  auto *h = static_cast<AnyCells *>(p);
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    knn<D>(get<D>(h), k, nbr, dist2);
    return 0;
  });
}

void orc_lsq_rows(void *p, const Kernel *kern, double radius, double shift, int k,
                  const int32_t *nbr, double *w) {
  auto *h = static_cast<AnyCells *>(p);
  const double r2 = radius > 0.0 ? radius * radius : -1.0;
  dispatch(h->dim, [&](auto DD) {
    constexpr int D = decltype(DD)::value;
    lsq_rows<D>(get<D>(h), *kern, r2, shift, k, nbr, w);
    return 0;
  });
}

}  // extern "C"
