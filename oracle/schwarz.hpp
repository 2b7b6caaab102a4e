// CPU restatement of SchwartzPreconditioner<Eigen::LLT<MatrixXd>>
// (src/Preconditioners.h:1121-1537) over a CellListOrdered, plus the
// least-squares approximate-inverse rows and the k-nearest-neighbour pattern
// used by the learned preconditioner.  TEST INFRASTRUCTURE ONLY.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "cells_search.hpp"
#include "kernels.hpp"

namespace orc {

struct Schwarz {
  std::vector<int64_t> off{0};
  std::vector<int32_t> idx;
  std::vector<int32_t> nint;
  std::vector<std::vector<double>> rows;  // interior rows of each local inverse
  int64_t n = 0;
  int64_t failed = 0;
  bool coarse = false;  // the last domain is the coarse grid, solved last
};

// analyze_impl_block (src/Preconditioners.h:1169-1230): for every non-empty
// bucket, the buckets returned by get_buckets_near_point<-1>(middle, 1.0,
// scale 1/(half_side + buffer)) — a bucket is kept when the L-infinity norm
// of max(|t * (centre - middle)| - t * side / 2, 0) is <= 1 (the scaled box
// distance of lattice_iterator_within_distance,
// src/NeighbourSearchBase.h:1884-1892) — and every particle of those buckets
// inside the leaf bounds (src/CellListOrdered.h:446-451) is interior, the
// rest buffer.  Buckets are taken in lattice order here (the reference
// walks them by quadrant: the same set, another order).
template <int D> void schwarz_analyse(const Cells<D> &c, double buffer, Schwarz &s) {
  const Grid<D> &g = c.g;
  s.n = int64_t(c.pos.size() / D);
  s.off.assign(1, 0);
  s.idx.clear();
  s.nint.clear();
  const size_t nb = g.nbuckets();
  std::vector<int32_t> buf;
  for (size_t b = 0; b < nb; ++b) {
    if (c.end[b] <= c.begin[b]) continue;
    long iv[D];
    size_t r = b;
    for (int d = D - 1; d >= 0; --d) {
      iv[d] = long(r % g.size[d]);
      r /= g.size[d];
    }
    double bmin[D], bmax[D], middle[D], half[D], t[D], hb[D];
    for (int d = 0; d < D; ++d) {
      bmin[d] = double(iv[d]) * g.side[d] + g.lo[d];
      bmax[d] = double(iv[d] + 1) * g.side[d] + g.lo[d];
      middle[d] = 0.5 * (bmax[d] + bmin[d]);
      half[d] = 0.5 * (bmax[d] - bmin[d]);
      t[d] = 1.0 / (half[d] + buffer);
      hb[d] = 0.5 * (g.side[d] * t[d]);
    }
    const size_t first = s.idx.size();
    buf.clear();
    // candidate buckets: a window of +-3 around the leaf covers every bucket
    // the scaled unit ball can reach
    long lo[D], hi[D], cur[D];
    for (int d = 0; d < D; ++d) {
      lo[d] = std::max(iv[d] - 3, 0L);
      hi[d] = std::min(iv[d] + 3, long(g.size[d]) - 1);
      cur[d] = lo[d];
    }
    for (;;) {
      double dist = 0.0;
      for (int d = 0; d < D; ++d) {
        const double centre = (double(cur[d]) + 0.5) * g.side[d] + g.lo[d];
        const double dx = (centre - middle[d]) * t[d];
        dist = std::max(dist, std::max(std::fabs(dx) - hb[d], 0.0));
      }
      if (dist <= 1.0) {
        size_t bi = 0;
        for (int d = 0; d < D; ++d) bi = bi * g.size[d] + size_t(cur[d]);
        for (uint32_t p = c.begin[bi]; p < c.end[bi]; ++p) {
          bool inside = true;
          for (int d = 0; d < D; ++d) {
            const double v = c.pos[size_t(p) * D + d];
            if (v < bmin[d] || v >= bmax[d]) inside = false;
          }
          if (inside) s.idx.push_back(int32_t(p));
          else buf.push_back(int32_t(p));
        }
      }
      int d = D - 1;
      while (d >= 0 && ++cur[d] > hi[d]) {
        cur[d] = lo[d];
        --d;
      }
      if (d < 0) break;
    }
    const int32_t ni = int32_t(s.idx.size() - first);
    if (ni == 0) {
      s.idx.resize(first);
      continue;
    }
    s.idx.insert(s.idx.end(), buf.begin(), buf.end());
    s.nint.push_back(ni);
    s.off.push_back(int64_t(s.idx.size()));
  }
}

// coefficient of A = K + shift I: the operator's entry (kernel within the
// radius, FP64 squared distance summed in dimension order) plus the
// diagonal shift of the reference's self kernels
// (tests/rbf_interpolation.h:673-680)
template <int D>
double schwarz_coeff(const double *pos, int32_t i, int32_t j, const Kernel &k, double r2,
                     double shift) {
  double s = 0.0;
  for (int d = 0; d < D; ++d) {
    const double dx = pos[size_t(j) * D + d] - pos[size_t(i) * D + d];
    s = s + dx * dx;
  }
  double v = (r2 < 0.0 || s <= r2) ? phi(k, s) : 0.0;
  if (i == j) v = v + shift;
  return v;
}

// In-place textbook Cholesky (lower, column-major m x m); false on a
// non-positive pivot.
inline bool cholesky(std::vector<double> &A, int m) {
  for (int j = 0; j < m; ++j) {
    double d = A[size_t(j) * m + j];
    for (int q = 0; q < j; ++q) d -= A[size_t(q) * m + j] * A[size_t(q) * m + j];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    A[size_t(j) * m + j] = d;
    for (int i = j + 1; i < m; ++i) {
      double v = A[size_t(j) * m + i];
      for (int q = 0; q < j; ++q) v -= A[size_t(q) * m + i] * A[size_t(q) * m + j];
      A[size_t(j) * m + i] = v / d;
    }
  }
  return true;
}

// factorize (src/Preconditioners.h:1402-1467): local matrix via coeff,
// LLT; the interior rows of the local inverse are kept (the only rows
// _solve_impl writes back).
template <int D>
void schwarz_factorise(const Cells<D> &c, const Kernel &k, double r2, double shift, Schwarz &s) {
  const int64_t nd = int64_t(s.nint.size());
  s.rows.assign(size_t(nd), {});
  int64_t failed = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : failed)
  for (int64_t d = 0; d < nd; ++d) {
    const int32_t *id = &s.idx[size_t(s.off[d])];
    const int m = int(s.off[d + 1] - s.off[d]);
    const int ni = s.nint[size_t(d)];
    std::vector<double> A(size_t(m) * m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i)
        A[size_t(j) * m + i] = schwarz_coeff<D>(c.pos.data(), id[i], id[j], k, r2, shift);
    if (!cholesky(A, m)) {
      ++failed;
      continue;
    }
    // column a of A^-1 for a < ni: forward then backward substitution
    std::vector<double> &R = s.rows[size_t(d)];
    R.assign(size_t(ni) * m, 0.0);
    std::vector<double> y(m);
    for (int a = 0; a < ni; ++a) {
      for (int i = 0; i < m; ++i) {
        double v = i == a ? 1.0 : 0.0;
        for (int q = 0; q < i; ++q) v -= A[size_t(q) * m + i] * y[q];
        y[i] = v / A[size_t(i) * m + i];
      }
      for (int i = m - 1; i >= 0; --i) {
        double v = y[i];
        for (int q = i + 1; q < m; ++q) v -= A[size_t(i) * m + q] * y[q];
        y[i] = v / A[size_t(i) * m + i];
      }
      std::copy(y.begin(), y.end(), R.begin() + size_t(a) * m);
    }
  }
  s.failed = failed;
}

// _solve_impl (src/Preconditioners.h:1476-1522): x = b, then each domain's
// local solution overwrites its interior rows; the coarse-grid domain, the
// last one, after all others (the serial loop's order).
inline void schwarz_apply(const Schwarz &s, const double *b, double *x) {
  for (int64_t i = 0; i < s.n; ++i) x[i] = b[i];
  const int64_t nd = int64_t(s.nint.size());
  auto solve = [&](int64_t d) {
    if (s.rows[size_t(d)].empty()) return;
    const int32_t *id = &s.idx[size_t(s.off[d])];
    const int m = int(s.off[d + 1] - s.off[d]);
    for (int a = 0; a < s.nint[size_t(d)]; ++a) {
      const double *Ra = &s.rows[size_t(d)][size_t(a) * m];
      double v = 0.0;
      for (int j = 0; j < m; ++j) v += Ra[j] * b[id[j]];
      x[id[a]] = v;
    }
  };
  const int64_t nmain = s.coarse ? nd - 1 : nd;
#pragma omp parallel for schedule(dynamic)
  for (int64_t d = 0; d < nmain; ++d) solve(d);
  if (s.coarse) solve(nd - 1);
}

// The coarse-grid domain (SchwartzPreconditioner::set_coarse_grid_n,
// src/Preconditioners.h:1230-1283) appended to the analysed domains: the
// lattice points lo + i * ((hi - lo) / n - DBL_EPSILON), i in [0, n]^D,
// counted per bucket (bucket index clamped to the grid), then from each
// bucket with c points the first min(c, m) of its m particles after a
// shuffle.  The reference's random_shuffle is unseeded and its buckets come
// in unordered_map order; the sampler pinned here walks buckets in
// ascending index and shuffles local indices 0..m-1 forward (swap i with
// r mod (i+1), i = 1..m-1), r from SplitMix64 (state += golden gamma,
// two xor-shift-multiply rounds) seeded with seed ^ bucket * 0xD1B54A32D192ED03.
template <int D> void schwarz_coarse(const Cells<D> &c, int n, uint64_t seed, Schwarz &s) {
  const Grid<D> &g = c.g;
  const size_t nb = g.nbuckets();
  std::vector<long> hits(nb, 0);
  double step[D];
  for (int d = 0; d < D; ++d) step[d] = (g.hi[d] - g.lo[d]) / n - 2.220446049250313e-16;
  int it[D] = {};
  for (;;) {
    int v[D];
    for (int d = 0; d < D; ++d) {
      const double p = it[d] * step[d] + g.lo[d];
      long q = long(std::floor((p - g.lo[d]) * g.inv[d]));
      if (q < 0) q = 0;
      if (q > long(g.size[d]) - 1) q = long(g.size[d]) - 1;
      v[d] = int(q);
    }
    ++hits[collapse<D>(g, v)];
    int d = D - 1;
    while (d >= 0 && ++it[d] > n) it[d--] = 0;
    if (d < 0) break;
  }
  std::vector<int32_t> dom;
  for (size_t b = 0; b < nb; ++b) {
    const long m = long(c.end[b]) - long(c.begin[b]);
    if (hits[b] == 0 || m <= 0) continue;
    std::vector<int32_t> perm(static_cast<size_t>(m));
    std::iota(perm.begin(), perm.end(), 0);
    uint64_t state = seed ^ (uint64_t(b) * 0xD1B54A32D192ED03ull);
    for (long i = 1; i < m; ++i) {
      uint64_t z = (state += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      std::swap(perm[size_t(i)], perm[size_t(z % uint64_t(i + 1))]);
    }
    for (long i = 0; i < std::min(hits[b], m); ++i)
      dom.push_back(int32_t(c.begin[b]) + perm[size_t(i)]);
  }
  s.idx.insert(s.idx.end(), dom.begin(), dom.end());
  s.nint.push_back(int32_t(dom.size()));
  s.off.push_back(int64_t(s.idx.size()));
  s.coarse = true;
}

// k nearest particles of every particle: itself first, then the k-1
// nearest others by FP64 squared distance, ties by index; -1 where fewer
// exist.  Brute force over a block of buckets grown until the farthest kept
// distance lies inside it.
template <int D> void knn(const Cells<D> &c, int k, int32_t *nbr, double *dist2) {
  const int64_t n = int64_t(c.pos.size() / D);
  const Grid<D> &g = c.g;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const double *pi = &c.pos[size_t(i) * D];
    std::vector<std::pair<double, int32_t>> cand;
    for (long w = 1;; ++w) {
      cand.clear();
      long iv[D], lo[D], hi[D], cur[D];
      bool whole = true;
      for (int d = 0; d < D; ++d) {
        iv[d] = std::min(long(std::floor((pi[d] - g.lo[d]) * g.inv[d])), long(g.size[d]) - 1);
        lo[d] = std::max(iv[d] - w, 0L);
        hi[d] = std::min(iv[d] + w, long(g.size[d]) - 1);
        if (lo[d] > 0 || hi[d] < long(g.size[d]) - 1) whole = false;
        cur[d] = lo[d];
      }
      for (;;) {
        size_t bi = 0;
        for (int d = 0; d < D; ++d) bi = bi * g.size[d] + size_t(cur[d]);
        for (uint32_t p = c.begin[bi]; p < c.end[bi]; ++p) {
          if (int64_t(p) == i) continue;
          double s = 0.0;
          for (int d = 0; d < D; ++d) {
            const double dx = c.pos[size_t(p) * D + d] - pi[d];
            s = s + dx * dx;
          }
          cand.push_back({s, int32_t(p)});
        }
        int d = D - 1;
        while (d >= 0 && ++cur[d] > hi[d]) {
          cur[d] = lo[d];
          --d;
        }
        if (d < 0) break;
      }
      if (int64_t(cand.size()) < k - 1 && !whole) continue;
      std::sort(cand.begin(), cand.end());
      if (whole || k == 1) break;
      // the farthest kept distance must not reach outside the searched block
      const double rk = std::sqrt(cand[size_t(k - 2)].first);
      bool inside = true;
      for (int d = 0; d < D; ++d) {
        const double blo = double(lo[d]) * g.side[d] + g.lo[d];
        const double bhi = double(hi[d] + 1) * g.side[d] + g.lo[d];
        if ((lo[d] > 0 && pi[d] - rk < blo) || (hi[d] < long(g.size[d]) - 1 && pi[d] + rk >= bhi))
          inside = false;
      }
      if (inside) break;
    }
    nbr[size_t(i) * k] = int32_t(i);
    if (dist2) dist2[size_t(i) * k] = 0.0;
    for (int q = 1; q < k; ++q) {
      const bool ok = q - 1 < int(cand.size());
      nbr[size_t(i) * k + q] = ok ? cand[size_t(q - 1)].second : -1;
      if (dist2) dist2[size_t(i) * k + q] = ok ? cand[size_t(q - 1)].first : -1.0;
    }
  }
}

// Least-squares approximate-inverse row of particle i on its pattern S
// (k indices): w = argmin || A[:, S] w - e_i ||_2, from the normal
// equations G w = A[S, i] with G = A[:, S]^T A[:, S], the sum running over
// every particle within reach of S (radius truncation) or all particles.
template <int D>
void lsq_rows(const Cells<D> &c, const Kernel &kern, double r2, double shift, int k,
              const int32_t *nbr, double *w) {
  const int64_t n = int64_t(c.pos.size() / D);
  const double *pos = c.pos.data();
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) {
    const int32_t *S = nbr + size_t(i) * k;
    std::vector<double> G(size_t(k) * k, 0.0), rhs(k, 0.0), col(k);
    for (int64_t p = 0; p < n; ++p) {
      bool any = false;
      for (int a = 0; a < k; ++a) {
        col[a] = schwarz_coeff<D>(pos, int32_t(p), S[a], kern, r2, shift);
        any |= col[a] != 0.0;
      }
      if (!any) continue;
      for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) G[size_t(a) * k + b] += col[a] * col[b];
      if (p == i)
        for (int a = 0; a < k; ++a) rhs[a] += col[a];
    }
    if (!cholesky(G, k)) {
      for (int a = 0; a < k; ++a) w[size_t(i) * k + a] = 0.0;
      continue;
    }
    std::vector<double> y(k);
    for (int a = 0; a < k; ++a) {
      double v = rhs[a];
      for (int q = 0; q < a; ++q) v -= G[size_t(q) * k + a] * y[q];
      y[a] = v / G[size_t(a) * k + a];
    }
    for (int a = k - 1; a >= 0; --a) {
      double v = y[a];
      for (int q = a + 1; q < k; ++q) v -= G[size_t(a) * k + q] * y[q];
      y[a] = v / G[size_t(a) * k + a];
    }
    for (int a = 0; a < k; ++a) w[size_t(i) * k + a] = y[a];
  }
}

}  // namespace orc
