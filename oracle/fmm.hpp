// CPU oracle: HyperOctree build and the black-box Chebyshev FMM restated from
// the reference's CPU path.  TEST INFRASTRUCTURE ONLY.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <numeric>
#include <vector>

#include "kernels.hpp"

namespace orc {

constexpr double kPi = 3.141592653589793;  // boost pi<double>()
constexpr double kPi2 = kPi / 2;           // src/detail/Chebyshev.h:51-52

// chebyshev_node (src/detail/Chebyshev.h:101-104)
inline double cheb_node(unsigned i, unsigned n) {
  return std::cos((2.0 * i + 1.0) * kPi / (2.0 * n));
}

// chebyshev_Sn by Clenshaw for one coordinate (src/detail/Chebyshev.h:137-153)
inline double cheb_Sn(double x, unsigned i, unsigned n) {
  double b1 = 0.0, b2 = 0.0;
  const double invn = 1.0 / n;
  for (unsigned k = n - 1; k >= 1; --k) {
    const double bk = 2 * invn * std::cos(k * (2.0 * i + 1.0) * kPi2 * invn) + 2.0 * x * b1 - b2;
    b2 = b1;
    b1 = bk;
  }
  return invn + x * b1 - b2;
}

template <int D> struct OBox {
  double lo[D], hi[D];
};

// ChebyshevRnSingle (src/detail/Chebyshev.h:245-294): S_n of the position
// mapped to [-1,1] in every dimension; R_n(m) = prod_d Sn[m_d][d].
template <int D, int N> struct RnSingle {
  double sn[N][D];
  RnSingle(const double *p, const OBox<D> &b) {
    double shift[D];
    for (int d = 0; d < D; ++d) {
      const double span = b.hi[d] - b.lo[d];
      shift[d] = span > 0.0 ? (2 * p[d] - b.lo[d] - b.hi[d]) / span : 0.0;
    }
    for (int m = 0; m < N; ++m)
      for (int d = 0; d < D; ++d) sn[m][d] = cheb_Sn(shift[d], unsigned(m), unsigned(N));
  }
  // m is the lattice index of a node, last dimension fastest
  double operator()(int m) const {
    int idx[D];
    for (int d = D - 1; d >= 0; --d) {
      idx[d] = m % N;
      m /= N;
    }
    double r = 1.0;
    for (int d = 0; d < D; ++d) r *= sn[idx[d]][d];
    return r;
  }
};

// point_to_tag (src/detail/SpatialUtil.h:255-286)
template <int D> inline int point_to_tag(const double *p, OBox<D> box, int max_level) {
  int result = 0;
  for (int level = 1; level <= max_level; ++level) {
    int hi_half[D];
    double mid[D];
    for (int i = 0; i < D; ++i) {
      mid[i] = 0.5f * (box.lo[i] + box.hi[i]);
      hi_half[i] = (p[i] < mid[i]) ? 0 : 1;
      result |= hi_half[i];
      result <<= 1;
    }
    for (int i = 0; i < D; ++i) {
      box.lo[i] = hi_half[i] ? mid[i] : box.lo[i];
      box.hi[i] = hi_half[i] ? box.hi[i] : mid[i];
    }
  }
  result >>= 1;
  return result;
}

enum : int { kEmpty = int(0xffffffff) };
inline bool is_leaf_id(int id) { return id < 0; }  // leaves and empties
inline bool is_empty_id(int id) { return id == kEmpty; }
inline int leaf_offset(int id) { return int(0x80000000u ^ uint32_t(id)); }

// HyperOctree (src/OctTree.h:102-475): tags, stable sort, then level by level
// the 2^D children of every active node, counted by lower/upper bound and
// classified EMPTY / LEAF (count <= n_leaf or last level) / NODE.
template <int D> struct Octree {
  OBox<D> bounds;
  int max_level = 32 / D - 2;
  std::vector<int> tags;       // sorted
  std::vector<int32_t> order;  // sorted -> input index
  std::vector<int> nodes;      // child codes; nodes[0..2^D) are the root's children
  std::vector<std::pair<int, int>> leaves;
  std::vector<double> pos;     // sorted positions
  int levels = 0;

  void build(const double *p, size_t n, double n_leaf) {
    constexpr int nchild = 1 << D;
    tags.resize(n);
    std::vector<int> t0(n);
    for (size_t i = 0; i < n; ++i) t0[i] = point_to_tag<D>(p + i * D, bounds, max_level);
    order.resize(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return t0[a] < t0[b]; });
    for (size_t i = 0; i < n; ++i) tags[i] = t0[order[i]];
    pos.resize(n * D);
    for (size_t i = 0; i < n; ++i)
      for (int d = 0; d < D; ++d) pos[i * D + d] = p[size_t(order[i]) * D + d];
    nodes.clear();
    leaves.clear();
    std::vector<int> active(1, 0);
    for (int level = 1; !active.empty(); ++level) {
      std::vector<int> children(nchild * active.size());
      for (size_t c = 0; c < children.size(); ++c) {
        const int tag = active[c / nchild];
        const int shift = (max_level - level) * D;
        children[c] = tag | (int(c & (nchild - 1)) << shift);
      }
      const int length = (1 << (max_level - level) * D) - 1;
      const int children_begin = int(nodes.size());
      nodes.resize(nodes.size() + children.size());
      std::vector<int> next;
      int node_idx = 0;
      for (size_t c = 0; c < children.size(); ++c) {
        const int lb = int(std::lower_bound(tags.begin(), tags.end(), children[c]) - tags.begin());
        const int ub = int(std::upper_bound(tags.begin(), tags.end(), children[c] + length) - tags.begin());
        const int count = ub - lb;
        if (count == 0) {
          nodes[children_begin + c] = kEmpty;
        } else if (level == max_level || count <= n_leaf) {
          nodes[children_begin + c] = int(0x80000000u | uint32_t(leaves.size()));
          leaves.push_back({lb, ub});
        } else {
          nodes[children_begin + c] = int(nodes.size()) + nchild * node_idx;  // after this level
          ++node_idx;
          next.push_back(children[c]);
        }
      }
      active.swap(next);
      levels = level;
    }
  }

  // child box c of box b (octree_child_iterator::get_bounds, src/OctTree.h:633-644)
  static OBox<D> child_box(const OBox<D> &b, int c) {
    OBox<D> r = b;
    for (int i = 0; i < D; ++i) {
      const bool high = c & (1 << (D - 1 - i));
      if (high) r.lo[i] = 0.5 * (r.hi[i] + r.lo[i]);
      else r.hi[i] = 0.5 * (r.hi[i] + r.lo[i]);
    }
    return r;
  }
};

// theta_condition (src/detail/FastMultipoleMethod.h:1249-1279): true when the
// source box is too close for M2L (strongly connected).
template <int D> struct Theta {
  const OBox<D> &t;
  double max_diam;
  explicit Theta(const OBox<D> &target) : t(target), max_diam(std::numeric_limits<double>::min()) {
    for (int i = 0; i < D; ++i)
      if (t.hi[i] - t.lo[i] > max_diam) max_diam = t.hi[i] - t.lo[i];
  }
  bool check(const OBox<D> &s) const {
    double dist[D];
    double md = max_diam;
    for (int i = 0; i < D; ++i) {
      dist[i] = 0.0;
      if (t.hi[i] < s.lo[i]) dist[i] = s.lo[i] - t.hi[i];
      else if (s.hi[i] < t.lo[i]) dist[i] = t.lo[i] - s.hi[i];
      if (s.hi[i] - s.lo[i] > md) md = s.hi[i] - s.lo[i];
    }
    double sq = 0.0;
    for (int i = 0; i < D; ++i) sq += dist[i] * dist[i];
    return md * md > 1.0 * sq;
  }
};

// The dual traversal's interaction lists (H2Matrix::generate_matrices,
// src/H2Matrix.h:998-1090; the same walk as the FMM's
// calculate_dive_M2L_and_L2L, src/FastMultipoleMethod.h:442-548): per node
// its weak (M2L) source nodes; per non-empty leaf its P2P source leaves --
// the strong connections in order, each expanded depth first (children in
// order) to its leaves, empty leaves dropped.
template <int D> struct Lists {
  std::vector<std::vector<int>> weak;  // per node
  std::vector<std::vector<int>> p2p;   // per leaf index
};

template <int D> void interaction_lists(const Octree<D> &t, Lists<D> &L) {
  constexpr int nchild = 1 << D;
  L.weak.assign(t.nodes.size(), {});
  L.p2p.assign(t.leaves.size(), {});
  struct R {
    int k;
    OBox<D> box;
  };
  std::function<void(int, std::vector<int> &)> expand = [&](int k, std::vector<int> &dst) {
    const int id = t.nodes[k];
    if (is_leaf_id(id)) {
      if (!is_empty_id(id)) dst.push_back(leaf_offset(id));
      return;
    }
    for (int c = 0; c < nchild; ++c) expand(id + c, dst);
  };
  std::function<void(const R &, const std::vector<R> &)> gen = [&](const R &tg,
                                                                   const std::vector<R> &ps) {
    Theta<D> theta(tg.box);
    std::vector<R> strong;
    auto consider = [&](const R &s) {
      if (theta.check(s.box)) strong.push_back(s);
      else L.weak[tg.k].push_back(s.k);
    };
    if (ps.empty()) {
      for (int c = 0; c < nchild; ++c) consider(R{c, Octree<D>::child_box(t.bounds, c)});
    } else {
      for (const R &src : ps) {
        const int id = t.nodes[src.k];
        if (is_leaf_id(id)) strong.push_back(src);
        else
          for (int c = 0; c < nchild; ++c) consider(R{id + c, Octree<D>::child_box(src.box, c)});
      }
    }
    const int tid = t.nodes[tg.k];
    if (!is_leaf_id(tid)) {
      for (int c = 0; c < nchild; ++c) gen(R{tid + c, Octree<D>::child_box(tg.box, c)}, strong);
      return;
    }
    if (is_empty_id(tid)) return;
    std::vector<int> &dst = L.p2p[leaf_offset(tid)];
    for (const R &s : strong) expand(s.k, dst);
  };
  const std::vector<R> none;
  for (int c = 0; c < nchild; ++c) gen(R{c, Octree<D>::child_box(t.bounds, c)}, none);
}

// FastMultipoleMethod::matrix_vector_multiply with BlackBoxExpansions
// (src/FastMultipoleMethod.h:110-548, src/detail/FastMultipoleMethod.h:96-368,
// 720-1101), serial, one right-hand side, rows == cols (same tree).
template <int D, int N> struct FMM {
  static constexpr int ncheb = [] {
    int r = 1;
    for (int d = 0; d < D; ++d) r *= N;
    return r;
  }();
  const Octree<D> &tree;
  Kernel kern;
  std::vector<std::vector<double>> W, g;
  double unit[ncheb][D];  // Chebyshev nodes on [-1,1]^D in lattice order

  FMM(const Octree<D> &t, const Kernel &k) : tree(t), kern(k) {
    W.assign(t.nodes.size(), std::vector<double>(ncheb, 0.0));
    g.assign(t.nodes.size(), std::vector<double>(ncheb, 0.0));
    for (int m = 0; m < ncheb; ++m) {
      int rem = m;
      for (int d = D - 1; d >= 0; --d) {
        unit[m][d] = cheb_node(unsigned(rem % N), unsigned(N));
        rem /= N;
      }
    }
  }

  double K(const double *a, const double *b) const {
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
      const double dx = b[d] - a[d];
      s = s + dx * dx;
    }
    return phi(kern, s);
  }

  void map(const double *u, const OBox<D> &b, double *p) const {
    for (int d = 0; d < D; ++d) p[d] = 0.5 * (u[d] + 1) * (b.hi[d] - b.lo[d]) + b.lo[d];
  }

  void up(int k, const OBox<D> &box, const double *q) {
    std::vector<double> &w = W[k];
    std::fill(w.begin(), w.end(), 0.0);
    const int id = tree.nodes[k];
    if (is_leaf_id(id)) {
      if (is_empty_id(id)) return;
      const auto lr = tree.leaves[leaf_offset(id)];
      for (int i = lr.first; i < lr.second; ++i) {  // P2M
        RnSingle<D, N> rn(&tree.pos[size_t(i) * D], box);
        for (int j = 0; j < ncheb; ++j) w[j] += rn(j) * q[i];
      }
      return;
    }
    for (int c = 0; c < (1 << D); ++c) {
      const OBox<D> cb = Octree<D>::child_box(box, c);
      up(id + c, cb, q);
      const std::vector<double> &wc = W[id + c];
      for (int j = 0; j < ncheb; ++j) {  // M2M
        double pj[D];
        map(unit[j], cb, pj);
        RnSingle<D, N> rn(pj, box);
        for (int i = 0; i < ncheb; ++i) w[i] += rn(i) * wc[j];
      }
    }
  }

  void m2l(std::vector<double> &acc, const OBox<D> &tb, const OBox<D> &sb, const std::vector<double> &src) const {
    for (int i = 0; i < ncheb; ++i) {
      double pi[D];
      map(unit[i], tb, pi);
      for (int j = 0; j < ncheb; ++j) {
        double pj[D];
        map(unit[j], sb, pj);
        acc[i] += K(pi, pj) * src[j];
      }
    }
  }

  struct Ref {
    int k;
    OBox<D> box;
  };

  void leaves_of(const Ref &r, std::vector<Ref> &out) const {
    const int id = tree.nodes[r.k];
    if (is_leaf_id(id)) {
      out.push_back(r);
      return;
    }
    for (int c = 0; c < (1 << D); ++c) leaves_of(Ref{id + c, Octree<D>::child_box(r.box, c)}, out);
  }

  void p2p(const Ref &t, const Ref &s, const double *q, double *y) const {
    const int ti = tree.nodes[t.k], si = tree.nodes[s.k];
    if (is_empty_id(ti) || is_empty_id(si)) return;
    const auto tr = tree.leaves[leaf_offset(ti)];
    const auto sr = tree.leaves[leaf_offset(si)];
    for (int i = tr.first; i < tr.second; ++i)
      for (int j = sr.first; j < sr.second; ++j)
        y[i] += K(&tree.pos[size_t(i) * D], &tree.pos[size_t(j) * D]) * q[j];
  }

  void down(const Ref &t, const std::vector<Ref> &conn_parent, const std::vector<double> &g_parent,
            const OBox<D> &box_parent, const double *q, double *y) {
    std::vector<double> &gt = g[t.k];
    std::fill(gt.begin(), gt.end(), 0.0);
    std::vector<Ref> conn;
    Theta<D> theta(t.box);
    if (conn_parent.empty()) {
      for (int c = 0; c < (1 << D); ++c) {
        const Ref s{c, Octree<D>::child_box(tree.bounds, c)};
        if (theta.check(s.box)) conn.push_back(s);
        else m2l(gt, t.box, s.box, W[s.k]);
      }
    } else {
      for (int i = 0; i < ncheb; ++i) {  // L2L
        double pi[D];
        map(unit[i], t.box, pi);
        RnSingle<D, N> rn(pi, box_parent);
        for (int j = 0; j < ncheb; ++j) gt[i] += rn(j) * g_parent[j];
      }
      for (const Ref &src : conn_parent) {
        const int sid = tree.nodes[src.k];
        if (is_leaf_id(sid)) {
          conn.push_back(src);
        } else {
          for (int c = 0; c < (1 << D); ++c) {
            const Ref s{sid + c, Octree<D>::child_box(src.box, c)};
            if (theta.check(s.box)) conn.push_back(s);
            else m2l(gt, t.box, s.box, W[s.k]);
          }
        }
      }
    }
    const int tid = tree.nodes[t.k];
    if (!is_leaf_id(tid)) {
      for (int c = 0; c < (1 << D); ++c)
        down(Ref{tid + c, Octree<D>::child_box(t.box, c)}, conn, gt, t.box, q, y);
      return;
    }
    if (is_empty_id(tid)) return;
    const auto tr = tree.leaves[leaf_offset(tid)];
    for (int i = tr.first; i < tr.second; ++i) {  // L2P
      RnSingle<D, N> rn(&tree.pos[size_t(i) * D], t.box);
      double sum = 0.0;
      for (int j = 0; j < ncheb; ++j) sum += rn(j) * gt[j];
      y[i] += sum;
    }
    for (const Ref &s : conn) {  // P2P
      std::vector<Ref> ls;
      leaves_of(s, ls);
      for (const Ref &l : ls) p2p(t, l, q, y);
    }
  }

  // y += K q, q and y in tree (sorted) order
  void matvec(const double *q, double *y) {
    for (int c = 0; c < (1 << D); ++c) up(c, Octree<D>::child_box(tree.bounds, c), q);
    std::vector<Ref> none;
    std::vector<double> g0(ncheb, 0.0);
    for (int c = 0; c < (1 << D); ++c)
      down(Ref{c, Octree<D>::child_box(tree.bounds, c)}, none, g0, OBox<D>(), q, y);
  }
};

}  // namespace orc
