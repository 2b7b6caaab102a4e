// aboria_hip.hpp -- header-only C++17 layer over the C ABI of
// libaboria_hip.so, with the names of Aboria's user-facing hot-path API:
//
//   CellListOrdered<D>        init_neighbour_search / update_positions /
//                             get_query (src/Particles.h:445-540,
//                             src/CellListOrdered.h:132-263)
//   euclidean_search(...)     bulk form of src/Search.h:839-845
//   create_sparse_operator    src/Operators.h:472-516
//   create_dense_operator     src/Operators.h:309-317
//   create_fmm_operator<N>    src/Operators.h:396-407
//   SchwartzPreconditioner    src/Preconditioners.h:1121-1537
//   gmres(...)                the Eigen::GMRES solve of the reference's tests
//
// Operators implement the reference's evaluate semantics, lhs += K rhs
// (src/detail/Operators.h:218-231), on device vectors.  Errors surface as
// aboria_hip::Error (status + message); nothing is retried or emulated on
// the host.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "aboria_hip.h"

namespace aboria_hip {

class Error : public std::runtime_error {
 public:
  Error(abh_status s, const std::string &m)
      : std::runtime_error(std::string(abh_status_string(s)) + ": " + m), status(s) {}
  abh_status status;
};

inline void check(abh_status s, const abh_context *ctx = nullptr) {
  if (s != ABH_OK) throw Error(s, ctx ? abh_last_error(ctx) : "");
}

// Device + stream.  stream == nullptr: the HIP null stream.
class Context {
 public:
  explicit Context(int device = 0, void *stream = nullptr) {
    check(abh_context_create(device, stream, &h_));
  }
  ~Context() { abh_context_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  abh_context *get() const { return h_; }
  void sync() const { check(abh_context_sync(h_), h_); }

  // owning device buffer of T
  template <class T> struct Buffer {
    Context *ctx = nullptr;
    T *p = nullptr;
    size_t n = 0;
    Buffer() = default;
    Buffer(Context &c, size_t count) : ctx(&c), n(count) {
      void *q = nullptr;
      if (count) check(abh_malloc(c.get(), &q, count * sizeof(T)), c.get());
      p = static_cast<T *>(q);
    }
    Buffer(Buffer &&o) noexcept : ctx(o.ctx), p(o.p), n(o.n) { o.p = nullptr; }
    Buffer &operator=(Buffer &&o) noexcept {
      std::swap(ctx, o.ctx);
      std::swap(p, o.p);
      std::swap(n, o.n);
      return *this;
    }
    ~Buffer() {
      if (p) abh_free(ctx->get(), p);
    }
    void upload(const std::vector<T> &h) {
      check(abh_memcpy_h2d(ctx->get(), p, h.data(), h.size() * sizeof(T)), ctx->get());
    }
    std::vector<T> download() const {
      std::vector<T> h(n);
      if (n) check(abh_memcpy_d2h(ctx->get(), h.data(), p, n * sizeof(T)), ctx->get());
      return h;
    }
  };

 private:
  abh_context *h_ = nullptr;
};

struct Kernel {
  abh_kernel k{};
  static Kernel make(abh_kernel_type t, double p) {
    Kernel r;
    r.k.type = t;
    r.k.param = p;
    return r;
  }
  static Kernel gaussian(double scale) { return make(ABH_KERNEL_GAUSSIAN, scale); }
  static Kernel wendland_c2(double support) { return make(ABH_KERNEL_WENDLAND_C2, support); }
  static Kernel multiquadric(double c) { return make(ABH_KERNEL_MULTIQUADRIC, c); }
  static Kernel smooth_inv_r(double eps) { return make(ABH_KERNEL_SMOOTH_INV_R, eps); }
  static Kernel constant() { return make(ABH_KERNEL_CONSTANT, 0.0); }
};

// The neighbour-search structure of a particle set; doubles as its own
// query object (get_query()).
template <unsigned D> class CellListOrdered {
 public:
  CellListOrdered(Context &ctx, const double (&low)[D], const double (&high)[D],
                  const bool (&periodic)[D], double n_particles_in_leaf = 10.0)
      : ctx_(&ctx) {
    uint8_t per[D];
    for (unsigned d = 0; d < D; ++d) per[d] = periodic[d] ? 1 : 0;
    check(abh_cells_create(ctx.get(), int(D), low, high, per, n_particles_in_leaf, &h_),
          ctx.get());
  }
  ~CellListOrdered() { abh_cells_destroy(h_); }
  CellListOrdered(const CellListOrdered &) = delete;
  CellListOrdered &operator=(const CellListOrdered &) = delete;

  // pos: device, n x D (wrapped in place); alive: device or nullptr.
  int64_t update_positions(double *pos, int64_t n, uint8_t *alive = nullptr) {
    int64_t na = 0;
    check(abh_cells_update(h_, n, pos, alive, &na), ctx_->get());
    n_alive_ = na;
    return na;
  }
  const CellListOrdered &get_query() const { return *this; }
  int64_t size() const { return n_alive_; }
  const double *positions() const { return abh_cells_positions(h_); }
  const int32_t *order() const { return abh_cells_order(h_); }
  const uint32_t *bucket_begin() const { return abh_cells_bucket_begin(h_); }
  const uint32_t *bucket_end() const { return abh_cells_bucket_end(h_); }
  abh_cells_info info() const {
    abh_cells_info i{};
    check(abh_cells_get_info(h_, &i), ctx_->get());
    return i;
  }
  abh_cells *get() const { return h_; }
  Context &context() const { return *ctx_; }

 private:
  Context *ctx_;
  abh_cells *h_ = nullptr;
  int64_t n_alive_ = 0;
};

// Neighbour lists in CSR form (device): row i is col[row_ptr[i] ..
// row_ptr[i+1]), container indices of the cell list after reorder.
struct SearchResult {
  Context::Buffer<int64_t> row_ptr;
  Context::Buffer<int32_t> col;
  int64_t total = 0;
};

template <unsigned D>
SearchResult euclidean_search(const CellListOrdered<D> &query, const double *centres,
                              int64_t nq, double radius) {
  Context &ctx = query.context();
  abh_search *s = nullptr;
  check(abh_search_create(ctx.get(), query.get(), nq, centres, radius, &s), ctx.get());
  std::unique_ptr<abh_search, abh_status (*)(abh_search *)> guard(s, abh_search_destroy);
  SearchResult r;
  check(abh_search_total(s, &r.total), ctx.get());
  r.row_ptr = Context::Buffer<int64_t>(ctx, size_t(nq + 1));
  r.col = Context::Buffer<int32_t>(ctx, size_t(r.total));
  check(abh_memcpy_d2d(ctx.get(), r.row_ptr.p, abh_search_row_ptr(s), size_t(nq + 1) * 8),
        ctx.get());
  check(abh_search_fill(s, r.col.p, r.total), ctx.get());
  return r;
}

// lhs += K rhs with K_ij = f(|x_j - y_i|) over the neighbours within radius.
class SparseOperator {
 public:
  template <unsigned D>
  SparseOperator(const CellListOrdered<D> &cols, int64_t nrows, const double *row_pos,
                 double radius, const Kernel &k)
      : ctx_(&cols.context()), nrows_(nrows), ncols_(cols.size()) {
    check(abh_sparse_op_create(ctx_->get(), cols.get(), nrows, row_pos, radius, &k.k, nullptr,
                               nullptr, &h_),
          ctx_->get());
  }
  ~SparseOperator() { abh_sparse_op_destroy(h_); }
  SparseOperator(const SparseOperator &) = delete;
  SparseOperator &operator=(const SparseOperator &) = delete;
  int64_t rows() const { return nrows_; }
  int64_t cols() const { return ncols_; }
  // column-major lhs (rows x nrhs, ld) += K rhs (cols x nrhs, ld)
  void evaluate(double *lhs, int64_t ldl, const double *rhs, int64_t ldr, int nrhs = 1,
                abh_precision p = ABH_FP64) const {
    check(abh_sparse_op_apply(h_, nrhs, rhs, ldr, lhs, ldl, p), ctx_->get());
  }
  abh_gmres_stats gmres(const double *b, double *x, double diag_shift = 0.0,
                        const abh_schwarz *prec = nullptr, int restart = 50,
                        int max_iters = 1000, double tol = 1e-8) const {
    abh_gmres_options o{restart, max_iters, tol, ABH_FP64, 0, 0};
    abh_gmres_stats st{};
    check(abh_sparse_op_gmres(h_, diag_shift, prec, nullptr, b, x, &o, &st), ctx_->get());
    return st;
  }
  abh_sparse_op *get() const { return h_; }

 private:
  Context *ctx_;
  int64_t nrows_, ncols_;
  abh_sparse_op *h_ = nullptr;
};

// create_sparse_operator(rows, cols, radius, f) with the rows at the column
// set's own particles (the square operators of the reference's solvers).
template <unsigned D>
std::unique_ptr<SparseOperator> create_sparse_operator(const CellListOrdered<D> &rows,
                                                       const CellListOrdered<D> &cols,
                                                       double radius, const Kernel &k) {
  return std::make_unique<SparseOperator>(cols, rows.size(), rows.positions(), radius, k);
}

// create_dense_operator(rows, cols, f): lhs += K rhs over all pairs.
class DenseOperator {
 public:
  DenseOperator(Context &ctx, int dim, int64_t nrows, const double *row_pos, int64_t ncols,
                const double *col_pos, const Kernel &k)
      : ctx_(&ctx), dim_(dim), nrows_(nrows), ncols_(ncols), rp_(row_pos), cp_(col_pos), k_(k) {}
  void evaluate(double *lhs, int64_t ldl, const double *rhs, int64_t ldr, int nrhs = 1,
                abh_precision p = ABH_FP64) const {
    check(abh_dense_apply(ctx_->get(), dim_, nrows_, rp_, ncols_, cp_, &k_.k, nrhs, rhs, ldr,
                          lhs, ldl, p),
          ctx_->get());
  }

 private:
  Context *ctx_;
  int dim_;
  int64_t nrows_, ncols_;
  const double *rp_, *cp_;
  Kernel k_;
};

template <unsigned D>
DenseOperator create_dense_operator(const CellListOrdered<D> &rows,
                                    const CellListOrdered<D> &cols, const Kernel &k) {
  return DenseOperator(rows.context(), int(D), rows.size(), rows.positions(), cols.size(),
                       cols.positions(), k);
}

// create_fmm_operator<N>(particles, ...): black-box Chebyshev FMM, 3D.
class FmmOperator {
 public:
  FmmOperator(Context &ctx, const double (&low)[3], const double (&high)[3], int64_t n,
              const double *pos, double n_leaf, int order, const Kernel &k)
      : ctx_(&ctx) {
    check(abh_fmm_create(ctx.get(), 3, low, high, n, pos, n_leaf, order, &k.k, &h_), ctx.get());
  }
  ~FmmOperator() { abh_fmm_destroy(h_); }
  FmmOperator(const FmmOperator &) = delete;
  FmmOperator &operator=(const FmmOperator &) = delete;
  void evaluate(double *lhs, int64_t ldl, const double *rhs, int64_t ldr, int nrhs = 1,
                abh_precision p = ABH_FP64) const {
    check(abh_fmm_apply(h_, nrhs, rhs, ldr, lhs, ldl, p), ctx_->get());
  }
  abh_fmm_info info() const {
    abh_fmm_info i{};
    check(abh_fmm_get_info(h_, &i), ctx_->get());
    return i;
  }

 private:
  Context *ctx_;
  abh_fmm *h_ = nullptr;
};

template <int N>
std::unique_ptr<FmmOperator> create_fmm_operator(Context &ctx, const double (&low)[3],
                                                 const double (&high)[3], int64_t n,
                                                 const double *pos, double n_leaf,
                                                 const Kernel &k) {
  return std::make_unique<FmmOperator>(ctx, low, high, n, pos, n_leaf, N, k);
}

// SchwartzPreconditioner<LLT>: compute() = analyzePattern + factorize,
// solve() = _solve_impl.
class SchwartzPreconditioner {
 public:
  template <unsigned D>
  SchwartzPreconditioner(const CellListOrdered<D> &particles, const Kernel &k, double radius,
                         double diag_shift = 0.0)
      : ctx_(&particles.context()) {
    check(abh_schwarz_create(ctx_->get(), particles.get(), &k.k, radius, diag_shift, &h_),
          ctx_->get());
  }
  ~SchwartzPreconditioner() { abh_schwarz_destroy(h_); }
  SchwartzPreconditioner(const SchwartzPreconditioner &) = delete;
  SchwartzPreconditioner &operator=(const SchwartzPreconditioner &) = delete;
  void solve(const double *b, double *x) const { check(abh_schwarz_apply(h_, b, x), ctx_->get()); }
  abh_schwarz_info info() const {
    abh_schwarz_info i{};
    check(abh_schwarz_get_info(h_, &i), ctx_->get());
    return i;
  }
  const abh_schwarz *get() const { return h_; }

 private:
  Context *ctx_;
  abh_schwarz *h_ = nullptr;
};

}  // namespace aboria_hip
