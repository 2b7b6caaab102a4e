// Exercises include/aboria_hip.hpp the way a C++ caller of the reference API
// would: build a cell list, run euclidean_search over every particle and a
// sparse Gaussian operator, and check both against a brute-force host loop.
// Exit status 0 = all checks passed.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "aboria_hip.hpp"

namespace ah = aboria_hip;

int main() {
  try {
    ah::Context ctx(0);
    const int64_t n = 3000;
    const double r = 0.08;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> pos(n * 3);
    for (auto &v : pos) v = u(rng);
    ah::Context::Buffer<double> dpos(ctx, pos.size());
    dpos.upload(pos);
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    const bool per[3] = {false, false, false};
    ah::CellListOrdered<3> cells(ctx, lo, hi, per, 10.0);
    const int64_t na = cells.update_positions(dpos.p, n);
    if (na != n) return std::printf("alive %ld\n", long(na)), 1;
    // sorted positions, host copy
    std::vector<double> sp(n * 3);
    ah::check(abh_memcpy_d2h(ctx.get(), sp.data(), cells.positions(), sp.size() * 8), ctx.get());
    ah::SearchResult res = ah::euclidean_search(cells.get_query(), cells.positions(), n, r);
    std::vector<int64_t> rp = res.row_ptr.download();
    std::vector<int32_t> col = res.col.download();
    int64_t brute = 0;
    for (int64_t i = 0; i < n; ++i) {
      int64_t cnt = 0;
      for (int64_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
          const double dx = sp[j * 3 + d] - sp[i * 3 + d];
          s = s + dx * dx;
        }
        cnt += s <= r * r;
      }
      if (cnt != rp[i + 1] - rp[i]) return std::printf("row %ld count mismatch\n", long(i)), 1;
      brute += cnt;
    }
    // sparse operator: y += K x
    auto K = ah::create_sparse_operator(cells, cells, r, ah::Kernel::gaussian(0.04));
    std::vector<double> x(n), y(n, 0.0);
    for (auto &v : x) v = u(rng);
    ah::Context::Buffer<double> dx(ctx, n), dy(ctx, n);
    dx.upload(x);
    dy.upload(y);
    K->evaluate(dy.p, n, dx.p, n);
    std::vector<double> yg = dy.download();
    double err = 0.0, nrm = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
        const int64_t j = col[q];
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
          const double dd = sp[j * 3 + d] - sp[i * 3 + d];
          s = s + dd * dd;
        }
        acc += std::exp(-(s / (0.04 * 0.04))) * x[j];
      }
      err += (acc - yg[i]) * (acc - yg[i]);
      nrm += acc * acc;
    }
    const double rel = std::sqrt(err / nrm);
    std::printf("pairs %ld (brute force %ld), sparse matvec rel err %.3e\n", long(res.total),
                long(brute), rel);
    return (res.total == brute && rel < 1e-13) ? 0 : 1;
  } catch (const ah::Error &e) {
    std::printf("error: %s\n", e.what());
    return 2;
  }
}
