// CPU AddressSanitizer / UBSan run of the oracle (oracle/oracle_capi.cpp,
// compiled together with this program): every entry point on small seeded
// inputs in 1, 2 and 3 dimensions, periodic and not, with non-finite and
// out-of-domain particles.  Any out-of-bounds access or undefined operation
// aborts; this program also checks that count and fill agree.
//
// usage: asan_oracle   (prints "ok" on success)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
struct KernelC {
  int32_t type, reserved;
  double param, w0, w_row, w_col;
};
void *orc_cells_create(int, const double *, const double *, const uint8_t *, double);
void orc_cells_destroy(void *);
int64_t orc_cells_update(void *, int64_t, double *, uint8_t *);
int64_t orc_cells_grid(void *, uint32_t *, double *);
void orc_cells_get(void *, int32_t *, uint32_t *, uint32_t *, uint32_t *, double *);
int64_t orc_search_count(void *, int64_t, const double *, double, int64_t *, int);
void orc_search_fill(void *, int64_t, const double *, double, const int64_t *, int32_t *, int);
void orc_sparse_matvec(void *, int64_t, const double *, double, const KernelC *, const double *,
                       const double *, int, const double *, int64_t, double *, int64_t, int);
void orc_dense_matvec(int, int64_t, const double *, int64_t, const double *, const KernelC *, int,
                      const double *, int64_t, double *, int64_t, int);
void *orc_octree_create(int, const double *, const double *, int64_t, const double *, double);
void orc_octree_destroy(void *);
void orc_octree_info(void *, int64_t *);
int orc_octree_lists(void *, int64_t *, int32_t *, int32_t *, int32_t *, int32_t *);
int orc_fmm_matvec(void *, int, const KernelC *, const double *, double *);
void *orc_schwarz_create(void *, const KernelC *, double, double, double);
void orc_schwarz_destroy(void *);
void orc_schwarz_apply(void *, const double *, double *);
void orc_knn(void *, int, int32_t *, double *);
void orc_lsq_rows(void *, const KernelC *, double, double, int, const int32_t *, double *);
}

static int fail(const char *m) {
  std::fprintf(stderr, "FAIL: %s\n", m);
  return 1;
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(-0.1, 1.1);
  for (int dim = 1; dim <= 3; ++dim)
    for (int per = 0; per < 2; ++per) {
      const int n = 3000;
      std::vector<double> pos(static_cast<size_t>(n) * dim);
      for (double &v : pos) v = u(rng);
      pos[3] = NAN;
      std::vector<uint8_t> alive(static_cast<size_t>(n), 1);
      alive[10] = 0;
      const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
      const uint8_t pr[3] = {uint8_t(per), uint8_t(per), uint8_t(per)};
      void *c = orc_cells_create(dim, lo, hi, pr, 5.0);
      const int64_t na = orc_cells_update(c, n, pos.data(), alive.data());
      uint32_t grid[3];
      double side[3];
      const int64_t nb = orc_cells_grid(c, grid, side);
      std::vector<int32_t> order(static_cast<size_t>(na));
      std::vector<uint32_t> keys(static_cast<size_t>(na)), beg(static_cast<size_t>(nb)), end(static_cast<size_t>(nb));
      std::vector<double> sp(static_cast<size_t>(na) * dim);
      orc_cells_get(c, order.data(), keys.data(), beg.data(), end.data(), sp.data());
      const double r = dim == 1 ? 0.01 : dim == 2 ? 0.05 : 0.12;
      std::vector<int64_t> cnt(static_cast<size_t>(na));
      const int64_t tot = orc_search_count(c, na, sp.data(), r, cnt.data(), 4);
      std::vector<int64_t> rp(static_cast<size_t>(na) + 1, 0);
      for (int64_t i = 0; i < na; ++i) rp[size_t(i) + 1] = rp[size_t(i)] + cnt[size_t(i)];
      if (rp.back() != tot) return fail("count total");
      std::vector<int32_t> col(static_cast<size_t>(tot) + 1);
      orc_search_fill(c, na, sp.data(), r, rp.data(), col.data(), 4);
      for (int64_t e = 0; e < tot; ++e)
        if (col[size_t(e)] < 0 || col[size_t(e)] >= na) return fail("column out of range");
      KernelC kw{2, 0, r, 0, 0, 0};
      std::vector<double> x(static_cast<size_t>(na) * 2), y(static_cast<size_t>(na) * 2, 0.0);
      for (double &v : x) v = u(rng);
      orc_sparse_matvec(c, na, sp.data(), r, &kw, nullptr, nullptr, 2, x.data(), na, y.data(), na, 4);
      KernelC kg{1, 0, 0.05, 0, 0, 0};
      std::vector<double> yd(static_cast<size_t>(na), 0.0);
      orc_dense_matvec(dim, 200, sp.data(), na, sp.data(), &kg, 1, x.data(), na, yd.data(), na, 4);
      if (!per) {
        void *s = orc_schwarz_create(c, &kg, r, 1e-3, 1e5 * 2.220446049250313e-16);
        std::vector<double> bx(static_cast<size_t>(na));
        orc_schwarz_apply(s, x.data(), bx.data());
        orc_schwarz_destroy(s);
        std::vector<int32_t> nbr(static_cast<size_t>(na) * 8);
        std::vector<double> d2(static_cast<size_t>(na) * 8), w(static_cast<size_t>(na) * 8);
        orc_knn(c, 8, nbr.data(), d2.data());
        orc_lsq_rows(c, &kg, r, 1e-3, 8, nbr.data(), w.data());
      }
      orc_cells_destroy(c);
    }
  // octrees in 1 and 2 dimensions
  for (int dim = 1; dim <= 2; ++dim) {
    std::uniform_real_distribution<double> v(0.0, 1.0);
    const int n = 2000;
    std::vector<double> pos(static_cast<size_t>(n) * dim);
    for (double &p : pos) p = v(rng);
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    void *t = orc_octree_create(dim, lo, hi, n, pos.data(), 8.0);
    int64_t info[3];
    orc_octree_info(t, info);
    if (info[1] <= 0) return fail("octree leaves");
    orc_octree_destroy(t);
  }
  // octree + interaction lists + FMM at orders 3 and 6
  {
    std::uniform_real_distribution<double> v(0.0, 1.0);
    const int n = 1500;
    std::vector<double> pos(static_cast<size_t>(n) * 3), q(static_cast<size_t>(n)), y(static_cast<size_t>(n), 0.0);
    for (double &p : pos) p = v(rng);
    for (double &p : q) p = v(rng) - 0.5;
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    void *t = orc_octree_create(3, lo, hi, n, pos.data(), 64.0);
    int64_t info[3];
    orc_octree_info(t, info);
    int64_t counts[3];
    if (orc_octree_lists(t, counts, nullptr, nullptr, nullptr, nullptr) != 0) return fail("lists");
    std::vector<int32_t> pp(static_cast<size_t>(info[1]) + 1), ps(static_cast<size_t>(counts[0])),
        mp(static_cast<size_t>(info[0]) + 1), ms(static_cast<size_t>(counts[1]));
    orc_octree_lists(t, counts, pp.data(), ps.data(), mp.data(), ms.data());
    if (pp.back() != counts[0] || mp.back() != counts[1]) return fail("list sizes");
    KernelC k{4, 0, 1e-3, 0, 0, 0};
    if (orc_fmm_matvec(t, 3, &k, q.data(), y.data()) != 0) return fail("fmm order 3");
    if (orc_fmm_matvec(t, 6, &k, q.data(), y.data()) != 0) return fail("fmm order 6");
    orc_octree_destroy(t);
  }
  std::printf("ok\n");
  return 0;
}
