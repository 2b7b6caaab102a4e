// CPU AddressSanitizer / UBSan run of the Schwarz host analysis
// (aboria_amd/csrc/schwarz_analysis.hpp, the code the library runs before
// its device factorisation) at full workload size: the 512-problem batch of
// 2D point sets with sizes log-uniform in [1,000, 64,000) and 64 particles
// per bucket, and a 400,056-point cell list.  Checks the invariants every
// domain list must satisfy; any out-of-bounds access aborts under ASan.
//
// usage: asan_schwarz_analysis [n_problems]   (prints "ok ..." on success)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../aboria_amd/csrc/schwarz_analysis.hpp"

static int fail(const char *m) {
  std::fprintf(stderr, "FAIL: %s\n", m);
  return 1;
}

// every particle interior to exactly one domain, indices in range, each
// domain's list free of duplicates
static int check(const abh::SchwarzDomains &dm, int64_t n) {
  if (dm.off.size() != dm.nint.size() + 1) return fail("offset count");
  std::vector<int> seen(size_t(n), 0), mark(size_t(n), -1);
  for (size_t d = 0; d + 1 < dm.off.size(); ++d) {
    const int64_t a = dm.off[d], b = dm.off[d + 1];
    if (b < a || b > int64_t(dm.idx.size())) return fail("offsets");
    if (dm.nint[d] <= 0 || dm.nint[d] > b - a) return fail("interior count");
    for (int64_t e = a; e < b; ++e) {
      const int32_t p = dm.idx[size_t(e)];
      if (p < 0 || p >= n) return fail("index out of range");
      if (mark[size_t(p)] == int(d)) return fail("duplicate in a domain");
      mark[size_t(p)] = int(d);
      if (e < a + dm.nint[d]) ++seen[size_t(p)];
    }
  }
  for (int64_t i = 0; i < n; ++i)
    if (seen[size_t(i)] != 1) return fail("particle not interior to exactly one domain");
  return 0;
}

int main(int argc, char **argv) {
  const int nprob = argc > 1 ? std::atoi(argv[1]) : 512;
  std::mt19937_64 rng(4000);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  // batch
  std::vector<int64_t> off(1, 0);
  for (int p = 0; p < nprob; ++p) {
    const double l = std::log(1000.0) + u(rng) * (std::log(64000.0) - std::log(1000.0));
    off.push_back(off.back() + int64_t(std::floor(std::exp(l))));
  }
  std::vector<double> pos(size_t(off.back()) * 2);
  for (double &v : pos) v = u(rng);
  pos[0] = 1.0;  // a particle on the upper face: clamped into the last bucket
  const double lo[2] = {0.0, 0.0}, hi[2] = {1.0, 1.0};
  abh::SchwarzDomains dm;
  if (const char *e = abh::analyse_batch(2, off, pos, lo, hi, 64.0, dm)) return fail(e);
  if (check(dm, off.back())) return 1;
  // rejected input: a particle outside [low, high]
  {
    std::vector<double> bad(pos.begin(), pos.begin() + 2000);
    bad[5] = 1.5;
    std::vector<int64_t> o2 = {0, 1000};
    abh::SchwarzDomains d2;
    if (!abh::analyse_batch(2, o2, bad, lo, hi, 64.0, d2)) return fail("outside point accepted");
  }
  std::printf("batch: %lld points, %zu domains, %zu list entries\n", (long long)off.back(),
              dm.nint.size(), dm.idx.size());
  // cell list: 632 x 633 jittered grid, 24 x 24 buckets, counting sort by bucket
  const int nx = 632, ny = 633;
  abh::GridView g;
  g.dim = 2;
  g.size[0] = 24;
  g.size[1] = 24;
  g.side[0] = g.side[1] = 1.0 / 24;
  g.inv[0] = g.inv[1] = 24.0;
  std::vector<double> p2;
  std::vector<int64_t> key;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const double x = (i + 0.5 + 0.6 * (u(rng) - 0.5)) / nx;
      const double y = (j + 0.5 + 0.6 * (u(rng) - 0.5)) / ny;
      p2.push_back(x);
      p2.push_back(y);
      key.push_back(int64_t(std::floor(x * 24)) * 24 + int64_t(std::floor(y * 24)));
    }
  const int64_t n = int64_t(key.size());
  std::vector<uint32_t> beg(576, 0), end(576, 0);
  std::vector<int64_t> cnt(577, 0);
  for (int64_t k : key) ++cnt[size_t(k) + 1];
  for (int b = 0; b < 576; ++b) cnt[size_t(b) + 1] += cnt[size_t(b)];
  std::vector<double> sorted(p2.size());
  {
    std::vector<int64_t> f(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < n; ++i) {
      const int64_t t = f[size_t(key[size_t(i)])]++;
      sorted[size_t(t) * 2] = p2[size_t(i) * 2];
      sorted[size_t(t) * 2 + 1] = p2[size_t(i) * 2 + 1];
    }
  }
  for (int b = 0; b < 576; ++b) {
    beg[size_t(b)] = uint32_t(cnt[size_t(b)]);
    end[size_t(b)] = uint32_t(cnt[size_t(b) + 1]);
  }
  abh::SchwarzDomains dc;
  if (const char *e = abh::analyse_cells(g, sorted, beg, end, dc)) return fail(e);
  if (check(dc, n)) return 1;
  // rejected input: bucket ranges past the particle count
  {
    std::vector<uint32_t> e2 = end;
    e2[575] = uint32_t(n + 5);
    abh::SchwarzDomains d3;
    if (!abh::analyse_cells(g, sorted, beg, e2, d3)) return fail("bad bucket range accepted");
  }
  // the coarse-grid sample: distinct particles, each in a bucket; sizes
  // from a single lattice point to more points than particles
  size_t ncoarse = 0;
  for (int cn : {1, 7, 23, 60}) {
    std::vector<int32_t> cs;
    if (const char *e = abh::coarse_grid(g, beg, end, cn, 11, cs)) return fail(e);
    std::vector<uint8_t> seen(size_t(n), 0);
    for (int32_t i : cs) {
      if (i < 0 || i >= n || seen[size_t(i)]) return fail("coarse sample out of range or repeated");
      seen[size_t(i)] = 1;
    }
    ncoarse += cs.size();
  }
  {
    std::vector<int32_t> cs;
    if (!abh::coarse_grid(g, beg, end, 0, 11, cs)) return fail("coarse size 0 accepted");
  }
  std::printf("ok batch %zu domains, cells %zu domains, coarse %zu\n", dm.nint.size(),
              dc.nint.size(), ncoarse);
  return 0;
}
