// Host analysis of the restricted additive Schwarz preconditioner: the
// domain lists (interior + buffer particles) of every domain.  Plain C++
// with no HIP dependency, so that the same code the library runs can be
// built and exercised on the CPU under AddressSanitizer / UBSan.
//
// Domains (SchwartzPreconditioner::analyzePattern,
// src/Preconditioners.h:1169-1230): one per non-empty bucket; its interior
// is the particles of the bucket, its buffer every other particle of the
// 3^dim block of buckets around it (get_buckets_near_point with the default
// neighbourhood buffer).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace abh {

struct SchwarzDomains {
  std::vector<int64_t> off;   // n_domains + 1 offsets into idx
  std::vector<int32_t> idx;   // domain lists, interior first
  std::vector<int32_t> nint;  // interior count per domain
  void clear() {
    off.assign(1, 0);
    idx.clear();
    nint.clear();
  }
};

// Bucket grid of a cell list: size per dimension (last dimension fastest in
// the collapsed index), side, lower bound.
struct GridView {
  int dim = 2;
  int64_t size[3] = {1, 1, 1};
  double side[3] = {1, 1, 1};
  double lo[3] = {0, 0, 0};
  double hi[3] = {1, 1, 1};   // domain bounds given to the cell list
  double inv[3] = {1, 1, 1};  // 1 / side, as the bucket index uses it
  int64_t nb() const { return size[0] * size[1] * size[2]; }
};

// Domains of a cell list (particles in gather order, pos n x dim, bucket
// begin/end).  A particle of the 3^dim block is interior when it lies in the
// half-open bounds [lo + i side, lo + (i+1) side) of the domain's bucket
// (the reference's classification, src/Preconditioners.h:1196-1214).
// Returns nullptr, or a message for inconsistent input.
inline const char *analyse_cells(const GridView &g, const std::vector<double> &pos,
                                 const std::vector<uint32_t> &beg,
                                 const std::vector<uint32_t> &end, SchwarzDomains &out) {
  const int D = g.dim;
  const int64_t nb = g.nb();
  out.clear();
  if (D < 1 || D > 3) return "dimension must be 1, 2 or 3";
  if (int64_t(beg.size()) != nb || int64_t(end.size()) != nb) return "bucket arrays do not match the grid";
  const int64_t n = int64_t(pos.size()) / D;
  const int64_t sz[3] = {g.size[0], g.size[1], g.size[2]};
  std::vector<int32_t> buf;
  for (int64_t b = 0; b < nb; ++b) {
    if (end[size_t(b)] <= beg[size_t(b)]) continue;  // empty bucket: no domain
    if (int64_t(end[size_t(b)]) > n) return "bucket range past the particles";
    int64_t iv[3] = {0, 0, 0}, r = b;
    for (int d = D - 1; d >= 0; --d) {
      iv[d] = r % sz[d];
      r /= sz[d];
    }
    double bmin[3], bmax[3];
    for (int d = 0; d < D; ++d) {
      bmin[d] = double(iv[d]) * g.side[d] + g.lo[d];
      bmax[d] = double(iv[d] + 1) * g.side[d] + g.lo[d];
    }
    buf.clear();
    const size_t first = out.idx.size();
    int64_t lo3[3] = {0, 0, 0}, hi3[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
      lo3[d] = std::max<int64_t>(iv[d] - 1, 0);
      hi3[d] = std::min<int64_t>(iv[d] + 1, sz[d] - 1);
    }
    for (int64_t x0 = lo3[0]; x0 <= hi3[0]; ++x0)
      for (int64_t x1 = lo3[1]; x1 <= hi3[1]; ++x1)
        for (int64_t x2 = lo3[2]; x2 <= hi3[2]; ++x2) {
          const int64_t nbk = (x0 * sz[1] + x1) * sz[2] + x2;
          if (int64_t(end[size_t(nbk)]) > n) return "bucket range past the particles";
          for (uint32_t p = beg[size_t(nbk)]; p < end[size_t(nbk)]; ++p) {
            bool inside = true;
            for (int d = 0; d < D; ++d) {
              const double v = pos[size_t(p) * D + d];
              if (v < bmin[d] || v >= bmax[d]) inside = false;
            }
            if (inside) out.idx.push_back(int32_t(p));
            else buf.push_back(int32_t(p));
          }
        }
    const int32_t ni = int32_t(out.idx.size() - first);
    if (ni == 0) {  // no interior: the domain writes nothing
      out.idx.resize(first);
      continue;
    }
    out.idx.insert(out.idx.end(), buf.begin(), buf.end());
    out.nint.push_back(ni);
    out.off.push_back(int64_t(out.idx.size()));
  }
  return nullptr;
}

// Bucket grid CellListOrdered::set_domain_impl sizes for np particles over
// [lo, hi] (src/CellListOrdered.h:132-186).
inline void size_grid(int D, int64_t np, const double *lo, const double *hi, double n_leaf,
                      int64_t *sz, double *side) {
  for (int d = 0; d < 3; ++d) {
    sz[d] = 1;
    side[d] = 1.0;
  }
  if (n_leaf <= double(np)) {
    double volume = 1;
    for (int d = 0; d < D; ++d) volume *= (hi[d] - lo[d]);
    const double box_side = std::pow(n_leaf / double(np) * volume, 1.0 / D);
    for (int d = 0; d < D; ++d) {
      sz[d] = int64_t(std::floor((hi[d] - lo[d]) / box_side));
      if (sz[d] == 0) sz[d] = 1;
    }
  }
  for (int d = 0; d < D; ++d) side[d] = (hi[d] - lo[d]) / double(sz[d]);
}

// Domains of a batch of independent particle sets (problem p owns points
// boff[p] .. boff[p+1]-1 of pos, n x D).  Every problem gets the bucket grid
// a cell list would size for its particle count; a domain per non-empty
// bucket, interior = the bucket's particles, buffer = the particles of the
// other buckets of the 3^dim block around it.  Bucket membership is
// floor((x - lo) / side) clamped to the grid, so every particle is interior
// to exactly one domain.  Lists are ascending per bucket, neighbour buckets
// in row-major order; indices are global.  Returns nullptr or a message.
inline const char *analyse_batch(int D, const std::vector<int64_t> &boff,
                                 const std::vector<double> &pos, const double *lo,
                                 const double *hi, double n_leaf, SchwarzDomains &out) {
  out.clear();
  if (D < 1 || D > 3) return "dimension must be 1, 2 or 3";
  if (boff.empty() || boff[0] != 0) return "offsets must start at 0";
  const int64_t ntot = int64_t(pos.size()) / D;
  if (boff.back() != ntot) return "offsets do not match the positions";
  if (ntot >= (int64_t(1) << 31)) return "batch too large for int32 indices";
  std::vector<int64_t> key, cnt;
  std::vector<int32_t> sorted;
  for (size_t p = 0; p + 1 < boff.size(); ++p) {
    const int64_t p0 = boff[p], np = boff[p + 1] - p0;
    if (np < 0) return "offsets must be non-decreasing";
    if (np == 0) continue;
    int64_t sz[3];
    double side[3];
    size_grid(D, np, lo, hi, n_leaf, sz, side);
    const int64_t nb = sz[0] * sz[1] * sz[2];
    if (nb >= (int64_t(1) << 31)) return "too many buckets";
    key.resize(size_t(np));
    cnt.assign(size_t(nb) + 1, 0);
    for (int64_t i = 0; i < np; ++i) {
      int64_t kk = 0;
      for (int d = 0; d < D; ++d) {
        const double v = pos[size_t(p0 + i) * D + d];
        if (!(std::isfinite(v) && v >= lo[d] && v <= hi[d])) return "batch particle outside [low, high]";
        const int64_t c = std::min<int64_t>(int64_t(std::floor((v - lo[d]) / side[d])), sz[d] - 1);
        kk = kk * sz[d] + c;
      }
      key[size_t(i)] = kk;
      ++cnt[size_t(kk) + 1];
    }
    for (int64_t b = 0; b < nb; ++b) cnt[size_t(b) + 1] += cnt[size_t(b)];
    sorted.resize(size_t(np));
    {
      std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
      for (int64_t i = 0; i < np; ++i) sorted[size_t(fill[size_t(key[size_t(i)])]++)] = int32_t(p0 + i);
    }
    for (int64_t b = 0; b < nb; ++b) {
      if (cnt[size_t(b) + 1] == cnt[size_t(b)]) continue;
      int64_t iv[3] = {0, 0, 0}, r = b;
      for (int d = D - 1; d >= 0; --d) {
        iv[d] = r % sz[d];
        r /= sz[d];
      }
      out.idx.insert(out.idx.end(), sorted.begin() + cnt[size_t(b)], sorted.begin() + cnt[size_t(b) + 1]);
      const int32_t ni = int32_t(cnt[size_t(b) + 1] - cnt[size_t(b)]);
      int64_t l3[3] = {0, 0, 0}, h3[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) {
        l3[d] = std::max<int64_t>(iv[d] - 1, 0);
        h3[d] = std::min<int64_t>(iv[d] + 1, sz[d] - 1);
      }
      for (int64_t x0 = l3[0]; x0 <= h3[0]; ++x0)
        for (int64_t x1 = l3[1]; x1 <= h3[1]; ++x1)
          for (int64_t x2 = l3[2]; x2 <= h3[2]; ++x2) {
            const int64_t nbk = (x0 * sz[1] + x1) * sz[2] + x2;
            if (nbk == b) continue;
            out.idx.insert(out.idx.end(), sorted.begin() + cnt[size_t(nbk)],
                           sorted.begin() + cnt[size_t(nbk) + 1]);
          }
      out.nint.push_back(ni);
      out.off.push_back(int64_t(out.idx.size()));
    }
  }
  return nullptr;
}

// SplitMix64 (Steele, Lea, Flood 2014): the sampler's random stream.
inline uint64_t splitmix64(uint64_t &state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// The coarse-grid domain of SchwartzPreconditioner::set_coarse_grid_n
// (src/Preconditioners.h:1230-1283): the (n+1)^dim lattice points
// lo + i dx, dx = (hi - lo) / n - DBL_EPSILON, are counted per bucket
// (floor((p - lo) * inv), clamped to the grid), and from every bucket
// holding c of them the first min(c, particles) particles of a shuffle of
// its particles are taken.  The reference shuffles with the unseeded
// std::random_shuffle and walks its buckets in std::unordered_map order;
// here buckets are walked in ascending index and each bucket's particles
// (local indices 0..m-1) are shuffled by a forward Fisher-Yates pass,
// a[i] <-> a[u mod (i+1)] for i = 1..m-1, with u drawn from SplitMix64
// seeded with seed ^ (bucket * 0xD1B54A32D192ED03).  out: indices in
// gather order.  Returns nullptr or a message.
inline const char *coarse_grid(const GridView &g, const std::vector<uint32_t> &beg,
                               const std::vector<uint32_t> &end, int n, uint64_t seed,
                               std::vector<int32_t> &out) {
  out.clear();
  const int D = g.dim;
  if (D < 1 || D > 3) return "dimension must be 1, 2 or 3";
  if (n <= 0) return "coarse grid size must be positive";
  const int64_t nb = g.nb();
  if (int64_t(beg.size()) != nb || int64_t(end.size()) != nb) return "bucket arrays do not match the grid";
  double dx[3] = {0, 0, 0};
  for (int d = 0; d < D; ++d) dx[d] = (g.hi[d] - g.lo[d]) / double(n) - 2.220446049250313e-16;
  std::vector<int64_t> count(size_t(nb), 0);
  int64_t npts = 1;
  for (int d = 0; d < D; ++d) npts *= int64_t(n) + 1;
  for (int64_t t = 0; t < npts; ++t) {
    int64_t r = t, b = 0;
    int64_t iv[3] = {0, 0, 0};
    for (int d = D - 1; d >= 0; --d) {
      iv[d] = r % (int64_t(n) + 1);
      r /= int64_t(n) + 1;
    }
    for (int d = 0; d < D; ++d) {
      const double p = double(iv[d]) * dx[d] + g.lo[d];
      int64_t v = int64_t(std::floor((p - g.lo[d]) * g.inv[d]));
      v = std::min<int64_t>(std::max<int64_t>(v, 0), g.size[d] - 1);
      b = b * g.size[d] + v;
    }
    ++count[size_t(b)];
  }
  std::vector<int32_t> loc;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t c = count[size_t(b)];
    const int64_t m = int64_t(end[size_t(b)]) - int64_t(beg[size_t(b)]);
    if (c == 0 || m <= 0) continue;
    loc.resize(size_t(m));
    for (int64_t i = 0; i < m; ++i) loc[size_t(i)] = int32_t(i);
    uint64_t st = seed ^ (uint64_t(b) * 0xD1B54A32D192ED03ull);
    for (int64_t i = 1; i < m; ++i) {
      const int64_t j = int64_t(splitmix64(st) % uint64_t(i + 1));
      std::swap(loc[size_t(i)], loc[size_t(j)]);
    }
    const int64_t take = std::min(c, m);
    for (int64_t i = 0; i < take; ++i) out.push_back(int32_t(beg[size_t(b)]) + loc[size_t(i)]);
  }
  if (out.empty()) return "coarse grid holds no particles";
  return nullptr;
}

}  // namespace abh
