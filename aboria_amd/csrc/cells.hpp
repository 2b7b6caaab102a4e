// CellListOrdered state shared by the build (cells.hip) and the search /
// operator kernels (search.hip).
#pragma once

#include "common.hpp"

namespace abh {

// Grid parameters passed by value to kernels.  Mirrors the fields of
// CellListOrderedQuery used by the search (src/CellListOrdered.h:284-364).
struct GridParams {
  double lo[3], hi[3], L[3];  // L = hi - lo
  double side[3], inv[3], half[3];
  int32_t size[3];
  int32_t end_bucket[3];
  int32_t periodic[3];
  int64_t nb;
};

struct BuildCounters {
  uint32_t n_alive;
  uint32_t n_wrap_fail;
};

// enforce_domain_lambda (src/NeighbourSearchBase.h:185-238) over n
// particles: periodic dimensions of pos wrapped in place (the others are not
// written), flags[i] = 1 for a kept particle, 0 for a non-finite one or one
// outside a non-periodic dimension; alive (may be null) gets the same;
// counters->n_wrap_fail counts wraps abandoned after 64 steps.
void enforce_domain(abh_context *ctx, int dim, const GridParams &g, int64_t n, double *pos,
                    uint8_t *alive, uint32_t *flags, BuildCounters *counters);

}  // namespace abh

struct abh_cells {
  abh_context *ctx = nullptr;
  int dim = 3;
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  bool periodic[3] = {false, false, false};
  double n_leaf = 10.0;
  uint32_t size[3] = {1, 1, 1};
  double side[3] = {1, 1, 1}, inv[3] = {1, 1, 1};
  size_t sized_with_n = size_t(-1);
  int64_t n_alive = 0;
  int64_t nb = 1;
  abh::DevBuf order, keys, begin, end, pos;
  abh::DevBuf flags, offs, counters;  // build temporaries
  // bucket-pair sequence (pairs.hip): per-source run lengths, total, pairs
  abh::DevBuf pair_src, pair_total, pair_i, pair_j, pair_img;
  int64_t npairs = 0;

  abh::GridParams params() const {
    abh::GridParams g{};
    for (int d = 0; d < 3; ++d) {
      const bool on = d < dim;
      g.lo[d] = on ? lo[d] : 0.0;
      g.hi[d] = on ? hi[d] : 1.0;
      g.L[d] = g.hi[d] - g.lo[d];
      g.side[d] = on ? side[d] : 1.0;
      g.inv[d] = on ? inv[d] : 1.0;
      g.half[d] = 0.5 * g.side[d];
      g.size[d] = on ? int32_t(size[d]) : 1;
      g.end_bucket[d] = g.size[d] - 1;
      g.periodic[d] = on && periodic[d];
    }
    g.nb = nb;
    return g;
  }
};
