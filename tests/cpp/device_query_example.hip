// A kernel written outside the library against include/aboria_hip_device.hpp
// (built with hipcc's default flags, FP contraction on): counts and sums the
// neighbour indices of every centre from an abh_query_view passed by value,
// the way the reference's device lambdas use a copied CellListOrderedQuery.
#include "aboria_hip_device.hpp"

__global__ void k_count(abh_query_view q, const double *centres, int64_t nq, double r,
                        int64_t *counts, int64_t *index_sums) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int64_t c = 0, s = 0;
  abh::for_each_neighbour(q, centres + i * q.dim, r, [&](uint32_t j, const double *, double) {
    ++c;
    s += j;
  });
  counts[i] = c;
  index_sums[i] = s;
}

extern "C" int example_count_neighbours(const abh_query_view *q, const double *centres,
                                        int64_t nq, double r, int64_t *counts,
                                        int64_t *index_sums) {
  if (nq <= 0) return 0;
  k_count<<<unsigned((nq + 255) / 256), 256>>>(*q, centres, nq, r, counts, index_sums);
  if (hipGetLastError() != hipSuccess) return 1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}

// the same counts through a HyperOctree view; index sums over input indices
__global__ void k_octree_count(abh_octree_view q, const double *centres, int64_t nq, double r,
                               int64_t *counts, int64_t *index_sums) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int64_t c = 0, s = 0;
  abh::octree_for_each_neighbour(q, centres + i * q.dim, r,
                                 [&](uint32_t j, const double *, double) {
                                   ++c;
                                   s += q.order[j];
                                 });
  counts[i] = c;
  index_sums[i] = s;
}

extern "C" int example_octree_count_neighbours(const abh_octree_view *q, const double *centres,
                                               int64_t nq, double r, int64_t *counts,
                                               int64_t *index_sums) {
  if (nq <= 0) return 0;
  k_octree_count<<<unsigned((nq + 255) / 256), 256>>>(*q, centres, nq, r, counts, index_sums);
  if (hipGetLastError() != hipSuccess) return 1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}

// the leaf bucket (node index) of every point (points inside the domain)
template <int D>
__global__ void k_octree_bucket(abh_octree_view q, const double *pts, int64_t np, int64_t *node) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= np) return;
  node[i] = abh::octree_get_bucket<D>(q, pts + i * D);
}

extern "C" int example_octree_get_bucket(const abh_octree_view *q, const double *pts, int64_t np,
                                         int64_t *node) {
  if (np <= 0) return 0;
  const unsigned g = unsigned((np + 255) / 256);
  if (q->dim == 1) k_octree_bucket<1><<<g, 256>>>(*q, pts, np, node);
  else if (q->dim == 2) k_octree_bucket<2><<<g, 256>>>(*q, pts, np, node);
  else k_octree_bucket<3><<<g, 256>>>(*q, pts, np, node);
  if (hipGetLastError() != hipSuccess) return 1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : 2;
}
