/*
 * aboria_hip.h -- C ABI of libaboria_hip.so, the MI355X (gfx950) particle
 * hot path: CellListOrdered build, radius search, sparse / dense kernel
 * operators, Chebyshev black-box FMM and the Schwarz-preconditioned GMRES.
 *
 * Conventions
 *   - Every entry point returns an abh_status; nothing throws across the ABI.
 *     abh_last_error() gives a message for the last failure on a context.
 *   - Pointers documented as "device" are HIP device pointers (hipMalloc,
 *     torch CUDA tensors, abh_malloc).  Host pointers are marked "host".
 *   - Positions are double[n][dim], row-major (the layout of Aboria's
 *     position column, a contiguous array of Vector<double,D>).
 *   - All work is enqueued on the context's stream.  Calls that must return a
 *     size to the host (number of alive particles, number of neighbour pairs)
 *     synchronise that stream, as the reference reads the same values on the
 *     host.
 *   - Handles are not thread-safe; use one context per host thread.
 *
 * Reference interfaces each group replaces are cited per declaration
 * (paths are relative to the Aboria source tree).
 */
#ifndef ABORIA_HIP_H_
#define ABORIA_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status / context                                                    */
/* ------------------------------------------------------------------ */

typedef enum abh_status {
  ABH_OK = 0,
  ABH_ERR_INVALID_ARG = 1,   /* bad dimension, null pointer, wrong size      */
  ABH_ERR_HIP = 2,           /* a HIP runtime call failed                    */
  ABH_ERR_OUT_OF_MEMORY = 3, /* device allocation failed                     */
  ABH_ERR_CAPACITY = 4,      /* caller-provided output buffer too small      */
  ABH_ERR_UNSUPPORTED = 5,   /* kernel type / precision not implemented      */
  ABH_ERR_DOMAIN = 6,        /* position too far outside a periodic domain   */
  ABH_ERR_NOT_CONVERGED = 7, /* iterative solve hit its iteration limit      */
  ABH_ERR_NUMERICAL = 8      /* a dense factorisation met a bad pivot        */
} abh_status;

typedef struct abh_context abh_context;

/* Creates a context on HIP device `device`.  `stream` is the hipStream_t all
 * work is enqueued on; NULL is the null (default) stream.  The caller keeps
 * ownership of the stream. */
abh_status abh_context_create(int device, void *stream, abh_context **out);
abh_status abh_context_destroy(abh_context *ctx);
abh_status abh_context_sync(abh_context *ctx);
const char *abh_status_string(abh_status s);
const char *abh_last_error(const abh_context *ctx);

/* Device memory helpers, so a host without HIP headers (the C++ API layer,
 * a cgo / ctypes binding) can own device buffers.  Copies are synchronous
 * with respect to the context stream. */
abh_status abh_malloc(abh_context *ctx, void **ptr, size_t bytes);
abh_status abh_free(abh_context *ctx, void *ptr);
abh_status abh_memcpy_h2d(abh_context *ctx, void *dst_dev, const void *src_host,
                          size_t bytes);
abh_status abh_memcpy_d2h(abh_context *ctx, void *dst_host, const void *src_dev,
                          size_t bytes);
/* Device-to-device copy, asynchronous on the context stream. */
abh_status abh_memcpy_d2d(abh_context *ctx, void *dst_dev, const void *src_dev,
                          size_t bytes);
abh_status abh_memset(abh_context *ctx, void *dst_dev, int value, size_t bytes);

/* ------------------------------------------------------------------ */
/* CellListOrdered                                                     */
/* ------------------------------------------------------------------ */

typedef struct abh_cells abh_cells;

typedef struct abh_cells_info {
  int32_t dim;
  uint32_t grid[3];     /* buckets per dimension (m_size)                   */
  double side[3];       /* bucket side length (m_bucket_side_length)        */
  double low[3];        /* domain lower bound                               */
  double high[3];       /* domain upper bound                               */
  uint8_t periodic[3];  /* periodicity per dimension                        */
  int64_t n_buckets;    /* prod(grid)                                       */
  int64_t n_alive;      /* particles kept by the last update                */
} abh_cells_info;

/* Replaces neighbour_search_base::set_domain (src/NeighbourSearchBase.h:252-268)
 * as called by Particles::init_neighbour_search (src/Particles.h:445-455).
 * dim in {1,2,3}.  low/high/periodic are host arrays of length dim. */
abh_status abh_cells_create(abh_context *ctx, int dim, const double *low,
                            const double *high, const uint8_t *periodic,
                            double n_particles_in_leaf, abh_cells **out);
abh_status abh_cells_destroy(abh_cells *cells);

/* Replaces neighbour_search_base::update_positions + CellListOrdered::
 * update_positions_impl (src/NeighbourSearchBase.h:350-495,
 * src/CellListOrdered.h:190-259).
 *   pos   device, n x dim, read and written: periodic dimensions are wrapped
 *         in place (enforce_domain_lambda, src/NeighbourSearchBase.h:185-238)
 *   alive device, n bytes, may be NULL (= all alive); written in place:
 *         non-finite positions and positions outside a non-periodic domain
 *         are marked dead.
 * On return the handle holds the gather order (m_alive_indices sorted by
 * bucket, stable), the bucket index of every kept particle, bucket
 * begin/end and the positions in gather order.  *n_alive_out (host) gets the
 * number of kept particles; this synchronises the stream. */
abh_status abh_cells_update(abh_cells *cells, int64_t n, double *pos,
                            uint8_t *alive, int64_t *n_alive_out);

abh_status abh_cells_get_info(const abh_cells *cells, abh_cells_info *info);

/* Device arrays owned by the handle, valid until the next update/destroy. */
const int32_t *abh_cells_order(const abh_cells *cells);        /* n_alive   */
const uint32_t *abh_cells_bucket_index(const abh_cells *cells); /* n_alive  */
const uint32_t *abh_cells_bucket_begin(const abh_cells *cells); /* n_buckets*/
const uint32_t *abh_cells_bucket_end(const abh_cells *cells);   /* n_buckets*/
const double *abh_cells_positions(const abh_cells *cells); /* n_alive x dim */

/* The query as a plain struct of device pointers and grid scalars, the
 * counterpart of CellListOrderedQuery (src/CellListOrdered.h:284-364), which
 * the reference copies by value into device lambdas (tests/parallel.h:
 * 220-228).  include/aboria_hip_device.hpp walks it from HIP kernels with the
 * reference's visit rule.  Valid until the next update/destroy of the cell
 * list. */
typedef struct abh_query_view {
  int32_t dim;
  int32_t periodic[3];
  int32_t size[3];        /* buckets per dimension                         */
  int32_t end_bucket[3];  /* size - 1                                      */
  double low[3], high[3];
  double side[3];         /* bucket side                                   */
  double inv_side[3];     /* 1 / side, as the bucket index uses it         */
  const double *pos;      /* n x dim, gather order                         */
  const uint32_t *bucket_begin, *bucket_end;
  int64_t n;              /* particles                                     */
  int64_t n_buckets;
} abh_query_view;

/* Fills *view from the cell list's current state (no device work). */
abh_status abh_cells_query_view(const abh_cells *cells, abh_query_view *view);

/* get_neighbouring_buckets(query) (src/Search.h:857-860): the sequence of
 * (bucket i, bucket j, periodic image) triples bucket_pair_iterator yields
 * (src/Search.h:498-770), in its order -- every bucket but the last against
 * the buckets of its 3^dim block that come after it, then for each periodic
 * quadrant after 0 the face buckets against the block around their ghost
 * image -- so that each pair of adjacent buckets is visited once (the
 * symmetric half-shell of pair loops).  Pair p's position offset is
 * image[p * dim + d] * (high - low)[d], added to bucket i's particles.
 * Buckets are collapsed indices.  *npairs (host) gets the count (this
 * synchronises); with bi, bj, image all NULL nothing else is written,
 * otherwise they are device arrays of capacity (image: capacity * dim)
 * entries, ABH_ERR_CAPACITY when that is below the count. */
abh_status abh_cells_bucket_pairs(abh_cells *cells, int64_t capacity, int32_t *bi,
                                  int32_t *bj, int8_t *image, int64_t *npairs);
/* The neighbour counts of the reference's symmetric pair loop over those
 * bucket pairs (tests/neighbours.h:892-951): counts[k] (device, n_alive,
 * gather order) = 1 + the particles b != k with |x_k - x_b + offset|^2 <
 * radius^2 (strict), each pair tested once and counted for both.  Needs
 * every bucket side >= radius (ABH_ERR_INVALID_ARG otherwise), as the 3^dim
 * traversal does. */
abh_status abh_cells_pair_counts(abh_cells *cells, double radius, int32_t *counts);

/* dst[k] = src[order[k]] for k < n, elements of elem_bytes bytes (the
 * per-column gather of Particles::reorder, src/Particles.h:694-732). */
abh_status abh_gather(abh_context *ctx, const int32_t *order, int64_t n,
                      const void *src, void *dst, size_t elem_bytes);

/* ------------------------------------------------------------------ */
/* find-by-id map                                                      */
/* ------------------------------------------------------------------ */

typedef struct abh_id_map abh_id_map;

/* Replaces neighbour_search_base::init_id_map and the id-map maintenance of
 * update_positions (src/NeighbourSearchBase.h:294-298, :440-486), switched
 * on by Particles::init_id_search (src/Particles.h:460-465): the (id,
 * container index) pairs of a particle set sorted by id (ids compared as
 * unsigned 64-bit values, as the reference's size_t; equal ids keep
 * container order).  ids: device, n int64 -- the id column in container
 * order (after abh_gather of the id column by the cell list's order when
 * the particles were reordered).  The map keeps no reference to ids. */
abh_status abh_id_map_create(abh_context *ctx, int64_t n, const int64_t *ids,
                             abh_id_map **out);
/* Rebuilds the map for a new id column (reusing its device buffers). */
abh_status abh_id_map_update(abh_id_map *map, int64_t n, const int64_t *ids);
abh_status abh_id_map_destroy(abh_id_map *map);
abh_status abh_id_map_size(const abh_id_map *map, int64_t *n);
/* Device arrays owned by the map (m_id_map_key / m_id_map_value), n entries,
 * valid until the next update/destroy. */
const int64_t *abh_id_map_keys(const abh_id_map *map);
const int32_t *abh_id_map_values(const abh_id_map *map);
/* CellListOrderedQuery::find (src/CellListOrdered.h:379-388) for every
 * query: index[i] = container index of the particle whose id is ids[i] (the
 * lowest such index when several share it), or n -- the end of the
 * container -- when no particle has that id.  ids, index: device, nq
 * entries.  Asynchronous on the context stream. */
abh_status abh_id_map_find(const abh_id_map *map, int64_t nq, const int64_t *ids,
                           int32_t *index);

/* ------------------------------------------------------------------ */
/* radius search (euclidean_search, bulk form)                         */
/* ------------------------------------------------------------------ */

typedef struct abh_search abh_search;

/* Bulk replacement of euclidean_search(query, centre, r)
 * (src/Search.h:839-845) for every centre q[i], i < nq (device, nq x dim).
 * The neighbour set of centre i is exactly the set the reference search
 * iterator yields: every particle of the cell list, under every periodic
 * image the reference visits, with sum_d dx_d^2 <= r^2 in FP64.
 * Creating the handle counts the neighbours (synchronises). */
abh_status abh_search_create(abh_context *ctx, const abh_cells *cells,
                             int64_t nq, const double *q, double radius,
                             abh_search **out);
/* Re-runs the search with new centres / cells / radius, reusing the handle's
 * device buffers (no allocation once they are large enough). */
abh_status abh_search_update(abh_search *s, const abh_cells *cells, int64_t nq,
                             const double *q, double radius);
abh_status abh_search_destroy(abh_search *s);
abh_status abh_search_total(const abh_search *s, int64_t *total_pairs);
/* device, nq+1 exclusive offsets into the neighbour index array */
const int64_t *abh_search_row_ptr(const abh_search *s);
/* Writes the neighbour indices (positions in the cell list's gather order,
 * i.e. container indices after reorder) of centre i to
 * col[row_ptr[i] .. row_ptr[i+1]).  Within a row the indices are ascending
 * except where one bucket is reached through two periodic images.
 * col is device memory of at least `capacity` entries. */
abh_status abh_search_fill(abh_search *s, int32_t *col, int64_t capacity);
/* The search of abh_search_update + abh_search_fill in one traversal of the
 * candidates: an upper bound of every row (the candidates of its visited
 * buckets, no distance tests, rounded up to a multiple of 16 entries so
 * that every padded row starts on a 64-byte boundary) places the rows in a
 * padded device buffer owned by the handle, the traversal writes each row
 * there and counts it, and the rows are then packed into col, with the same
 * row_ptr and lists as the two-pass form.  The padded buffer is grow-only
 * and holds at most 4 * (sum of the row bounds + 15 * nq) bytes: about 4-5x the
 * neighbour lists at dozens of neighbours per centre, dominated by the
 * 15 * nq rounding term when centres have few neighbours.
 * *total_pairs (host) gets the number of pairs; synchronises the stream.
 * When capacity < total this returns ABH_ERR_CAPACITY after *total_pairs,
 * abh_search_total and abh_search_row_ptr have been updated: the caller can
 * size col from them and call abh_search_fill (no second count pass). */
abh_status abh_search_run_fused(abh_search *s, const abh_cells *cells, int64_t nq,
                                const double *q, double radius, int32_t *col,
                                int64_t capacity, int64_t *total_pairs);

/* ------------------------------------------------------------------ */
/* kernel descriptors and sparse operator                              */
/* ------------------------------------------------------------------ */

typedef enum abh_kernel_type {
  ABH_KERNEL_CONSTANT = 0,     /* phi = 1                                   */
  ABH_KERNEL_GAUSSIAN = 1,     /* phi = exp(-|dx|^2 / param^2)               */
  ABH_KERNEL_WENDLAND_C2 = 2,  /* q=|dx|/param, phi=(1-q)^4(4q+1) for q<1    */
  ABH_KERNEL_MULTIQUADRIC = 3, /* phi = sqrt(|dx|^2 + param^2)               */
  ABH_KERNEL_SMOOTH_INV_R = 4  /* phi = 1 / (|dx| + param)                   */
} abh_kernel_type;

/* The reference takes an arbitrary C++ functor f(dx, a, b)
 * (src/Operators.h:472-516).  Across the C ABI a kernel is a radial profile
 * times an optional per-particle affine weight:
 *   f(dx, a_i, b_j) = phi(|dx|) * (w0 + w_row * row_scalar[i]
 *                                      + w_col * col_scalar[j])
 * when scalars are supplied, and phi(|dx|) otherwise. */
typedef struct abh_kernel {
  int32_t type;
  int32_t reserved;
  double param;
  double w0, w_row, w_col;
} abh_kernel;

typedef enum abh_precision {
  ABH_FP64 = 0,               /* FP64 kernel and accumulation               */
  ABH_FP32_COMPENSATED = 1,   /* FP64 radius test, FP32 kernel, FP32 sum
                                 with Kahan-Babuska compensation            */
  ABH_MIXED = 2               /* FMM: M2L on BF16 (uncompressed) or FP16
                                 (compressed, power-of-two scaled) MFMA with
                                 a three-term split and FP32 accumulation,
                                 other stages FP64                          */
} abh_precision;

typedef struct abh_sparse_op abh_sparse_op;

/* Replaces create_sparse_operator(rows, cols, radius, f)
 * (src/Operators.h:508-516, KernelSparseConst src/Kernels.h:754-776).
 * row_pos: device, nrows x dim.  row_scalar (nrows) / col_scalar (n_alive of
 * cols) are device arrays or NULL. */
abh_status abh_sparse_op_create(abh_context *ctx, const abh_cells *cols,
                                int64_t nrows, const double *row_pos,
                                double radius, const abh_kernel *kernel,
                                const double *row_scalar,
                                const double *col_scalar, abh_sparse_op **out);
abh_status abh_sparse_op_destroy(abh_sparse_op *op);
/* number of stored pairs (neighbours within the radius, with multiplicity) */
abh_status abh_sparse_op_nnz(const abh_sparse_op *op, int64_t *nnz);
/* Replaces KernelSparse::assemble into triplets (src/Kernels.h:650-682):
 * the operator as CSR, rows in row order, every row's entries in the
 * search's order, val = f(dx) of each pair (a column reached through two
 * periodic images appears twice, as the reference's triplets do).
 * row_ptr: device, nrows+1; col / val: device, capacity >= nnz entries. */
abh_status abh_sparse_op_assemble(abh_sparse_op *op, int64_t *row_ptr,
                                  int32_t *col, double *val, int64_t capacity);
/* y += K x for nrhs right-hand sides (KernelSparse::evaluate,
 * src/Kernels.h:720-751).  x: device, column-major ncols x nrhs with leading
 * dimension ldx; y: device, nrows x nrhs, leading dimension ldy. */
abh_status abh_sparse_op_apply(abh_sparse_op *op, int nrhs, const double *x,
                               int64_t ldx, double *y, int64_t ldy,
                               abh_precision precision);
/* No counterpart in the reference (which evaluates the kernel at every
 * apply): keep the operator's CSR form (columns and kernel values of the
 * given precision, 12 bytes per stored pair for ABH_FP64, 8 for
 * ABH_FP32_COMPENSATED, plus the row pointers) on the device, so that later
 * applies in that precision stream it instead of searching and evaluating
 * again.  The values are exactly those the on-the-fly forms compute; the
 * sums run in another order (within the same tolerances).  enable = 0 frees
 * that precision's values.  ABH_ERR_OUT_OF_MEMORY when they would take more
 * than 80% of the free device memory.  Valid until the cell list changes. */
abh_status abh_sparse_op_cache(abh_sparse_op *op, abh_precision precision,
                               int enable);

/* Matrix-valued kernel: the reference's functor may return an Eigen block
 * of BlockRows x BlockCols entries (src/detail/Kernels.h:336-367, e.g. the
 * 2-vector kernel of tests/operators.h:904-933).  Entry (ii, jj), e = ii *
 * block_cols + jj, of the block of pair (i, j) is
 *   phi(|dx|) * (w0[e] + w_row[e] * row_scalar[i] + w_col[e] * col_scalar[j])
 * (a missing scalar array counts as zeros).  block_rows, block_cols in 1..4. */
typedef struct abh_block_kernel {
  int32_t type;         /* abh_kernel_type of phi                          */
  int32_t block_rows;
  int32_t block_cols;
  int32_t reserved;
  double param;
  double w0[16], w_row[16], w_col[16];
} abh_block_kernel;

typedef struct abh_sparse_block_op abh_sparse_block_op;

/* Replaces create_sparse_operator(rows, cols, radius_function, f) and
 * (radius, f) with a block-valued f (src/Operators.h:472-516, KernelSparse
 * src/Kernels.h:578-751): row i couples to the euclidean_search neighbours
 * of row_pos[i] within row_radius[i] (device, nrows; NULL = radius for every
 * row), each through the kernel's block.  The general form of
 * abh_sparse_op: one thread per row walking the neighbours with the device
 * query (include/aboria_hip_device.hpp).  row_pos, row_radius, row_scalar,
 * col_scalar stay referenced (not copied) and must not change while the
 * operator lives; the neighbour counts are fixed at creation (synchronises). */
abh_status abh_sparse_block_op_create(abh_context *ctx, const abh_cells *cols,
                                      int64_t nrows, const double *row_pos,
                                      double radius, const double *row_radius,
                                      const abh_block_kernel *kernel,
                                      const double *row_scalar,
                                      const double *col_scalar,
                                      abh_sparse_block_op **out);
abh_status abh_sparse_block_op_destroy(abh_sparse_block_op *op);
/* neighbour pairs (the operator stores block_rows * block_cols entries each) */
abh_status abh_sparse_block_op_pairs(const abh_sparse_block_op *op, int64_t *npairs);
/* y += K x (KernelSparse::evaluate, src/Kernels.h:727-751): x device,
 * ncols * block_cols; y device, nrows * block_rows (segment of particle i at
 * i * block_rows). */
abh_status abh_sparse_block_op_apply(abh_sparse_block_op *op, const double *x, double *y);
/* KernelSparse::assemble (src/Kernels.h:629-685) as CSR of the scalar
 * matrix (nrows * block_rows rows, row_ptr device of that + 1): block row
 * i * block_rows + ii holds, per neighbour j in search order, the entries of
 * columns j * block_cols .. + block_cols - 1.  col / val: device, capacity >=
 * pairs * block_rows * block_cols. */
abh_status abh_sparse_block_op_assemble(abh_sparse_block_op *op, int64_t *row_ptr,
                                        int32_t *col, double *val, int64_t capacity);

/* ------------------------------------------------------------------ */
/* HyperOctree                                                          */
/* ------------------------------------------------------------------ */

typedef struct abh_octree abh_octree;

/* Replaces HyperOctree construction + set_domain (src/OctTree.h:129-144,
 * src/NeighbourSearchBase.h:252-268; init_neighbour_search,
 * src/Particles.h:445-455): domain [low, high) per dimension, periodic
 * (dim bytes, may be NULL = none), a leaf holds at most
 * int(n_particles_in_leaf) particles unless it lies on the deepest level
 * 32 / dim - 2.  dim 1..3.  The new tree is empty: one level of 2^dim empty
 * boxes, as the reference's constructor builds. */
abh_status abh_octree_create(abh_context *ctx, int dim, const double *low,
                             const double *high, const uint8_t *periodic,
                             double n_particles_in_leaf, abh_octree **out);
abh_status abh_octree_destroy(abh_octree *tree);

/* Replaces neighbour_search_base::update_positions + HyperOctree::
 * update_positions_impl / build_tree (src/NeighbourSearchBase.h:350-495,
 * src/OctTree.h:168-225, 328-475).  pos / alive exactly as for
 * abh_cells_update (periodic dimensions wrapped in place, non-finite and
 * out-of-domain particles dropped and marked dead).  On return the handle
 * holds the tags of the kept particles sorted stably (tree order), the
 * positions in tree order, the child codes and the leaves.  *n_alive_out
 * (host, may be NULL) gets the kept count; synchronises. */
abh_status abh_octree_update(abh_octree *tree, int64_t n, double *pos,
                             uint8_t *alive, int64_t *n_alive_out);

typedef struct abh_octree_info {
  int32_t dim;
  int32_t levels;    /* m_number_of_levels                              */
  int64_t n_alive;   /* particles in the tree                           */
  int64_t n_nodes;   /* boxes below the root, empty ones included       */
  int64_t n_leaves;  /* non-empty leaves (m_leaves)                     */
} abh_octree_info;
abh_status abh_octree_get_info(const abh_octree *tree, abh_octree_info *info);

/* Device arrays owned by the handle, valid until the next update/destroy:
 * tree order -> input index (n_alive), positions in tree order
 * (n_alive x dim), sorted tags (n_alive), child codes (m_nodes, n_nodes:
 * index of the first of the 2^dim children, 0x80000000 | leaf for a leaf,
 * 0xffffffff for an empty box; the root's children are nodes 0..2^dim-1)
 * and leaves (m_leaves: n_leaves x 2, the [begin, end) particle range). */
const int32_t *abh_octree_order(const abh_octree *tree);
const double *abh_octree_positions(const abh_octree *tree);
const uint32_t *abh_octree_tags(const abh_octree *tree);
const int32_t *abh_octree_nodes(const abh_octree *tree);
const int32_t *abh_octree_leaves(const abh_octree *tree);

/* The query as a plain struct of device pointers, the counterpart of
 * HyperOctreeQuery (src/OctTree.h:699-911), passed by value into HIP
 * kernels; include/aboria_hip_device.hpp walks it (get_bucket,
 * get_buckets_near_point and the euclidean_search visit).  Valid until the
 * next update/destroy. */
typedef struct abh_octree_view {
  int32_t dim;
  int32_t levels;
  int32_t periodic[3];
  int32_t reserved;
  double low[3], high[3];  /* m_bounds                                   */
  const int32_t *nodes;    /* m_nodes                                    */
  const int32_t *leaves;   /* m_leaves, 2 per leaf                       */
  const double *pos;       /* n x dim, tree order                        */
  const int32_t *order;    /* tree order -> input index                  */
  int64_t n;
  int64_t n_nodes;
  int64_t n_leaves;
} abh_octree_view;
abh_status abh_octree_query_view(const abh_octree *tree, abh_octree_view *view);

/* The interaction lists of the dual tree traversal over the tree (the
 * strong / weak connectivity of H2Matrix::generate_matrices,
 * src/H2Matrix.h:998-1090, the same walk as
 * FastMultipoleMethod's calculate_dive_M2L_and_L2L,
 * src/FastMultipoleMethod.h:442-548, with theta_condition,
 * src/detail/FastMultipoleMethod.h:1249-1279), built on the device.
 *   p2p_ptr (n_leaves + 1) / p2p_src: per non-empty leaf the non-empty
 *     source leaves of its P2P products, in the traversal's order (strong
 *     connections in order, each expanded depth first, children in order);
 *   m2l_ptr (n_nodes + 1) / m2l_src: per node the source nodes of its M2L
 *     products (empty boxes included, as the reference lists them).
 * All four are device int32 arrays and may be NULL (a first call with NULL
 * arrays sizes them).  counts (host, 3): P2P pairs, M2L pairs, particle
 * pairs the P2P lists cover.  Synchronises. */
abh_status abh_octree_interaction_lists(const abh_octree *tree, int32_t *p2p_ptr,
                                        int32_t *p2p_src, int32_t *m2l_ptr,
                                        int32_t *m2l_src, int64_t *counts);

/* ------------------------------------------------------------------ */
/* black-box Chebyshev FMM over a HyperOctree                          */
/* ------------------------------------------------------------------ */

typedef struct abh_fmm abh_fmm;

typedef struct abh_fmm_info {
  int32_t order;        /* Chebyshev nodes per dimension                  */
  int32_t levels;       /* deepest octree level                           */
  int64_t n_nodes;      /* octree boxes below the root (incl. empty)      */
  int64_t n_leaves;     /* non-empty leaves                               */
  int64_t m2l_pairs;    /* M2L (target, source) box pairs                 */
  int64_t p2p_pairs;    /* P2P (target leaf, source leaf) pairs           */
  int64_t p2p_interactions; /* particle pairs evaluated by P2P            */
} abh_fmm_info;

/* Replaces create_fmm_operator<N>(particles, particles, f_pos, f)
 * (src/Operators.h:396-407, KernelFMM src/Kernels.h:533-576) for a particle
 * set searched by a HyperOctree (src/OctTree.h:102-475) built here with
 * n_particles_in_leaf.  dim in 1..3; order (Chebyshev nodes per dimension)
 * in 3..8.  pos: device, n x dim.  Builds the tree, the interaction lists
 * and the M2L operators (synchronises; the M2L operators take
 * levels x 7^dim x order^(2 dim) x 8 bytes, 3.6 GB for order 8, dim 3 and
 * 5 levels).  In 1D and 2D the applies run the M2L over the dual
 * traversal's weak lists in FP64 on the vector units. */
abh_status abh_fmm_create(abh_context *ctx, int dim, const double *low,
                          const double *high, int64_t n, const double *pos,
                          double n_particles_in_leaf, int order,
                          const abh_kernel *kernel, abh_fmm **out);
abh_status abh_fmm_destroy(abh_fmm *fmm);
abh_status abh_fmm_get_info(const abh_fmm *fmm, abh_fmm_info *info);
/* device, n: tree order (sorted position -> index of the input position) */
const int32_t *abh_fmm_order(const abh_fmm *fmm);
/* y += K x (FastMultipoleMethod::matrix_vector_multiply,
 * src/FastMultipoleMethod.h:110-167) for nrhs right-hand sides; x, y are
 * indexed like the positions given at creation (column-major, leading
 * dimensions ldx / ldy).  precision: ABH_FP64 or ABH_MIXED.  MIXED runs the
 * uncompressed M2L of orders up to 6 as a BF16 split and, after
 * abh_fmm_compress with a padded rank up to 192, the compressed M2L as an
 * FP16 split with power-of-two scaling per level; otherwise (uncompressed
 * orders 7 and 8, whose BF16 split would floor the error near 1e-6) its M2L
 * is FP64.  P2M, L2P and P2P run in FP64 on the matrix cores in both.  In
 * 1D and 2D both precisions compute in FP64. */
abh_status abh_fmm_apply(abh_fmm *fmm, int nrhs, const double *x, int64_t ldx,
                         double *y, int64_t ldy, abh_precision precision);

/* Replaces KernelH2::compress(tol) (src/Kernels.h:523, the H2 matrix's
 * low-rank M2L blocks): per level, the M2L operators K_o become
 * U (U^T K_o U) U^T with U the left singular vectors of the level's stacked
 * far-offset operators whose singular values exceed eps times the largest
 * (eps in (0, 1); eps <= 0 restores the uncompressed operators; dim 3
 * only, ABH_ERR_UNSUPPORTED otherwise).  Later
 * applies (either precision) contract the rank-k operators in FP64 on the
 * matrix cores.  padded_rank (host, may be NULL): the rank the kernels run
 * at (the largest level rank rounded up to a multiple of 16 up to 192, then
 * to 224, 256, 320, 384, 448 or 512); ranks (host, levels - 1 ints, may be NULL): the rank per M2L level
 * 2..levels.  Synchronises. */
abh_status abh_fmm_compress(abh_fmm *fmm, double eps, int *padded_rank,
                            int *ranks);
/* The M2L autotuner's choice (no counterpart in the reference, which runs
 * its M2L on the CPU): the first MIXED apply after abh_fmm_compress times
 * one pass of the FP16 compressed M2L over every level with 1 and with 2
 * sibling tiles per block and keeps the faster.  tiles_per_block (host, may
 * be NULL): 0 before tuning, else 1 or 2; ms (host, 2 floats, may be NULL):
 * the two passes' times in milliseconds. */
abh_status abh_fmm_m2l_tuning(const abh_fmm *fmm, int *tiles_per_block, float *ms);
/* Fixes the FP16 compressed M2L's tiles per block (1 or 2), or 0: tune
 * again at the next MIXED apply.  Results do not depend on it (the same
 * products in the same order). */
abh_status abh_fmm_set_m2l_tuning(abh_fmm *fmm, int tiles_per_block);

/* ------------------------------------------------------------------ */
/* batched dense kernel operators and GMRES                            */
/* ------------------------------------------------------------------ */

typedef struct abh_dense_batch abh_dense_batch;

/* Replaces create_dense_operator(particles, particles, f)
 * (src/Operators.h:309-317, KernelDense src/Kernels.h:138-268) for a batch
 * of independent particle sets evaluated together: problem p owns points
 * offsets[p] .. offsets[p+1]-1 of pos (device, total x dim) and the same
 * rows of every vector.  offsets: host, nprob+1 entries, offsets[0] == 0.
 * The kernel argument is the position difference (f(a,b) = phi(|b - a|)). */
abh_status abh_dense_batch_create(abh_context *ctx, int dim, int nprob,
                                  const int64_t *offsets, const double *pos,
                                  const abh_kernel *kernel,
                                  abh_dense_batch **out);
abh_status abh_dense_batch_destroy(abh_dense_batch *op);
/* y += K x for every problem (KernelDense::evaluate, src/Kernels.h:212-267);
 * x, y device vectors of length total.  precision: ABH_FP64 or
 * ABH_FP32_COMPENSATED (FP64 differences, FP32 kernel, compensated sum). */
abh_status abh_dense_batch_apply(abh_dense_batch *op, const double *x, double *y,
                                 abh_precision precision);

/* Replaces create_dense_operator(rows, cols, f) with distinct row and column
 * particle sets (src/Operators.h:309-317, KernelDense::evaluate
 * src/Kernels.h:212-267): y += K x, K_ij = phi(|x_col_j - x_row_i|), for nrhs
 * right-hand sides.  row_pos (nrows x dim), col_pos (ncols x dim): device;
 * x: column-major ncols x nrhs (leading dimension ldx); y: nrows x nrhs
 * (ldy).  precision: ABH_FP64 or ABH_FP32_COMPENSATED.  With fewer rows
 * than 2,048 per CU the columns are split into ranges whose partial sums the
 * context keeps in a device buffer (ranges x 16 x nrows doubles, kept for
 * later calls) and adds in range order: results are deterministic. */
abh_status abh_dense_apply(abh_context *ctx, int dim, int64_t nrows,
                           const double *row_pos, int64_t ncols,
                           const double *col_pos, const abh_kernel *kernel,
                           int nrhs, const double *x, int64_t ldx, double *y,
                           int64_t ldy, abh_precision precision);

/* abh_dense_apply without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results as abh_dense_apply (bit for bit); the column split's
 * partial sums live in a static 32 MiB device array (one per device, held
 * until the process ends) instead of the context's grow-only buffer.  A
 * shape whose split needs more (ranges x min(nrhs, 16) x nrows doubles
 * above 4 Mi; on a 256-CU device only when nrows x min(nrhs, 16) exceeds
 * about 2 million) returns ABH_ERR_CAPACITY before any launch.  Every call
 * on one device shares the array: calls that overlap on one GPU (other
 * streams, other contexts) race. */
abh_status abh_dense_apply_noalloc(abh_context *ctx, int dim, int64_t nrows,
                                   const double *row_pos, int64_t ncols,
                                   const double *col_pos,
                                   const abh_kernel *kernel, int nrhs,
                                   const double *x, int64_t ldx, double *y,
                                   int64_t ldy, abh_precision precision);

/* Autotuner of the batched dense operator (no counterpart in the
 * reference): times every P2P tile shape (256 x {1, 2, 4} rows per
 * workgroup, each source staged in LDS serving 1, 2 or 4 rows per thread) on
 * this batch once and keeps the fastest for the precision; apply and gmres
 * tune on first use.  rows_per_thread (host, may be NULL) gets the choice,
 * variant_ms (host, 3 floats, may be NULL) the measured times.  A positive
 * *rows_per_thread on entry (1, 2 or 4) sets that shape instead of tuning. */
abh_status abh_dense_batch_tune(abh_dense_batch *op, abh_precision precision,
                                int *rows_per_thread, float *variant_ms);

typedef struct abh_gmres_options {
  int32_t restart;       /* Krylov dimension m of GMRES(m)                   */
  int32_t max_iters;     /* Arnoldi steps per problem (all restarts)         */
  double tol;            /* stop when ||b - A x|| <= tol ||b||               */
  int32_t precision;     /* operator precision inside the Arnoldi process;
                            residuals, orthogonalisation and updates are
                            always FP64.  With ABH_FP32_COMPENSATED every
                            restart recomputes the true residual in FP64
                            (iterative refinement).                          */
  int32_t stall_restarts; /* 0: iterate to max_iters.  k > 0 (no counterpart
                            in the reference): stop a problem after k
                            consecutive restarts that each fail to cut its
                            true residual by 10% (converged stays 0), so a
                            batch ends once every problem has converged or
                            stagnated, e.g. at its FP64 residual floor.     */
  int32_t graph;          /* 0: the host checks after every Arnoldi step
                            whether a problem is still iterating.  1 (dense
                            batches only; ABH_ERR_UNSUPPORTED for the sparse
                            operator): the steps of a restart cycle run as
                            HIP graphs of 5 steps each, captured once per
                            solve and replayed every cycle, the host checking
                            between graphs.                                  */
} abh_gmres_options;

typedef struct abh_gmres_stats {
  int32_t iterations;    /* Arnoldi steps taken (a problem whose true
                            residual grows over a restart cycle is returned
                            at its previous iterate and stops)              */
  int32_t converged;     /* 1 if the tolerance was met                       */
  double residual;       /* final ||b - A x|| / ||b||, operator in FP64      */
} abh_gmres_stats;

/* Solves K x = b for every problem with restarted GMRES (the Krylov solve
 * the reference delegates to Eigen::GMRES, e.g. tests/rbf_interpolation.h:
 * 360-369).  b, x: device, length total; x holds the initial guess on entry.
 * stats: host array of nprob entries (may be NULL). */
abh_status abh_dense_batch_gmres(abh_dense_batch *op, const double *b, double *x,
                                 const abh_gmres_options *opts,
                                 abh_gmres_stats *stats);

/* ------------------------------------------------------------------ */
/* restricted additive Schwarz preconditioner                          */
/* ------------------------------------------------------------------ */

typedef struct abh_schwarz abh_schwarz;

typedef struct abh_schwarz_info {
  int64_t n;               /* rows: particles of the cell list (batch)     */
  int64_t n_domains;       /* domains with a non-empty interior            */
  int64_t min_size, max_size;         /* interior + buffer particles       */
  int64_t min_interior, max_interior; /* interior particles                */
  int64_t factor_bytes;    /* device bytes of the stored inverse rows      */
  int64_t failed_domains;  /* local Cholesky factorisations that met a
                              non-positive pivot (a zero pivot for the
                              signed factorisation of a dense batch)       */
} abh_schwarz_info;

/* Replaces SchwartzPreconditioner<Eigen::LLT<MatrixXd>>::analyzePattern +
 * factorize (src/Preconditioners.h:1169-1316 and :1402-1467) for the
 * operator A = K + diag_shift * I over the cell list's particles (gather
 * order), with K_ij = phi(|x_i - x_j|) when |x_i - x_j|^2 <= radius^2 (every
 * pair when radius <= 0) — the matrix an abh_sparse_op of the same kernel
 * and radius applies, plus the diagonal shift of the reference's
 * "self kernel" (tests/rbf_interpolation.h:673-680).
 * One domain per non-empty bucket: its particles inside the bucket bounds
 * are the interior, every other particle of the 3^dim block of buckets
 * around it is the buffer (get_buckets_near_point with the default
 * neighbourhood buffer).  The interior rows of each local inverse are
 * computed once, in FP64 (blocked Cholesky), and kept on the device.
 * Non-periodic cell lists only; the kernel takes no per-particle scalars. */
abh_status abh_schwarz_create(abh_context *ctx, const abh_cells *cells,
                              const abh_kernel *kernel, double radius,
                              double diag_shift, abh_schwarz **out);
/* abh_schwarz_create after SchwartzPreconditioner::set_coarse_grid_n(n)
 * (src/Preconditioners.h:1167, the coarse domain :1230-1283): one more
 * domain, all interior and no buffer, of particles sampled near the
 * (n+1)^dim lattice lo + i dx, dx = (high - low) / n - DBL_EPSILON: each
 * bucket holding c lattice points gives min(c, its particles) of them.  It
 * is factorised with the others and solved last, its rows overwriting the
 * bucket domains' (the reference's order without OpenMP).  The reference
 * samples with the unseeded std::random_shuffle in std::unordered_map
 * bucket order; here buckets are taken in ascending index and shuffled by
 * a Fisher-Yates pass on SplitMix64 seeded with seed ^ (bucket *
 * 0xD1B54A32D192ED03).  coarse_grid_n in [0, 4096] (0: abh_schwarz_create).
 * The coarse domain is the last in abh_schwarz_get_info / get_domains;
 * abh_schwarz_factorize refactorises it too.  ABH_ERR_INVALID_ARG when no
 * lattice point falls in a non-empty bucket (the reference asserts a
 * non-empty coarse domain). */
abh_status abh_schwarz_create_coarse(abh_context *ctx, const abh_cells *cells,
                                     const abh_kernel *kernel, double radius,
                                     double diag_shift, int coarse_grid_n,
                                     uint64_t seed, abh_schwarz **out);
/* Replaces SchwartzPreconditioner::factorize alone (src/Preconditioners.h:
 * 1402-1467) on the pattern analyzePattern left: the domains of `s` are
 * kept and their local matrices rebuilt and factorised for a new kernel,
 * radius or diagonal shift (the reference's compute() after a change of the
 * operator's values).  cells must hold the particles `s` was analysed on
 * (same count, gather order); preconditioners of a dense batch return
 * ABH_ERR_UNSUPPORTED.  Synchronises. */
abh_status abh_schwarz_factorize(abh_schwarz *s, const abh_cells *cells,
                                 const abh_kernel *kernel, double radius,
                                 double diag_shift);
abh_status abh_schwarz_destroy(abh_schwarz *s);
abh_status abh_schwarz_get_info(const abh_schwarz *s, abh_schwarz_info *info);
/* Domain lists as host arrays (any pointer may be NULL): offsets has
 * n_domains+1 entries; domain d is indices[offsets[d] .. offsets[d+1]), its
 * first n_interior[d] entries the interior.  Domains are in bucket order and
 * particles in bucket order, the reference's visiting order being a
 * permutation of the same sets; a coarse-grid domain comes last. */
abh_status abh_schwarz_get_domains(const abh_schwarz *s, int64_t *offsets,
                                   int32_t *indices, int32_t *n_interior);
/* x = M^-1 b (SchwartzPreconditioner::_solve_impl,
 * src/Preconditioners.h:1476-1522): x = b, then every domain overwrites its
 * interior rows with its local solution.  b, x: device, length n, distinct. */
abh_status abh_schwarz_apply(const abh_schwarz *s, const double *b, double *x);

/* SchwartzPreconditioner::analyzePattern + factorize
 * (src/Preconditioners.h:1169-1316, :1402-1467) applied to every problem of
 * a dense batch (the reference builds one per particle set): the operator is
 * K + diag_shift * I with the batch's kernel over all pairs of the problem.
 * Problem p gets the bucket grid CellListOrdered would size for its particle
 * count over [low, high] (src/CellListOrdered.h:132-186); one domain per
 * non-empty bucket, interior = the bucket's particles (membership by
 * floor((x - low) / side), clamped to the grid), buffer = the particles of
 * the other buckets of the 3^dim block around it.  Every particle must lie
 * in [low, high] (host arrays of dim entries).  The local factorisation is
 * A = L S L^T with S = diag(+-1), the Cholesky factorisation when A is
 * positive definite, so that indefinite kernels such as the multiquadric
 * are handled (the reference's LLT meets negative pivots there);
 * failed_domains counts zero pivots.  Domain indices are global (offsets
 * of the batch).  Apply with abh_schwarz_apply over the batch vector, free
 * with abh_schwarz_destroy. */
abh_status abh_dense_batch_schwarz_create(abh_dense_batch *op, const double *low,
                                          const double *high,
                                          double n_particles_in_leaf,
                                          double diag_shift, abh_schwarz **out);
/* abh_dense_batch_gmres right-preconditioned by prec (from
 * abh_dense_batch_schwarz_create; NULL = none): Eigen::GMRES<...,
 * SchwartzPreconditioner> (tests/rbf_interpolation.h:358-369) for every
 * problem.  Convergence is judged on the true residual ||b - K x||. */
/* No counterpart in the reference: r = b - K x for every problem in
 * double-double arithmetic (exact products split by fma, two-sum
 * accumulation, one final rounding), K the FP64 matrix abh_dense_batch_apply
 * evaluates.  b, x, r: device, total points. */
abh_status abh_dense_batch_residual_dd(abh_dense_batch *op, const double *b,
                                       const double *x, double *r);
/* No counterpart in the reference (its GMRES stops at the FP64 residual):
 * iterative refinement of every problem of the batch.  Each outer step
 * forms r = b - K x in double-double arithmetic (the residual of the FP64
 * matrix without the summation's rounding), stops the problems with
 * ||r|| <= tol ||b||, and adds to the others the solution d of K d = r from
 * GMRES (inner options: restart, iteration budget, relative tolerance,
 * operator precision), right-preconditioned by prec (NULL = none).
 * max_outer bounds the residual evaluations.  stats (host, one per problem,
 * may be NULL): the inner iterations summed, converged, and the last
 * double-double relative residual. */
abh_status abh_dense_batch_refine(abh_dense_batch *op, const abh_schwarz *prec,
                                  const double *b, double *x,
                                  const abh_gmres_options *inner, double tol,
                                  int max_outer, abh_gmres_stats *stats);
abh_status abh_dense_batch_gmres_precond(abh_dense_batch *op,
                                         const abh_schwarz *prec,
                                         const double *b, double *x,
                                         const abh_gmres_options *opts,
                                         abh_gmres_stats *stats);

/* ------------------------------------------------------------------ */
/* learned approximate-inverse preconditioner                          */
/* ------------------------------------------------------------------ */

/* No counterpart in the reference (the input of the learned
 * preconditioner): for every particle i of the cell list (gather order) its
 * k nearest particles, itself first, by FP64 squared distance with ties
 * broken by index.  nbr: device, n x k int32; dist2: device n x k or NULL.
 * Non-periodic cell lists only; k <= 64. */
abh_status abh_knn(abh_context *ctx, const abh_cells *cells, int k, int32_t *nbr,
                   double *dist2);

typedef struct abh_mlp abh_mlp;

/* No counterpart in the reference: the network that stands in for the
 * Schwarz local solves.  Fully connected n_in -> hidden (n_hidden layers,
 * SiLU) -> n_out (linear), FP32 parameters, initialised from seed.
 * Parameter layout (host and gradient arrays): for each layer its weight
 * matrix (out x in, row-major) then its bias. */
abh_status abh_mlp_create(abh_context *ctx, int n_in, int hidden, int n_hidden,
                          int n_out, uint64_t seed, abh_mlp **out);
abh_status abh_mlp_destroy(abh_mlp *m);
abh_status abh_mlp_num_params(const abh_mlp *m, int64_t *n);
abh_status abh_mlp_get_params(const abh_mlp *m, float *host);
abh_status abh_mlp_set_params(abh_mlp *m, const float *host);
/* y = net(x).  x: device, n x n_in FP32 row-major; y: device n x n_out.
 * bf16 = 0: FP32 arithmetic; bf16 = 1: weights and activations rounded to
 * BF16, products on MFMA with FP32 accumulation. */
abh_status abh_mlp_forward(abh_mlp *m, int64_t n, const float *x, float *y, int bf16);
/* loss = mean over rows and outputs of (net(x) - t)^2, and its gradient
 * with respect to the parameters (grad: device, num_params FP32, FP32
 * forward and backward kernels). */
abh_status abh_mlp_loss_grad(abh_mlp *m, int64_t n, const float *x, const float *t,
                             double *loss, float *grad);

#define ABH_TRAIN_GRAPH 1
#define ABH_TRAIN_FUSED 2
typedef struct abh_mlp_train_options {
  int32_t epochs;
  int32_t batch;         /* rows per Adam step                               */
  float lr, beta1, beta2, eps;  /* Adam                                      */
  float lr_decay;        /* learning rate multiplied by this every epoch    */
  int32_t flags;         /* ABH_TRAIN_GRAPH: training steps replayed from one
                            HIP graph (same kernels, same results);
                            ABH_TRAIN_FUSED: a step as one kernel per 8 rows 
                            through the whole network plus one reduction and
                            Adam kernel (at most 4 layers, widths <= 128;
                            another summation order)                        */
  uint64_t seed;         /* shuffling                                        */
} abh_mlp_train_options;

/* Adam on the mean squared error over (x, t) rows; loss_per_epoch: host
 * array of epochs entries (mean training loss of each epoch) or NULL. */
abh_status abh_mlp_train(abh_mlp *m, int64_t n, const float *x, const float *t,
                         const abh_mlp_train_options *opts, float *loss_per_epoch);
/* No counterpart in the reference: the patch-residual loss of the learned
 * local solves, loss = mean over rows of |M_r y_r - e_1|^2 with y_r =
 * net(x_r) and M_r the row's n_out x n_out patch matrix (mats: device,
 * n x n_out x n_out FP32, row-major, from abh_ainv_patch_matrices), and
 * its gradient; n_out <= 64.  The network is trained to be a patch's local
 * inverse where the operator weighs it, not to copy the inverse's
 * entries. */
abh_status abh_mlp_loss_grad_residual(abh_mlp *m, int64_t n, const float *x,
                                      const float *mats, double *loss, float *grad);
/* Adam on the patch-residual loss (options as abh_mlp_train). */
abh_status abh_mlp_train_residual(abh_mlp *m, int64_t n, const float *x,
                                  const float *mats, const abh_mlp_train_options *opts,
                                  float *loss_per_epoch);

typedef struct abh_ainv abh_ainv;

/* Approximate-inverse rows on the k-nearest-neighbour pattern: row i has
 * weights w_i on nbr_i (abh_knn) and applies as x_i = sum_k w_ik b_nbr_ik.
 * _lsq: w_i minimises || A[:, nbr_i] w - e_i ||_2 for the Schwarz operator
 * A = K + diag_shift I (same kernel/radius convention as abh_schwarz_create),
 * solved from the local dense normal equations in FP64: the training target
 * of the network and its exact counterpart.
 * _learned: w_i = out_scale * net(features_i), features_i = the relative
 * positions (x_nbr_ik - x_i) / feature_scale, k x dim FP32 (abh_ainv_features). */
abh_status abh_ainv_create_lsq(abh_context *ctx, const abh_cells *cells,
                               const abh_kernel *kernel, double radius,
                               double diag_shift, int k, abh_ainv **out);
/* _patch: the local solves of every particle's patch (its k nearest
 * neighbours, itself first; k <= 64): w_i = the particle's row of
 * (K_P + diag_shift I)^{-1}, the patch matrix of the same kernel / radius
 * convention, factorised in FP64 -- the local inverses the network is
 * trained on (the patch form of the Schwarz local solves,
 * src/Preconditioners.h:1476-1522, one overlapping domain per particle). */
abh_status abh_ainv_create_patch(abh_context *ctx, const abh_cells *cells,
                                 const abh_kernel *kernel, double radius,
                                 double diag_shift, int k, abh_ainv **out);
/* The patch matrices of rows built from the operator (_lsq / _patch), times
 * scale, in FP32: mats (device, n x k x k, row-major) -- the training input
 * of abh_mlp_train_residual. */
abh_status abh_ainv_patch_matrices(const abh_ainv *a, double scale, float *mats);
abh_status abh_ainv_create_learned(abh_context *ctx, const abh_cells *cells,
                                   abh_mlp *net, int k, double feature_scale,
                                   double out_scale, int bf16, abh_ainv **out);
abh_status abh_ainv_destroy(abh_ainv *a);
/* device copies of the pattern (n x k int32) and weights (n x k FP64);
 * either may be NULL */
abh_status abh_ainv_get(const abh_ainv *a, int32_t *nbr, double *w);
/* features of every particle (device, n x (k*dim) FP32) */
abh_status abh_ainv_features(const abh_ainv *a, double feature_scale, float *features);
abh_status abh_ainv_apply(const abh_ainv *a, const double *b, double *x);

/* ------------------------------------------------------------------ */
/* GMRES on a sparse operator                                          */
/* ------------------------------------------------------------------ */

/* Solves (K + diag_shift I) x = b for a square sparse operator (rows = the
 * cell list's particles) with restarted GMRES, right preconditioned by
 * schwarz or ainv (at most one non-NULL; both NULL = unpreconditioned) —
 * Eigen::GMRES<MatrixReplacement, SchwartzPreconditioner> in the reference
 * (tests/rbf_interpolation.h:351-369).  b, x: device, length n; x holds the
 * initial guess.  stats: host, one entry (may be NULL). */
abh_status abh_sparse_op_gmres(abh_sparse_op *op, double diag_shift,
                               const abh_schwarz *schwarz, const abh_ainv *ainv,
                               const double *b, double *x,
                               const abh_gmres_options *opts,
                               abh_gmres_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* ABORIA_HIP_H_ */
