// The device HyperOctree (src/OctTree.h:102-475) shared by the octree entry
// points (octree.hip) and the FMM (fmm.hip).
#pragma once

#include <vector>

#include "common.hpp"

namespace abh {

struct Octree {
  int dim = 3;
  int max_level = 8;  // 32 / dim - 2 (src/OctTree.h:131)
  double n_leaf = 1.0;
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  bool periodic[3] = {false, false, false};
  int64_t n_alive = 0;
  int levels = 0;
  int64_t n_nodes = 0, n_leaves = 0;
  // particles: sorted tags, tree order -> input index, positions in tree order
  DevBuf keys, order, pos;
  // nodes: child codes (m_nodes: first child index, 0x80000000 | leaf, or
  // 0xffffffff for an empty box), tag and parent node (-1 at level 1)
  DevBuf nodes, ntag, parent;
  // non-empty leaves: particle range (m_leaves) and node
  DevBuf leaves, leaf_node;
  // internal nodes, level by level, in node order
  DevBuf internal;
  std::vector<int64_t> level_begin;  // nodes of level l: [level_begin[l-1], level_begin[l])
  std::vector<int64_t> leaf_begin;   // leaves of level l: [leaf_begin[l-1], leaf_begin[l])
  std::vector<int64_t> par_off;      // internal nodes of level l: [par_off[l], par_off[l+1])
  // build temporaries
  DevBuf flags, offs, counters, act_a, act_b, range, cflags, total;
};

// neighbour_search_base::update_positions + HyperOctree::update_positions_impl
// (src/NeighbourSearchBase.h:350-495, src/OctTree.h:168-225): enforces the
// domain (periodic dimensions wrapped in place, non-finite positions and
// positions outside a non-periodic domain dropped and marked dead in alive
// when given), tags the kept particles, sorts them stably by tag, gathers
// their positions and builds the tree level by level.
void octree_update(abh_context *ctx, Octree &t, int64_t n, double *pos, uint8_t *alive);

// Per node its box (lo[dim] then hi[dim]) and integer box coordinates
// (dim per node), as octree_child_iterator::get_bounds halves the domain
// (src/OctTree.h:633-644).  box and coord are device arrays of n_nodes rows.
void octree_geometry(abh_context *ctx, const Octree &t, double *box, int32_t *coord);

// The dual traversal's interaction lists (src/H2Matrix.h:998-1090,
// src/FastMultipoleMethod.h:442-548 with theta_condition,
// src/detail/FastMultipoleMethod.h:1249-1279).
struct OctreeLists {
  DevBuf p2p_ptr;   // n_leaves + 1 (int32)
  DevBuf p2p_src;   // source leaf indices (int32)
  DevBuf weak_ptr;  // n_nodes + 1 (int32), filled when weak lists are asked for
  DevBuf weak_src;  // source nodes of the M2L pairs
  int64_t m2l = 0, p2p_pairs = 0, interactions = 0;
};
void octree_lists(abh_context *ctx, const Octree &t, const double *box, OctreeLists &L,
                  bool weak_lists);

}  // namespace abh
