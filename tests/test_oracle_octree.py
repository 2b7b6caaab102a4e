"""Pins the oracle's HyperOctree in dimensions 1-3 and its dual-traversal
interaction lists.

Tree: a 1-D known answer worked by hand from build_tree (src/OctTree.h:
328-475: children tabulated per active node, EMPTY / LEAF (count <=
n_leaf) / NODE, node codes first-child index / 0x80000000 | leaf / -1),
and the partition invariants (tests/data_structures.h:327-393: every
particle in exactly one leaf, inside the leaf's box).

Lists (src/H2Matrix.h:998-1090, src/FastMultipoleMethod.h:442-548):
  * a complete tree (one particle per finest box, n_leaf = 1) in 1-3
    dimensions against the closed form: same-level boxes are strongly
    connected exactly when they touch, so the M2L pairs of level l >= 2
    number (sum_x a(x))^D - (sum_x b(x))^D with a(x) the children of the
    parent's neighbours and b(x) the neighbours along one axis, and every
    leaf's P2P sources are its touching boxes;
  * on irregular trees in 1-3 dimensions, for every non-empty target leaf
    the M2L sources of the leaf and its ancestors and its P2P sources cover
    every particle exactly once, and every M2L pair is well separated
    (theta_condition false, recomputed here from the boxes) and on one
    level.
"""
import numpy as np
import pytest

import oracle

EMPTY = -1


def leaf_index(code):
    return int(np.uint32(np.int32(code)) ^ np.uint32(0x80000000))


def child_box(lo, hi, c, dim):
    lo, hi = list(lo), list(hi)
    for d in range(dim):
        mid = 0.5 * (hi[d] + lo[d])
        if (c >> (dim - 1 - d)) & 1:
            lo[d] = mid
        else:
            hi[d] = mid
    return lo, hi


def tree_geometry(t, low, high):
    """Per node: box, level, parent; walked from the root's children."""
    dim = len(low)
    nch = 1 << dim
    box = [None] * t.n_nodes
    level = np.zeros(t.n_nodes, np.int64)
    parent = np.full(t.n_nodes, -1, np.int64)
    stack = [(c, child_box(low, high, c, dim), 1, -1) for c in range(nch)]
    while stack:
        k, b, lv, p = stack.pop()
        box[k], level[k], parent[k] = b, lv, p
        code = int(t.nodes[k])
        if code >= 0:
            for c in range(nch):
                stack.append((code + c, child_box(b[0], b[1], c, dim), lv + 1, k))
    return box, level, parent


def subtree_range(t):
    """Particle range [lo, hi) of every node (children come after parents)."""
    nch = len(t.nodes)
    lo = np.full(nch, -1, np.int64)
    hi = np.full(nch, -1, np.int64)
    dim_children = None
    for k in range(t.n_nodes - 1, -1, -1):
        code = int(t.nodes[k])
        if code == EMPTY:
            continue
        if code < 0:
            lo[k], hi[k] = t.leaves[leaf_index(code)]
            continue
        if dim_children is None:
            dim_children = 1 << t.dim
        ch = [code + c for c in range(dim_children) if lo[code + c] >= 0]
        lo[k] = min(lo[c] for c in ch)
        hi[k] = max(hi[c] for c in ch)
    return lo, hi


def theta_close(tb, sb):
    dim = len(tb[0])
    md = np.finfo(np.float64).tiny
    for i in range(dim):
        md = max(md, tb[1][i] - tb[0][i])
    sq = 0.0
    for i in range(dim):
        dist = 0.0
        if tb[1][i] < sb[0][i]:
            dist = sb[0][i] - tb[1][i]
        elif sb[1][i] < tb[0][i]:
            dist = tb[0][i] - sb[1][i]
        md = max(md, sb[1][i] - sb[0][i])
        sq = sq + dist * dist
    return md * md > sq


def test_octree_1d_known_answer():
    # [0, 0.5) holds 3 > 2 particles: a node; [0.5, 1) one: leaf 0.  Its
    # children [0, 0.25) (2 particles) and [0.25, 0.5) (1) are leaves 1, 2.
    pos = np.array([[0.3], [0.8], [0.1], [0.2]])
    t = oracle.Octree([0.0], [1.0], pos, 2)
    assert t.levels == 2
    assert t.nodes.tolist() == [2, np.int32(np.uint32(0x80000000)), np.int32(np.uint32(0x80000001)),
                                np.int32(np.uint32(0x80000002))]
    assert t.leaves.tolist() == [[3, 4], [0, 2], [2, 3]]
    assert t.order.tolist() == [2, 3, 0, 1]


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_octree_partition_any_dim(dim):
    rng = np.random.default_rng(10 + dim)
    pos = np.concatenate([rng.random((3000, dim)),
                          0.3 + 1e-4 * rng.random((200, dim))])  # a dense cluster
    t = oracle.Octree([0.0] * dim, [1.0] * dim, pos, 12)
    assert np.all(np.diff(t.tags) >= 0)
    lv = t.leaves[np.argsort(t.leaves[:, 0])]
    assert lv[0, 0] == 0 and lv[-1, 1] == len(pos)
    assert np.all(lv[1:, 0] == lv[:-1, 1])
    box, level, _ = tree_geometry(t, [0.0] * dim, [1.0] * dim)
    max_level = 32 // dim - 2
    for k, code in enumerate(t.nodes):
        if code >= 0 or code == EMPTY:
            continue
        a, b = t.leaves[leaf_index(code)]
        assert b - a <= 12 or level[k] == max_level
        p = t.positions[a:b]
        assert np.all(p >= np.array(box[k][0])) and np.all(p < np.array(box[k][1]))


@pytest.mark.parametrize("dim,L", [(1, 6), (2, 4), (3, 3)])
def test_lists_complete_tree_closed_form(dim, L):
    m = 1 << L
    g = (np.arange(m) + 0.5) / m
    pos = np.array(np.meshgrid(*([g] * dim), indexing="ij")).reshape(dim, -1).T.copy()
    t = oracle.Octree([0.0] * dim, [1.0] * dim, pos, 1)
    assert t.levels == L and t.n_leaves == m ** dim
    lists = t.interaction_lists()
    want_m2l = 0
    for lev in range(2, L + 1):
        side = 1 << lev
        x = np.arange(side)
        px = x // 2
        a = 2 * (np.minimum(px + 1, side // 2 - 1) - np.maximum(px - 1, 0) + 1)
        b = np.minimum(x + 1, side - 1) - np.maximum(x - 1, 0) + 1
        want_m2l += int(a.sum()) ** dim - int(b.sum()) ** dim
    assert len(lists["m2l_src"]) == want_m2l
    x = np.arange(m)
    b = np.minimum(x + 1, m - 1) - np.maximum(x - 1, 0) + 1
    assert lists["p2p_ptr"][-1] == int(b.sum()) ** dim
    assert lists["interactions"] == int(b.sum()) ** dim  # one particle per leaf


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("case", ["uniform", "clustered", "tiny"])
def test_lists_partition_and_separation(case, dim):
    rng = np.random.default_rng({"uniform": 1, "clustered": 2, "tiny": 3}[case] + 10 * dim)
    if case == "uniform":
        pos, leaf = rng.random((3000, dim)), 16
    elif case == "clustered":
        c = rng.random((6, dim))
        pos = np.clip(c[rng.integers(0, 6, 4000)] + 0.03 * rng.standard_normal((4000, dim)),
                      0, np.nextafter(1.0, 0))
        leaf = 8
    else:
        pos, leaf = rng.random((20, dim)), 3
    t = oracle.Octree([0.0] * dim, [1.0] * dim, pos, leaf)
    L = t.interaction_lists()
    box, level, parent = tree_geometry(t, [0.0] * dim, [1.0] * dim)
    lo, hi = subtree_range(t)
    n = len(pos)
    mp, ms = L["m2l_ptr"], L["m2l_src"]
    for k in range(t.n_nodes):
        for s in ms[mp[k]:mp[k + 1]]:
            assert level[s] == level[k]
            assert not theta_close(box[k], box[s])
    pp, ps = L["p2p_ptr"], L["p2p_src"]
    inter = 0
    for k, code in enumerate(t.nodes):
        if code >= 0 or code == EMPTY:
            continue
        li = leaf_index(code)
        cover = np.zeros(n + 1, np.int64)
        a = k
        while a >= 0:
            for s in ms[mp[a]:mp[a + 1]]:
                if lo[s] >= 0:
                    cover[lo[s]] += 1
                    cover[hi[s]] -= 1
            a = parent[a]
        for sl in ps[pp[li]:pp[li + 1]]:
            sa, sb = t.leaves[sl]
            assert sb > sa
            cover[sa] += 1
            cover[sb] -= 1
            inter += (t.leaves[li][1] - t.leaves[li][0]) * (sb - sa)
        assert np.all(np.cumsum(cover)[:n] == 1), "leaf %d" % li
    assert inter == L["interactions"]


def test_lists_empty_tree():
    t = oracle.Octree([0.0] * 3, [1.0] * 3, np.zeros((0, 3)), 4)
    assert t.levels == 1 and t.n_nodes == 8 and t.n_leaves == 0
    assert np.all(t.nodes == EMPTY)
    L = t.interaction_lists()
    assert L["p2p_ptr"].tolist() == [0] and len(L["m2l_src"]) == 0
