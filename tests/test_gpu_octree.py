"""GPU HyperOctree against the CPU oracle.

Bar: bit-exact.  The device build (tags, stable sort, level-by-level node /
leaf structure) gives the oracle's tags, tree order, positions, child codes
and leaves in dimensions 1-3, for uniform, clustered (the deepest level
reached), tiny and empty inputs; the domain enforcement keeps the cell
list's rule (out-of-domain and non-finite particles dropped, periodic ones
wrapped); the device dual traversal gives the oracle's P2P and M2L lists
entry for entry in dimensions 1-3; and the device query header (tests/cpp, built with
contraction on) finds exactly the brute-force FP64 neighbours of every
centre through the tree, periodic images included.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

import aboria_amd as ab
import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _positions(kind, dim, rng):
    if kind == "uniform":
        return rng.random((20000, dim)), 16
    if kind == "clustered":
        # a cluster of 300 within 1e-7 (and 40 exact duplicates) drives
        # leaves to the deepest level
        p = rng.random((5000, dim))
        p[:300] = 0.37 + 1e-7 * rng.random((300, dim))
        p[300:340] = 0.61
        return p, 8
    if kind == "tiny":
        return rng.random((5, dim)), 2
    return np.zeros((0, dim)), 4


def _build(ctx, pos, leaf, periodic=None, low=None, high=None):
    dim = pos.shape[1]
    low = [0.0] * dim if low is None else low
    high = [1.0] * dim if high is None else high
    t = ab.HyperOctree(ctx, low, high, periodic, leaf)
    gp = torch.tensor(pos, device="cuda")
    na = t.update_positions(gp)
    return t, gp, na


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("kind", ["uniform", "clustered", "tiny", "empty"])
def test_octree_matches_oracle(ctx, dim, kind):
    print("octree", dim, kind, flush=True)
    rng = np.random.default_rng(100 * dim + len(kind))
    pos, leaf = _positions(kind, dim, rng)
    t, _, na = _build(ctx, pos, leaf)
    o = oracle.Octree([0.0] * dim, [1.0] * dim, pos, leaf)
    info = t.info
    assert na == len(pos)
    assert (info.levels, info.n_nodes, info.n_leaves) == (o.levels, o.n_nodes, o.n_leaves)
    assert np.array_equal(t.nodes().cpu().numpy(), o.nodes)
    assert np.array_equal(t.leaves().cpu().numpy(), o.leaves)
    if len(pos):
        assert np.array_equal(t.order().cpu().numpy(), o.order)
        assert np.array_equal(t.tags().cpu().numpy().astype(np.int64), o.tags.astype(np.int64))
        assert np.array_equal(t.positions().cpu().numpy(), o.positions)
    if kind == "clustered":
        assert o.levels == 32 // dim - 2


def test_octree_domain_enforcement(ctx):
    """Non-periodic: particles outside [low, high) and non-finite ones are
    dropped and marked dead; the tree is the oracle's over the kept ones,
    the order indexing the input.  Periodic: coordinates are wrapped in
    place first."""
    rng = np.random.default_rng(5)
    pos = rng.uniform(-0.2, 1.2, (6000, 3))
    pos[7, 1] = np.nan
    pos[8, 2] = np.inf
    t = ab.HyperOctree(ctx, [0.0] * 3, [1.0] * 3, None, 10)
    gp = torch.tensor(pos, device="cuda")
    alive = torch.ones(len(pos), dtype=torch.uint8, device="cuda")
    na = t.update_positions(gp, alive)
    keep = np.all((pos >= 0) & (pos < 1), axis=1)
    assert na == int(keep.sum())
    assert np.array_equal(alive.cpu().numpy(), keep.astype(np.uint8))
    kept = np.nonzero(keep)[0]
    o = oracle.Octree([0.0] * 3, [1.0] * 3, pos[kept], 10)
    assert np.array_equal(t.order().cpu().numpy(), kept[o.order])
    assert np.array_equal(t.nodes().cpu().numpy(), o.nodes)
    assert np.array_equal(t.leaves().cpu().numpy(), o.leaves)
    # periodic in x and z: wrapped in place, y still drops
    t2 = ab.HyperOctree(ctx, [0.0] * 3, [1.0] * 3, [True, False, True], 10)
    gp2 = torch.tensor(pos, device="cuda")
    na2 = t2.update_positions(gp2)
    wrapped = gp2.cpu().numpy()
    keep2 = np.isfinite(pos).all(axis=1) & (pos[:, 1] >= 0) & (pos[:, 1] < 1)
    assert na2 == int(keep2.sum())
    w = pos.copy()
    for d in (0, 2):  # enforce_domain_lambda's repeated +/- L
        col = w[:, d]
        fin = np.isfinite(col)
        for _ in range(4):
            col[fin & (col < 0)] += 1.0
        for _ in range(4):
            col[fin & (col >= 1)] -= 1.0
    assert np.array_equal(wrapped[keep2], w[keep2])
    k2 = np.nonzero(keep2)[0]
    o2 = oracle.Octree([0.0] * 3, [1.0] * 3, w[k2], 10)
    assert np.array_equal(t2.order().cpu().numpy(), k2[o2.order])
    assert np.array_equal(t2.nodes().cpu().numpy(), o2.nodes)


@pytest.mark.parametrize("case", [("uniform", 3000, 16, 3), ("clustered", 8000, 8, 3),
                                  ("large", 200000, 64, 3), ("uniform", 5000, 10, 2),
                                  ("clustered", 6000, 6, 2), ("uniform", 4000, 8, 1),
                                  ("clustered", 6000, 5, 1)])
def test_octree_interaction_lists_match_oracle(ctx, case):
    kind, n, leaf, dim = case
    print("octree lists", kind, n, dim, flush=True)
    rng = np.random.default_rng(n)
    if kind == "clustered":
        c = rng.random((8, dim))
        w = 0.04 if dim > 1 else 1e-4
        pos = np.clip(c[rng.integers(0, 8, n)] + w * rng.standard_normal((n, dim)), 0,
                      np.nextafter(1.0, 0))
    else:
        pos = rng.random((n, dim))
    t, _, _ = _build(ctx, pos, leaf)
    g = t.interaction_lists()
    o = oracle.Octree([0.0] * dim, [1.0] * dim, pos, leaf).interaction_lists()
    for k in ("p2p_ptr", "p2p_src", "m2l_ptr", "m2l_src"):
        assert np.array_equal(g[k].cpu().numpy(), o[k]), k
    assert g["interactions"] == o["interactions"]


def test_fmm_lists_equal_octree_lists(ctx):
    """The FMM's pair counts are the traversal's (the FMM builds the same
    device tree and lists)."""
    rng = np.random.default_rng(77)
    pos = rng.random((30000, 3))
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, 32, 4,
                       ab.Kernel.smooth_inv_r(1e-3))
    o = oracle.Octree([0.0] * 3, [1.0] * 3, pos, 32).interaction_lists()
    info = f.info
    assert info["p2p_pairs"] == len(o["p2p_src"])
    assert info["m2l_pairs"] == len(o["m2l_src"])
    assert info["p2p_interactions"] == o["interactions"]


def _brute_digest(pos, centres, r, periodic, low, high):
    """count and input-index sum of the particles within r of every
    periodic image of every centre (the search's FP64 rule, images in
    lattice order); non-finite centres have none."""
    dim = pos.shape[1]
    L = np.asarray(high) - np.asarray(low)
    rngs = [(-1, 0, 1) if periodic[d] else (0,) for d in range(dim)]
    imgs = np.array(np.meshgrid(*rngs, indexing="ij")).reshape(dim, -1).T
    r2 = r * r
    cnt = np.zeros(len(centres), np.int64)
    sm = np.zeros(len(centres), np.int64)
    idx = np.arange(len(pos), dtype=np.int64)
    for i, c in enumerate(centres):
        if not np.all(np.isfinite(c)):
            continue
        for im in imgs:
            p = c + im * L
            acc = np.zeros(len(pos))
            for d in range(dim):
                dx = pos[:, d] - p[d]
                acc = acc + dx * dx
            hit = ~(acc > r2)
            cnt[i] += int(hit.sum())
            sm[i] += int(idx[hit].sum())
    return cnt, sm


@pytest.mark.parametrize("dim,periodic", [(3, True), (2, False), (1, True)])
def test_octree_query_header_from_external_kernel(ctx, dim, periodic):
    """tests/cpp/libdevice_query_example.so (contraction on) counts the
    neighbours of every centre through abh::octree_for_each_neighbour on an
    octree view and finds every particle's leaf with abh::octree_get_bucket:
    counts and input-index sums equal the brute-force FP64 search, and each
    particle lies in the particle range of the leaf found for it."""
    lib = ctypes.CDLL(os.path.join(HERE, "cpp", "libdevice_query_example.so"))
    lib.example_octree_count_neighbours.restype = ctypes.c_int
    lib.example_octree_get_bucket.restype = ctypes.c_int
    n, r = 6000, {3: 0.08, 2: 0.03, 1: 0.002}[dim]
    rng = np.random.default_rng(30 + dim)
    pos = rng.random((n, dim))
    t, _, _ = _build(ctx, pos, 12, [periodic] * dim)
    nq = 500
    centres = rng.uniform(-0.1, 1.1, (nq, dim))
    centres[3, 0] = np.nan
    ct = torch.tensor(centres, device="cuda")
    cnt = torch.zeros(nq, dtype=torch.int64, device="cuda")
    sm = torch.zeros(nq, dtype=torch.int64, device="cuda")
    v = t.query_view()
    torch.cuda.synchronize()
    assert lib.example_octree_count_neighbours(ctypes.byref(v), ctypes.c_void_p(ct.data_ptr()),
                                               ctypes.c_int64(nq), ctypes.c_double(r),
                                               ctypes.c_void_p(cnt.data_ptr()),
                                               ctypes.c_void_p(sm.data_ptr())) == 0
    want_c, want_s = _brute_digest(pos, centres, r, [periodic] * dim, [0.0] * dim, [1.0] * dim)
    assert np.array_equal(cnt.cpu().numpy(), want_c)
    assert np.array_equal(sm.cpu().numpy(), want_s)
    # every particle's bucket holds it
    tp = t.positions().contiguous()
    node = torch.zeros(n, dtype=torch.int64, device="cuda")
    assert lib.example_octree_get_bucket(ctypes.byref(v), ctypes.c_void_p(tp.data_ptr()),
                                         ctypes.c_int64(n), ctypes.c_void_p(node.data_ptr())) == 0
    codes = t.nodes().cpu().numpy()[node.cpu().numpy()]
    assert np.all((codes < 0) & (codes != -1))
    li = (codes.astype(np.int64) & 0x7fffffff)
    lv = t.leaves().cpu().numpy()[li]
    j = np.arange(n)
    assert np.all((lv[:, 0] <= j) & (j < lv[:, 1]))
