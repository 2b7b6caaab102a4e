"""Pins the CPU oracle of the cell-list build and euclidean_search against the
reference's own known-answer tests and against brute force.

References: tests/utils.h:52-112 (bucket index KATs), tests/neighbours.h:520-605
(single / two particle searches), tests/neighbours.h:715-765 and :968-1147
(brute-force neighbour counts with periodic images and deletions),
tests/data_structures.h:327-393 (particles inside their bucket bounds).
"""
import itertools

import numpy as np
import pytest

import oracle


def brute_force_multiset(pos, centres, r, low, high, periodic):
    """Neighbour multiset per centre: every (j, image) with
    sum_d (p_j - (c + I*L))_d^2 <= r^2, evaluated in the search's order."""
    D = pos.shape[1]
    L = np.asarray(high) - np.asarray(low)
    ranges = [(-1, 0, 1) if periodic[d] else (0,) for d in range(D)]
    out = []
    r2 = r * r
    for c in centres:
        js = []
        for img in itertools.product(*ranges):
            q = c + np.asarray(img, np.float64) * L
            dx = pos - q
            s = np.zeros(pos.shape[0])
            for d in range(D):
                s = s + dx[:, d] * dx[:, d]
            js.extend(np.nonzero(~(s > r2))[0].tolist())
        out.append(sorted(js))
    return out


def test_bucket_index_kat_collapse():
    # tests/utils.h:52-62: size (4,7,2), index (1,2,1) -> 19
    cells = oracle.Cells([0, 0, 0], [4, 7, 2], [0, 0, 0], n_leaf=1.0 / 56.0)
    pos = np.array([[1.5, 2.5, 1.5]])
    cells.update(pos)
    assert cells.grid == (4, 7, 2)
    assert cells.keys.tolist() == [19]


def test_bucket_index_kat_point():
    # tests/utils.h:76-96: (0.5,0.5,0.5) in [0,1]^3 with 5^3 buckets -> (2,2,2) -> 62
    cells = oracle.Cells([0, 0, 0], [1, 1, 1], [0, 0, 0], n_leaf=0.0075)
    cells.update(np.array([[0.5, 0.5, 0.5]]))
    assert cells.grid == (5, 5, 5)
    assert cells.keys.tolist() == [2 * 25 + 2 * 5 + 2]


def test_grid_sizing_configs():
    # n_particles_in_leaf chosen so the cell side equals the search radius
    c1 = oracle.Cells([0, 0, 0], [1, 1, 1], [0, 0, 0], n_leaf=6.0)
    c1.update(np.random.default_rng(0).random((100_000, 3)))
    assert c1.grid == (25, 25, 25)
    c2 = oracle.Cells([0, 0, 0], [1, 1, 1], [1, 1, 1], n_leaf=3.95)
    # sizing only depends on the alive count
    c2.update(np.random.default_rng(0).random((4_000_000, 3)))
    assert c2.grid == (100, 100, 100)


def _search_counts(cells, centre, r):
    rp, col = cells.search(np.asarray([centre], np.float64), r)
    return int(rp[1] - rp[0])


def test_single_particle_kat():
    # tests/neighbours.h:520-560
    r = 0.1
    cells = oracle.Cells([-1, -1, -1], [1, 1, 1], [1, 1, 1])
    cells.update(np.zeros((1, 3)))
    assert _search_counts(cells, [r / 2, r / 2, 0], r) == 1
    assert _search_counts(cells, [2 * r, 0, 0], r) == 0


def test_two_particles_kat():
    # tests/neighbours.h:562-605
    r = 0.1
    cells = oracle.Cells([-1, -1, -1], [1, 1, 1], [1, 1, 1])
    cells.update(np.array([[0.0, 0, 0], [r / 2, 0, 0]]))
    rp, col = cells.search(np.array([[1.1 * r, 0, 0]]), r)
    assert rp[1] == 1
    # the found particle is the one at r/2 (id 1)
    assert cells.order[col[0]] == 1
    assert _search_counts(cells, [0.9 * r, 0, 0], r) == 2
    assert _search_counts(cells, [1.6 * r, 0, 0], r) == 0
    assert _search_counts(cells, [0.25 * r, 0.9 * r, 0], r) == 2
    assert _search_counts(cells, [0.25 * r, 0.99 * r, 0], r) == 0


def test_periodic_wrap_single():
    # a particle across the periodic boundary is found through an image
    cells = oracle.Cells([0, 0, 0], [1, 1, 1], [1, 1, 1], n_leaf=0.01)
    cells.update(np.array([[0.99, 0.5, 0.5]]))
    assert _search_counts(cells, [0.01, 0.5, 0.5], 0.05) == 1
    cells_np = oracle.Cells([0, 0, 0], [1, 1, 1], [0, 0, 0], n_leaf=0.01)
    cells_np.update(np.array([[0.99, 0.5, 0.5]]))
    assert _search_counts(cells_np, [0.01, 0.5, 0.5], 0.05) == 0


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("n_leaf", [1.0, 10.0])
def test_brute_force(D, periodic, n_leaf):
    # tests/neighbours.h:968-1147: N particles in [-1,1]^D, one random and the
    # first particle deleted, neighbour sets against brute force.
    rng = np.random.default_rng(100 * D + 10 * periodic + int(n_leaf))
    N = 400
    r = 0.2 if D > 1 else 0.05
    pos = rng.uniform(-1, 1, (N, D)).astype(np.float32).astype(np.float64)
    alive = np.ones(N, np.uint8)
    alive[rng.integers(0, N)] = 0
    alive[0] = 0
    low, high = [-1.0] * D, [1.0] * D
    per = [periodic] * D
    cells = oracle.Cells(low, high, per, n_leaf)
    na = cells.update(pos, alive)
    assert na == N - 1 - (alive.sum() != N - 1) or na == alive.sum()
    sp = cells.positions
    rp, col = cells.search(sp, r)
    bf = brute_force_multiset(sp, sp, r, low, high, per)
    for i in range(na):
        assert sorted(col[rp[i]:rp[i + 1]].tolist()) == bf[i]


@pytest.mark.parametrize("periodic", [False, True])
def test_build_invariants(periodic):
    # tests/data_structures.h:327-393 and the stable (bucket, index) order
    rng = np.random.default_rng(7)
    N = 5000
    pos = rng.random((N, 3)) * 1.2 - 0.1  # some outside the unit cube
    alive = np.ones(N, np.uint8)
    cells = oracle.Cells([0, 0, 0], [1, 1, 1], [periodic] * 3, n_leaf=4.0)
    wrapped = pos.copy()
    na = cells.update(wrapped, alive)
    inside = np.all((pos >= 0) & (pos < 1), axis=1)
    if periodic:
        assert na == N
        assert np.all((wrapped >= 0) & (wrapped < 1))
    else:
        assert na == inside.sum()
        assert np.array_equal(alive.astype(bool), inside)
    grid = np.array(cells.grid)
    side = 1.0 / grid
    inv = 1.0 / side
    # independent restatement of the bucket index in numpy
    v = np.floor((wrapped[cells.order] - 0.0) * inv).astype(np.int64)
    key = (v[:, 0] * grid[1] + v[:, 1]) * grid[2] + v[:, 2]
    assert np.array_equal(key.astype(np.uint32), cells.keys)
    # stable order: sorted by (key, original index)
    alive_idx = np.nonzero(alive)[0]
    kk = np.floor(wrapped[alive_idx] * inv).astype(np.int64)
    kk = (kk[:, 0] * grid[1] + kk[:, 1]) * grid[2] + kk[:, 2]
    expect = alive_idx[np.lexsort((alive_idx, kk))]
    assert np.array_equal(expect, cells.order)
    # begin/end partition and histogram
    hist = np.bincount(cells.keys, minlength=cells.nb)
    assert np.array_equal(cells.end.astype(np.int64) - cells.begin, hist[:cells.nb])
    assert cells.begin[0] == 0 and cells.end[-1] == na
    assert np.array_equal(cells.positions, wrapped[cells.order])
