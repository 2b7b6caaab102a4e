"""The bucket-pair oracle (bucket_pair_iterator restated as the reference's
state machine) against closed forms and brute force:

* the pairs visit every unordered pair of distinct adjacent buckets of the
  regular grid exactly once (j after i), then, for each periodic quadrant
  after 0 in lattice order, every face bucket against the buckets around
  its ghost image -- each pair of buckets adjacent across a periodic face
  once;
* neighbour counts by traversal (tests/neighbours.h:892-951, strict radius)
  equal a brute-force count over all periodic images, on the reference's
  test configurations (tests/neighbours.h:1326-1370).
"""
import itertools

import numpy as np
import pytest

import oracle


def grid_cells(dim, n, r, periodic, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, dim))
    n_leaf = n * r ** dim / 2.0 ** dim
    c = oracle.Cells([-1.0] * dim, [1.0] * dim, [periodic] * dim, n_leaf)
    c.update(pos.copy())
    return c


def brute_counts(pos, r, periodic, L=2.0):
    n, dim = pos.shape
    imgs = list(itertools.product(*([(-1, 0, 1)] if periodic else [(0,)]) * dim))
    cnt = np.zeros(n, np.int64)
    for im in imgs:
        d = pos[None, :, :] - pos[:, None, :] - np.array(im) * L
        cnt += ((d * d).sum(-1) < r * r).sum(1)
    return cnt


@pytest.mark.parametrize("dim,periodic", [(1, False), (1, True), (2, False), (2, True),
                                          (3, False), (3, True)])
def test_pairs_are_the_half_shell(dim, periodic):
    c = grid_cells(dim, 500, 0.45, periodic, dim)
    size = np.array(c.grid)
    bi, bj, q = c.bucket_pairs()
    coords = lambda b: np.stack(np.unravel_index(b, size), 1)
    ci, cj = coords(bi), coords(bj)
    dom = np.all(q == 0, axis=1)
    # domain-domain: distinct adjacent buckets, j after i, each once
    assert np.all(bj[dom] > bi[dom])
    assert np.all(np.abs(ci[dom] - cj[dom]) <= 1)
    got = set(zip(bi[dom].tolist(), bj[dom].tolist()))
    assert len(got) == dom.sum()
    want = set()
    nb = int(np.prod(size))
    for a in range(nb):
        for b in range(a + 1, nb):
            if np.all(np.abs(coords(np.array([a]))[0] - coords(np.array([b]))[0]) <= 1):
                want.add((a, b))
    assert got == want
    # periodic quadrants: only after 0 in lattice order, i on the face, j
    # adjacent to i's ghost image i + q * size
    if periodic:
        quads = [t for t in itertools.product((-1, 0, 1), repeat=dim)]
        after = quads[quads.index((0,) * dim) + 1:]
        assert set(map(tuple, q[~dom].tolist())) <= set(after)
        ghost = ci[~dom] + q[~dom] * size
        assert np.all(np.abs(ghost - cj[~dom]) <= 1)
    else:
        assert dom.all()


@pytest.mark.parametrize("dim,n,r,periodic", [(1, 14, 0.1, False), (1, 14, 0.1, True),
                                              (1, 1000, 0.1, True), (2, 1000, 0.1, False),
                                              (2, 1000, 0.5, True), (2, 1000, 0.2, True),
                                              (3, 1000, 0.2, True), (3, 1000, 0.2, False)])
def test_pair_counts_equal_brute_force(dim, n, r, periodic):
    c = grid_cells(dim, n, r, periodic, 100 + n + dim)
    assert min(c.grid) >= 3 or not periodic
    got = c.pair_counts(r)
    assert np.array_equal(got, brute_counts(c.positions, r, periodic))
