"""The batched Schwarz oracle (oracle/batch_schwarz.py) pinned against the
cell-list Schwarz oracle (the C restatement of SchwartzPreconditioner over a
CellListOrdered) and against the exact inverse; CPU only."""
import numpy as np
import pytest

import oracle
from oracle import batch_schwarz as bs

GAUSS, MQ = 1, 3


def cell_list_domains(pos, low, high, n_leaf, kernel=(GAUSS, 0.05), radius=0.0, shift=1e-3):
    oc = oracle.Cells(low, high, [0] * len(low), n_leaf)
    oc.update(pos.copy())
    sw = oc.schwarz(kernel, radius, shift)
    return oc, sw


@pytest.mark.parametrize("dim,n,leaf", [(2, 3000, 30.0), (3, 2500, 40.0), (1, 400, 12.0),
                                        (2, 300, 500.0)])
def test_domains_equal_cell_list_schwarz(dim, n, leaf):
    """Same domains, as sets of original particle indices, as the cell-list
    restatement (whose interior is the position test against the bucket
    bounds and whose buffer is the scaled-ball bucket search)."""
    rng = np.random.default_rng(40 + dim)
    pos = rng.random((n, dim))
    low, high = [0.0] * dim, [1.0] * dim
    off, idx, nint = bs.domains(np.array([0, n]), pos, low, high, leaf)
    oc, sw = cell_list_domains(pos, low, high, leaf)
    assert len(nint) == sw.n_domains
    order = oc.order
    for d in range(sw.n_domains):
        ids = order[sw.indices[sw.offsets[d]:sw.offsets[d + 1]]]
        ni = sw.n_interior[d]
        mine = idx[off[d]:off[d + 1]]
        assert nint[d] == ni
        assert sorted(mine[:ni]) == sorted(ids[:ni])
        assert sorted(mine[ni:]) == sorted(ids[ni:])
    # every particle is interior to exactly one domain
    interior = np.concatenate([idx[off[d]:off[d] + nint[d]] for d in range(len(nint))])
    assert np.array_equal(np.sort(interior), np.arange(n))


def test_batch_domains_are_the_problems_domains():
    rng = np.random.default_rng(7)
    sizes = [700, 1, 0, 1500, 64]
    off = np.concatenate([[0], np.cumsum(sizes)])
    pos = rng.random((off[-1], 2))
    low, high = [0.0, 0.0], [1.0, 1.0]
    o_all, i_all, n_all = bs.domains(off, pos, low, high, 20.0)
    parts_i, parts_n = [], []
    for p in range(len(sizes)):
        if sizes[p] == 0:
            continue
        o, i, n = bs.domains(np.array([0, sizes[p]]), pos[off[p]:off[p + 1]], low, high, 20.0)
        parts_i.append(i + off[p])
        parts_n.append(n)
    assert np.array_equal(i_all, np.concatenate(parts_i))
    assert np.array_equal(n_all, np.concatenate(parts_n))
    assert o_all[-1] == len(i_all)


def test_grid_matches_cell_list_sizing():
    for n, leaf, dim in [(1000, 10.0, 2), (4000, 3.95, 3), (50, 100.0, 2), (12345, 64.0, 2)]:
        oc = oracle.Cells([0.0] * dim, [1.0] * dim, [0] * dim, leaf)
        oc.update(np.random.default_rng(n).random((n, dim)))
        size, side = bs.grid(n, leaf, [0.0] * dim, [1.0] * dim)
        assert tuple(size) == oc.grid
        assert np.allclose(side, oc.side, rtol=0, atol=0)


@pytest.mark.parametrize("kernel,shift", [((GAUSS, 0.1), 1e-6), ((MQ, 0.05), 0.0)])
def test_single_domain_apply_is_the_exact_inverse(kernel, shift):
    """n_particles_in_leaf above the problem size: one domain holding every
    particle, so M^-1 = K^-1; checked with the C oracle's dense matvec."""
    rng = np.random.default_rng(3)
    pos = rng.random((400, 2))
    doms = bs.domains(np.array([0, 400]), pos, [0, 0], [1, 1], 1000.0)
    assert len(doms[2]) == 1 and doms[2][0] == 400
    b = rng.normal(size=400)
    x = bs.apply(pos, doms, kernel, shift, b)
    r = oracle.dense_matvec(pos, pos, kernel, x) + shift * x - b
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8


def test_apply_equals_cell_list_schwarz():
    """Gaussian, no truncation: the batch oracle's apply equals the cell-list
    restatement's (LLT local solves) on a single problem."""
    rng = np.random.default_rng(11)
    n = 1200
    pos = rng.random((n, 2))
    kern, shift = (GAUSS, 0.03), 1e-4
    doms = bs.domains(np.array([0, n]), pos, [0, 0], [1, 1], 40.0)
    oc, sw = cell_list_domains(pos, [0, 0], [1, 1], 40.0, kern, 0.0, shift)
    b = rng.normal(size=n)
    x = bs.apply(pos, doms, kern, shift, b)
    xo = sw.apply(b[oc.order])
    ref = np.empty(n)
    ref[oc.order] = xo
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-9
