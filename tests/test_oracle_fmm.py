"""Pins the oracle's Chebyshev core, HyperOctree and black-box FMM.

References: tests/chebyshev.h:190-253 (S_n / R_n against the slow sum
1/n + 2/n sum_k T_k(y_i) T_k(x) within 1e-10; T_k(cos t_i) = cos(k t_i)),
tests/data_structures.h:327-393 (particles inside their leaf boxes),
tests/fmm.h:165-345 (full FMM against the direct sum: smooth cosine source,
multiquadric sqrt(|dx|^2 + 0.01), order 3, sum (y_fmm - y)^2 / sum y^2 < 1e-2).
"""
import math

import numpy as np
import pytest

import oracle

KM, KI = 3, 4


def cheb_T(x, k):
    return math.cos(k * math.acos(max(-1.0, min(1.0, x)))) if abs(x) <= 1 else \
        math.cosh(k * math.acosh(abs(x))) * (np.sign(x) ** k)


@pytest.mark.parametrize("n", range(1, 10))
def test_sn_against_slow_sum(n):
    rng = np.random.default_rng(n)
    for x in rng.uniform(-1.5, 1.5, 20):
        for i in range(n):
            yi = oracle.cheb_node(i, n)
            slow = (1.0 + 2.0 * sum(cheb_T(yi, k) * cheb_T(x, k) for k in range(1, n))) / n
            assert abs(oracle.cheb_Sn(x, i, n) - slow) < 1e-10


def test_chebyshev_nodes_closed_form():
    n = 4
    for i in range(n):
        x = oracle.cheb_node(i, n)
        assert abs(x - math.cos((2 * i + 1) / (2 * n) * math.pi)) < 1e-15
        for k in range(n):
            assert abs(cheb_T(x, k) - math.cos(k * (2 * i + 1) / (2 * n) * math.pi)) < 1e-10


def test_sn_interpolates_node_values():
    # S_n(y_j, i) = delta_ij: the Lagrange property of the Chebyshev interpolant
    n = 6
    for i in range(n):
        for j in range(n):
            v = oracle.cheb_Sn(oracle.cheb_node(j, n), i, n)
            assert abs(v - (1.0 if i == j else 0.0)) < 1e-13


def test_octree_invariants():
    rng = np.random.default_rng(2)
    pos = rng.random((5000, 3))
    t = oracle.Octree([0, 0, 0], [1, 1, 1], pos, 20)
    # leaves partition the sorted particles
    lv = t.leaves[np.argsort(t.leaves[:, 0])]
    nonempty = lv[lv[:, 1] > lv[:, 0]]
    assert nonempty[0, 0] == 0 and nonempty[-1, 1] == 5000
    assert np.all(nonempty[1:, 0] == nonempty[:-1, 1])
    assert np.all(lv[:, 1] - lv[:, 0] <= 20)
    # sorted tags and positions in order
    assert np.all(np.diff(t.tags) >= 0)
    assert np.array_equal(t.positions, pos[t.order])
    # each leaf's particles share the leaf's tag prefix
    for a, b in t.leaves[:50]:
        if b > a:
            tags = t.tags[a:b]
            common = tags[0] ^ tags
            assert np.all(common < (1 << 24))


def direct(pos, q, kern):
    return oracle.dense_matvec(pos, pos, kern, q)


def test_fmm_multiquadric_order3():
    # tests/fmm.h:165-215 in 3D: N=1000, 10 per leaf, cosine source
    N = 1000
    rng = np.random.default_rng(7)
    pos = rng.random((N, 3))
    t = oracle.Octree([0, 0, 0], [1, 1, 1], pos, 10)
    p = t.positions
    q = np.prod(np.cos(2 * np.pi * p), axis=1) / N
    kern = (KM, 0.1)  # sqrt(|dx|^2 + 0.01)
    y = t.fmm_matvec(3, kern, q)
    ref = direct(p, q, kern)
    assert np.sum((y - ref) ** 2) / np.sum(ref ** 2) < 1e-2


def test_fmm_smooth_inverse_order6():
    N = 2000
    rng = np.random.default_rng(8)
    pos = rng.random((N, 3))
    t = oracle.Octree([0, 0, 0], [1, 1, 1], pos, 16)
    q = rng.uniform(-1, 1, N)
    kern = (KI, 1e-3)
    y = t.fmm_matvec(6, kern, q)
    ref = direct(t.positions, q, kern)
    rel = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    assert rel < 1e-4, rel
    # lower order is less accurate
    y4 = t.fmm_matvec(4, kern, q)
    assert np.linalg.norm(y4 - ref) > np.linalg.norm(y - ref)


def test_octree_leaf_rule_is_inclusive():
    """HyperOctree's leaf rule (src/OctTree.h:328-475): a box holding at most
    n_particles_in_leaf particles is a leaf.  With n_leaf = 4, the octant
    holding exactly 4 particles is one leaf at level 1 (a strict '<' would
    split it), the octant holding 5 is split, and the lone particle of a
    third octant is a leaf of its own."""
    octant0 = [[0.1, 0.1, 0.1], [0.2, 0.3, 0.1], [0.4, 0.2, 0.3], [0.3, 0.4, 0.4]]
    octant7 = [[0.6, 0.6, 0.6], [0.9, 0.9, 0.9], [0.6, 0.9, 0.6], [0.9, 0.6, 0.9],
               [0.7, 0.7, 0.8]]
    lone = [[0.1, 0.1, 0.9]]
    pos = np.array(octant0 + octant7 + lone, dtype=np.float64)
    t = oracle.Octree([0, 0, 0], [1, 1, 1], pos, 4)
    sizes = t.leaves[:, 1] - t.leaves[:, 0]
    # the four particles of octant 0 come first in tag order: one leaf of 4
    leaf0 = [tuple(l) for l in t.leaves if l[0] == 0 and l[1] > 0]
    assert leaf0 == [(0, 4)]
    assert sorted(t.order[:4]) == [0, 1, 2, 3]
    assert sizes.max() == 4
    # octant 7's five particles end up in several leaves, none holding 5
    in7 = [i for i in range(len(pos)) if np.all(pos[t.order[i]] >= 0.5)]
    leaves7 = {tuple(l) for l in t.leaves for i in in7 if l[0] <= i < l[1]}
    assert len(leaves7) >= 2
    assert t.levels >= 2


@pytest.mark.parametrize("dim", [1, 2])
def test_fmm_lower_dimensions_converge(dim):
    """tests/fmm.h runs the FMM in 1-3 dimensions (helper_fast_methods<D>):
    in 1D and 2D the black-box FMM against the direct sum, with the error
    falling as the order rises (errors of a Chebyshev interpolant of the
    smooth kernel over well-separated boxes)."""
    N = 3000
    rng = np.random.default_rng(20 + dim)
    pos = rng.random((N, dim))
    t = oracle.Octree([0.0] * dim, [1.0] * dim, pos, 12)
    q = rng.uniform(-1, 1, N)
    kern = (KI, 1e-3)
    ref = direct(t.positions, q, kern)
    errs = []
    for order in (3, 5, 8):
        y = t.fmm_matvec(order, kern, q)
        errs.append(np.linalg.norm(y - ref) / np.linalg.norm(ref))
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[2] < 1e-6, errs
