"""Pins the oracle's sparse / dense operator restatements.

References: tests/operators.h:810-964 (sparse operator KAT: 3 particles at
x = 0, 0.09, 0.18, radius 0.1, f = s1_a + s2_b -> C*(1,2,3) = (9,18,15),
7 non-zeros), tests/operators.h:703-808 (dense operator: (s1+s2)*sum(v)).
"""
import numpy as np

import oracle

KC, KG, KW, KM, KI = 0, 1, 2, 3, 4


def test_sparse_operator_kat():
    cells = oracle.Cells([-1, -1, -1], [1, 1, 1], [0, 0, 0])
    d = 0.1
    pos = np.array([[0, 0, 0], [d * 0.9, 0, 0], [d * 1.8, 0, 0]], np.float64)
    cells.update(pos)
    # container order after the reorder; ids follow cells.order
    ids = cells.order
    s1 = np.full(3, 1.0)
    s2 = np.full(3, 2.0)
    v_by_id = np.array([1.0, 2.0, 3.0])
    v = v_by_id[ids]
    y = cells.sparse_matvec(cells.positions, d, (KC, 0.0, 0.0, 1.0, 1.0), v, row_s=s1, col_s=s2)
    expect_by_id = np.array([9.0, 18.0, 15.0])
    assert np.array_equal(y, expect_by_id[ids])  # exact, like TS_ASSERT_EQUALS
    rp, col = cells.search(cells.positions, d)
    assert rp[-1] == 7  # C_sparse.nonZeros() == 7


def test_sparse_matches_explicit_sum():
    rng = np.random.default_rng(3)
    n = 3000
    pos = rng.random((n, 3))
    cells = oracle.Cells([0, 0, 0], [1, 1, 1], [1, 1, 1], n_leaf=5.0)
    cells.update(pos)
    sp = cells.positions
    r = 0.08
    x = rng.standard_normal((n, 2))
    for kern in [(KG, 0.05), (KW, r), (KM, 0.05), (KI, 1e-3)]:
        y = cells.sparse_matvec(sp, r, kern, x)
        rp, col = cells.search(sp, r)
        # explicit sum over the neighbour lists with minimum-image dx
        ref = np.zeros_like(x)
        for i in range(0, n, 97):
            js = col[rp[i]:rp[i + 1]]
            dx = sp[js] - sp[i]
            dx -= np.round(dx)
            s = (dx * dx).sum(1)
            if kern[0] == KG:
                f = np.exp(-s / kern[1] ** 2)
            elif kern[0] == KW:
                q = np.sqrt(s) / kern[1]
                f = np.where(q < 1, (1 - q) ** 4 * (4 * q + 1), 0.0)
            elif kern[0] == KM:
                f = np.sqrt(s + kern[1] ** 2)
            else:
                f = 1.0 / (np.sqrt(s) + kern[1])
            ref[i] = f @ x[js]
            np.testing.assert_allclose(y[i], ref[i], rtol=1e-12, atol=1e-12)


def test_dense_closed_form():
    # constant kernel: y_i = sum_j x_j for every i (dense KAT shape)
    rng = np.random.default_rng(4)
    a = rng.random((5, 2))
    b = rng.random((7, 2))
    x = np.arange(1.0, 8.0)
    y = oracle.dense_matvec(a, b, (KC, 0.0), x)
    assert np.array_equal(y, np.full(5, x.sum()))
    # multiquadric against numpy
    y = oracle.dense_matvec(a, b, (KM, 0.05), x)
    d2 = ((b[None, :, :] - a[:, None, :]) ** 2).sum(-1)
    np.testing.assert_allclose(y, np.sqrt(d2 + 0.05 ** 2) @ x, rtol=1e-13)
