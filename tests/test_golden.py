"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the CPU oracle): the oracle still reproduces them, and the Chebyshev
table agrees with its closed forms."""
import os

import numpy as np

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name))  # plain arrays, no pickle


def test_c1_mini_reproduced_by_oracle():
    g = load("c1_mini.npz")
    c = oracle.Cells([0, 0, 0], [1, 1, 1], [0, 0, 0], 6.0)
    c.update(g["positions"].copy())
    assert np.array_equal(np.array(c.grid), g["grid"])
    for k in ("order", "keys", "begin", "end"):
        assert np.array_equal(getattr(c, k), g[k]), k
    rp, col = c.search(c.positions, float(g["radius"]))
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(col, g["col"])
    y = c.sparse_matvec(c.positions, float(g["radius"]), (1, float(g["radius"]) / 2), g["b"])
    assert np.array_equal(y, g["y"])
    # about 27 neighbours per particle including itself (the C1 design point)
    assert 15 < g["row_ptr"][-1] / len(g["order"]) < 40


def test_chebyshev6_closed_forms():
    g = load("chebyshev6.npz")
    i = np.arange(6)
    assert np.allclose(g["nodes"], np.cos((2 * i + 1) * np.pi / 12), rtol=0, atol=1e-15)
    # S_n(x_j, i) at a node x_j equals the Lagrange delta
    for j in range(6):
        s = [oracle.cheb_Sn(g["nodes"][j], k, 6) for k in range(6)]
        assert np.allclose(s, np.eye(6)[j], atol=1e-12)
    sn = np.array([[oracle.cheb_Sn(v, k, 6) for v in g["x"]] for k in range(6)])
    assert np.array_equal(sn, g["sn"])
    # the interpolants sum to one (they reproduce constants)
    assert np.allclose(g["sn"].sum(axis=0), 1.0, atol=1e-12)
