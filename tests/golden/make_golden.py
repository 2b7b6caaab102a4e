"""Writes the small golden fixtures of tests/golden/ from the CPU oracle.

  c1_mini.npz   3D, N=2,000 uniform in the unit cube (PCG64 seed 1000),
                non-periodic, n_particles_in_leaf 6, radius = bucket side:
                gather order, bucket index per particle, bucket begin/end,
                the radius-search CSR of every particle, and y = K b for the
                Gaussian exp(-|dx|^2 / (radius/2)^2) with b ~ U[0,1) (seed 1001)
  chebyshev6.npz  Chebyshev nodes of order 6 and S_n(x_j, i) on 11 points

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def c1_mini():
    n = 2000
    pos = np.random.default_rng(1000).random((n, 3))
    c = oracle.Cells([0, 0, 0], [1, 1, 1], [0, 0, 0], 6.0)
    c.update(pos.copy())
    r = float(c.side[0])
    row_ptr, col = c.search(c.positions, r)
    b = np.random.default_rng(1001).random(n)
    y = c.sparse_matvec(c.positions, r, (1, r / 2), b)
    return dict(positions=pos, radius=np.float64(r), grid=np.array(c.grid, np.int64),
                order=c.order, keys=c.keys, begin=c.begin, end=c.end, row_ptr=row_ptr, col=col,
                b=b, y=y)


def chebyshev6():
    x = np.linspace(-1.0, 1.0, 11)
    nodes = np.array([oracle.cheb_node(i, 6) for i in range(6)])
    sn = np.array([[oracle.cheb_Sn(v, i, 6) for v in x] for i in range(6)])
    return dict(x=x, nodes=nodes, sn=sn)


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "c1_mini.npz"), **c1_mini())
    np.savez_compressed(os.path.join(HERE, "chebyshev6.npz"), **chebyshev6())
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
