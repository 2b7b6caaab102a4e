"""Restricted additive Schwarz over a batch of particle sets, restated on the
CPU with numpy (one SchwartzPreconditioner per problem of a dense batch).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; never by the
product package.

Follows src/Preconditioners.h:1169-1230 (domains: one per non-empty bucket,
interior = the bucket's particles, buffer = the particles of the other
buckets of the 3^dim block around it), src/CellListOrdered.h:132-186 (the
bucket grid sized from n_particles_in_leaf and the particle count) and
src/Preconditioners.h:1476-1522 (apply: x = b, then each domain overwrites
its interior rows with its local solution).  The local systems are solved by
LU with partial pivoting (numpy.linalg.solve), an algorithm independent of
the product's signed Cholesky factorisation.
"""
import math

import numpy as np


def grid(n, n_leaf, low, high):
    """(size per dimension, side per dimension) of the bucket grid that
    CellListOrdered::set_domain_impl sizes for n particles
    (src/CellListOrdered.h:132-186)."""
    dim = len(low)
    low = np.asarray(low, np.float64)
    high = np.asarray(high, np.float64)
    if n_leaf > n:
        size = [1] * dim
    else:
        volume = float(np.prod(high - low))
        box_side = (n_leaf / n * volume) ** (1.0 / dim)
        size = [max(int(math.floor((high[d] - low[d]) / box_side)), 1) for d in range(dim)]
    side = [(high[d] - low[d]) / size[d] for d in range(dim)]
    return size, side


def bucket_keys(pos, size, side, low):
    """Row-major bucket key (last dimension fastest) of every particle,
    floor((x - low) / side) clamped to the grid."""
    key = np.zeros(pos.shape[0], np.int64)
    for d in range(pos.shape[1]):
        c = np.floor((pos[:, d] - low[d]) / side[d]).astype(np.int64)
        c = np.minimum(c, size[d] - 1)
        key = key * size[d] + c
    return key


def domains(off, pos, low, high, n_leaf):
    """(offsets, indices, n_interior): domains of every problem in bucket
    order, the interior (ascending particle index) first, then the buffer
    bucket by bucket (row-major over the 3^dim block, each ascending);
    indices are global."""
    dim = pos.shape[1]
    offs, idx, nint = [0], [], []
    for p in range(len(off) - 1):
        p0, p1 = int(off[p]), int(off[p + 1])
        if p1 == p0:
            continue
        size, side = grid(p1 - p0, n_leaf, low, high)
        key = bucket_keys(pos[p0:p1], size, side, low)
        members = {}
        for i in np.argsort(key, kind="stable"):
            members.setdefault(int(key[i]), []).append(p0 + int(i))
        sz3 = list(size) + [1] * (3 - dim)
        for b in sorted(members):
            iv, r = [0, 0, 0], b
            for d in range(dim - 1, -1, -1):
                iv[d] = r % sz3[d]
                r //= sz3[d]
            lst = list(members[b])
            rng = [range(max(iv[d] - 1, 0), min(iv[d] + 1, sz3[d] - 1) + 1) for d in range(3)]
            for x0 in rng[0]:
                for x1 in rng[1]:
                    for x2 in rng[2]:
                        nbk = (x0 * sz3[1] + x1) * sz3[2] + x2
                        if nbk != b:
                            lst.extend(members.get(nbk, []))
            idx.extend(lst)
            nint.append(len(members[b]))
            offs.append(len(idx))
    return (np.asarray(offs, np.int64), np.asarray(idx, np.int32),
            np.asarray(nint, np.int32))


def phi(kernel, s):
    """Radial profile of the squared distance s (kernel = (type, param)),
    the same profiles as the C oracle's kernels.hpp."""
    t, c = kernel
    if t == 0:
        return np.ones_like(s)
    if t == 1:
        return np.exp(-(s / (c * c)))
    if t == 2:
        q = np.sqrt(s) / c
        return np.where(q < 1.0, (1.0 - q) ** 4 * (4.0 * q + 1.0), 0.0)
    if t == 3:
        return np.sqrt(s + c * c)
    if t == 4:
        return 1.0 / (np.sqrt(s) + c)
    raise ValueError("unknown kernel type")


def local_matrix(pos, ids, kernel, shift):
    p = pos[ids]
    s = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    return phi(kernel, s) + shift * np.eye(len(ids))


def apply(pos, doms, kernel, shift, b):
    """x = M^-1 b (src/Preconditioners.h:1476-1522)."""
    offs, idx, nint = doms
    x = np.array(b, np.float64, copy=True)
    for d in range(len(nint)):
        ids = idx[offs[d]:offs[d + 1]]
        y = np.linalg.solve(local_matrix(pos, ids, kernel, shift), b[ids])
        x[ids[:nint[d]]] = y[:nint[d]]
    return x
