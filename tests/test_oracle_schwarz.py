"""The CPU oracle of the Schwarz preconditioner, the k-nearest-neighbour
pattern, the least-squares approximate-inverse rows and the network, pinned
against independent numpy computations (dense solves, brute force, lstsq,
finite differences)."""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp

GAUSS = 1


def gaussian_matrix(pos, sigma, radius, shift):
    d2 = ((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    a = np.exp(-(d2 / sigma ** 2))
    if radius > 0:
        a[d2 > radius * radius] = 0.0
    return a + shift * np.eye(len(pos))


def jittered(n1, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n1
    g = (np.arange(n1) + 0.5) * h
    x, y = np.meshgrid(g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel()], 1) + rng.uniform(-0.3 * h, 0.3 * h, (n1 * n1, 2))
    return p, h


def test_mlp_oracle_gradient_finite_differences():
    rng = np.random.default_rng(3)
    dims = [6, 8, 8, 3]
    _, npar = omlp.layout(dims)
    p = rng.normal(0, 0.5, npar)
    x = rng.normal(0, 1, (5, 6))
    t = rng.normal(0, 1, (5, 3))
    loss, g = omlp.loss_grad(dims, p, x, t)
    assert abs(loss - omlp.loss(dims, p, x, t)) < 1e-15
    h = 1e-6
    fd = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (omlp.loss(dims, p + e, x, t) - omlp.loss(dims, p - e, x, t)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-8 * max(1.0, np.max(np.abs(g)))
    # a dropped SiLU derivative would not pass
    assert np.max(np.abs(g)) > 1e-3


def test_mlp_oracle_residual_gradient_finite_differences():
    """The patch-residual loss |M y - e_1|^2: its hand-written gradient
    against central differences, and its value at a known answer (y solving
    M y = e_1 exactly gives 0)."""
    rng = np.random.default_rng(4)
    dims = [6, 8, 8, 3]
    _, npar = omlp.layout(dims)
    p = rng.normal(0, 0.5, npar)
    x = rng.normal(0, 1, (5, 6))
    M = rng.normal(0, 1, (5, 3, 3))
    loss, g = omlp.loss_grad_residual(dims, p, x, M)
    assert abs(loss - omlp.loss_residual(dims, p, x, M)) < 1e-14
    h = 1e-6
    fd = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (omlp.loss_residual(dims, p + e, x, M) -
                 omlp.loss_residual(dims, p - e, x, M)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))
    # a transposed M in the gradient would not pass (M is not symmetric here)
    assert np.max(np.abs(g)) > 1e-3
    # zero residual: M_r = diag(1 / y_r0, 0, 0) makes M_r y_r = e_1
    y = omlp.forward(dims, p, x)
    Mz = np.zeros((5, 3, 3))
    Mz[:, 0, 0] = 1.0 / y[:, 0]
    assert omlp.loss_residual(dims, p, x, Mz) < 1e-24


def test_schwarz_single_domain_is_the_exact_inverse():
    rng = np.random.default_rng(5)
    pos = rng.random((300, 2))
    c = oracle.Cells([0, 0], [1, 1], [0, 0], 1000.0)  # one bucket
    c.update(pos)
    s = c.schwarz((GAUSS, 0.05), -1.0, 1e-3)
    assert s.n_domains == 1 and s.n_interior[0] == 300
    A = gaussian_matrix(c.positions, 0.05, -1.0, 1e-3)
    b = rng.normal(size=300)
    x = s.apply(b)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-9


def block_domains(c):
    """Domains by the 3^dim block rule, computed independently in numpy."""
    grid = c.grid
    side = c.side
    offs, idx, nint = [0], [], []
    for b in range(int(np.prod(grid))):
        if c.end[b] <= c.begin[b]:
            continue
        iv = np.unravel_index(b, grid)
        bmin = np.array(iv) * side
        bmax = (np.array(iv) + 1) * side
        inner, buf = [], []
        ranges = [range(max(v - 1, 0), min(v + 1, g - 1) + 1) for v, g in zip(iv, grid)]
        for nb in np.ndindex(*[len(r) for r in ranges]):
            cell = tuple(r[k] for r, k in zip(ranges, nb))
            bi = np.ravel_multi_index(cell, grid)
            for p in range(c.begin[bi], c.end[bi]):
                x = c.positions[p]
                (inner if np.all(x >= bmin) and np.all(x < bmax) else buf).append(p)
        if not inner:
            continue
        idx += inner + buf
        nint.append(len(inner))
        offs.append(len(idx))
    return np.array(offs), np.array(idx, np.int32), np.array(nint, np.int32)


@pytest.mark.parametrize("dim,n,leaf", [(2, 700, 40.0), (3, 900, 50.0), (1, 200, 15.0)])
def test_schwarz_domains_equal_the_block_rule(dim, n, leaf):
    rng = np.random.default_rng(dim)
    pos = rng.random((n, dim))
    c = oracle.Cells([0] * dim, [1] * dim, [0] * dim, leaf)
    c.update(pos)
    s = c.schwarz((GAUSS, 0.05), 0.1, 1e-3)
    off, idx, nint = block_domains(c)
    assert np.array_equal(s.offsets, off)
    assert np.array_equal(s.indices, idx)
    assert np.array_equal(s.n_interior, nint)
    # interiors partition the particles
    interior = np.concatenate([idx[off[d]:off[d] + nint[d]] for d in range(len(nint))])
    assert np.array_equal(np.sort(interior), np.arange(c.n_alive))


def test_schwarz_apply_equals_dense_local_solves():
    pos, h = jittered(24, 11)
    c = oracle.Cells([0, 0], [24 * h, 24 * h], [0, 0], 36.0)
    c.update(pos)
    sigma, radius, shift = 1.5 * h, 6 * 1.5 * h, 1e-4
    s = c.schwarz((GAUSS, sigma), radius, shift)
    assert s.failed == 0 and s.n_domains > 4
    A = gaussian_matrix(c.positions, sigma, radius, shift)
    b = np.random.default_rng(2).normal(size=c.n_alive)
    ref = b.copy()
    for d in range(s.n_domains):
        ids = s.indices[s.offsets[d]:s.offsets[d + 1]]
        sol = np.linalg.solve(A[np.ix_(ids, ids)], b[ids])
        ref[ids[:s.n_interior[d]]] = sol[:s.n_interior[d]]
    x = s.apply(b)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-10


M64 = (1 << 64) - 1


def splitmix64(state):
    """SplitMix64 (Steele, Lea, Flood 2014): (next state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def test_splitmix64_known_answer():
    # the generator's published first outputs from state 0
    st, a = splitmix64(0)
    st, b = splitmix64(st)
    assert (a, b) == (0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4)


def coarse_sample(c, n, seed):
    """The coarse-grid sample (set_coarse_grid_n, src/Preconditioners.h:
    1230-1283) with the seeded sampler: lattice points counted per bucket
    in numpy, buckets ascending, a forward Fisher-Yates shuffle per bucket."""
    dim = len(c.grid)
    lo, hi = c.low, c.high
    dx = (hi - lo) / n - np.finfo(float).eps
    pts = np.array(np.meshgrid(*[np.arange(n + 1)] * dim, indexing="ij")).reshape(dim, -1).T
    p = pts * dx + lo
    v = np.floor((p - lo) * (1.0 / np.asarray(c.side))).astype(np.int64)
    v = np.clip(v, 0, np.asarray(c.grid) - 1)
    counts = np.bincount(np.ravel_multi_index(v.T, c.grid), minlength=int(np.prod(c.grid)))
    out = []
    for b in range(len(counts)):
        m = int(c.end[b]) - int(c.begin[b])
        if counts[b] == 0 or m <= 0:
            continue
        perm = list(range(m))
        st = seed ^ ((b * 0xD1B54A32D192ED03) & M64)
        for i in range(1, m):
            st, u = splitmix64(st)
            j = u % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        out += [int(c.begin[b]) + q for q in perm[:min(int(counts[b]), m)]]
    return np.array(out, np.int32), counts


@pytest.mark.parametrize("dim,n", [(2, 8), (2, 40), (3, 5), (1, 30)])
def test_schwarz_coarse_grid(dim, n):
    """The coarse-grid domain: the last domain, all interior; its particles
    the seeded sample (distinct, min(points, particles) from every bucket
    the lattice reaches); the apply equals the bucket domains' apply except
    on the coarse particles, where it is the dense solve of their block,
    written last."""
    rng = np.random.default_rng(40 + dim + n)
    npts = {1: 300, 2: 900, 3: 1200}[dim]
    pos = rng.random((npts, dim))
    c = oracle.Cells([0] * dim, [1] * dim, [0] * dim, 12.0)
    c.update(pos)
    sigma, radius, shift = 0.08, 0.3, 1e-3
    s0 = c.schwarz((GAUSS, sigma), radius, shift)
    s = c.schwarz((GAUSS, sigma), radius, shift, coarse_grid_n=n, seed=7)
    assert s.n_domains == s0.n_domains + 1
    assert np.array_equal(s.offsets[:-1], s0.offsets)
    ids = s.indices[s.offsets[-2]:s.offsets[-1]]
    assert s.n_interior[-1] == len(ids)
    want, counts = coarse_sample(c, n, 7)
    assert np.array_equal(ids, want)
    assert len(np.unique(ids)) == len(ids)
    sizes = np.array(c.end, np.int64) - np.array(c.begin, np.int64)
    assert len(ids) == int(np.minimum(counts, sizes).sum())
    other = c.schwarz((GAUSS, sigma), radius, shift, coarse_grid_n=n, seed=8)
    if len(ids) < c.n_alive:
        assert not np.array_equal(other.indices[other.offsets[-2]:], ids)
    A = gaussian_matrix(c.positions, sigma, radius, shift)
    b = rng.normal(size=c.n_alive)
    want_x = s0.apply(b)
    want_x[ids] = np.linalg.solve(A[np.ix_(ids, ids)], b[ids])
    x = s.apply(b)
    assert np.linalg.norm(x - want_x) / np.linalg.norm(want_x) < 1e-10


@pytest.mark.parametrize("dim,k", [(2, 16), (3, 32), (2, 1)])
def test_knn_oracle_brute_force(dim, k):
    rng = np.random.default_rng(10 + k)
    n = 600
    pos = rng.random((n, dim))
    c = oracle.Cells([0] * dim, [1] * dim, [0] * dim, 4.0)
    c.update(pos)
    nbr, d2 = c.knn(k)
    P = c.positions
    for i in range(0, n, 7):
        dd = ((P - P[i]) ** 2).sum(1)
        others = np.array([j for j in range(n) if j != i])
        order = np.lexsort((others, dd[others]))[: k - 1]
        assert nbr[i, 0] == i and d2[i, 0] == 0.0
        assert np.array_equal(nbr[i, 1:], others[order])
        assert np.array_equal(d2[i, 1:], dd[others[order]])


def test_lsq_rows_oracle_vs_lstsq():
    pos, h = jittered(14, 4)
    c = oracle.Cells([0, 0], [14 * h, 14 * h], [0, 0], 9.0)
    c.update(pos)
    sigma, radius, shift, k = 1.2 * h, 5 * 1.2 * h, 1e-3, 8
    nbr, _ = c.knn(k)
    w = c.lsq_rows((GAUSS, sigma), radius, shift, nbr)
    A = gaussian_matrix(c.positions, sigma, radius, shift)
    for i in range(0, c.n_alive, 9):
        e = np.zeros(c.n_alive)
        e[i] = 1.0
        ref, *_ = np.linalg.lstsq(A[:, nbr[i]], e, rcond=None)
        assert np.linalg.norm(w[i] - ref) / np.linalg.norm(ref) < 1e-7
