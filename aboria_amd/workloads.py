"""Seeded synthetic inputs for the particle hot-path workloads.

numpy's PCG64 generator is deterministic across machines, so this container
and the GPU host regenerate identical inputs from the seed alone.

  cells_uniform   3D, N=100,000 i.i.d. uniform in the unit cube, non-periodic,
                  cell side = radius = 0.04 (n_particles_in_leaf = 6 -> 25^3)
  cells_mixture   3D, N=4,000,000 periodic unit cube, mixture of 64 isotropic
                  Gaussians (sigma 0.02, wrapped), radius 0.01
                  (n_particles_in_leaf = 3.95 -> 100^3), Wendland C2, 8 RHS
  fmm_uniform     3D, N=2,000,000 uniform, kernel 1/(r + 1e-3), leaves of at
                  most 64 particles, Chebyshev order 6, 16 RHS ~ U[-1, 1)
  rbf_batch       512 2D problems, N_k = floor(exp(U[ln 1e3, ln 6.4e4))),
                  positions U[0,1)^2, multiquadric c = 0.05, targets
                  sin(2 pi x) cos(2 pi y) + N(0, 1e-6) (variance), GMRES(50)
  schwarz_grid    2D, 632 x 633 = 400,056 points on a grid jittered by
                  U[-0.3h, 0.3h] per axis, Gaussian exp(-|dx|^2 / 0.01^2)
                  truncated at 6 sigma, diagonal shift 1e-5 (the reference's
                  self-kernel jitter), RHS ~ N(0, 1), GMRES to 1e-8
"""
import numpy as np

UNIFORM = dict(n=100_000, radius=0.04, n_leaf=6.0, periodic=(False, False, False),
               low=(0.0, 0.0, 0.0), high=(1.0, 1.0, 1.0), gauss_scale=0.02, seed=1000)
MIXTURE = dict(n=4_000_000, radius=0.01, n_leaf=3.95, periodic=(True, True, True),
               low=(0.0, 0.0, 0.0), high=(1.0, 1.0, 1.0), n_centres=64, sigma=0.02, nrhs=8,
               seed=2000)


def uniform_positions(n, dim=3, seed=1000):
    rng = np.random.default_rng(seed)
    return rng.random((n, dim))


def mixture_positions(n, n_centres=64, sigma=0.02, dim=3, seed=2000):
    """Gaussian mixture wrapped into [0,1)^dim (x - floor(x), 1.0 -> 0.0)."""
    rng = np.random.default_rng(seed)
    centres = rng.random((n_centres, dim))
    labels = rng.integers(0, n_centres, size=n)
    pos = centres[labels] + rng.normal(0.0, sigma, size=(n, dim))
    pos -= np.floor(pos)
    pos[pos >= 1.0] = 0.0
    return pos


def uniform_rhs(n, nrhs=1, seed=1001):
    rng = np.random.default_rng(seed)
    return rng.random((n, nrhs))


def normal_rhs(n, nrhs=8, seed=2001):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, nrhs))


FMM = dict(n=2_000_000, n_leaf=64, order=6, eps=1e-3, nrhs=16, n_check=10_000, seed=3000)
RBF_BATCH = dict(nprob=512, n_min=1000, n_max=64000, c=0.05, restart=50, tol=1e-8, seed=4000,
                 schwarz_n_leaf=64.0)
SCHWARZ = dict(nx=632, ny=633, jitter=0.3, sigma=0.01, cutoff=6.0, shift=1e-5, n_leaf=576.0,
               k=32, seed=5000)


def fmm_positions(n, seed=3000):
    return np.random.default_rng(seed).random((n, 3))


def fmm_rhs(n, nrhs=16, seed=3001):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, (nrhs, n))


def rbf_batch_problems(nprob=512, n_min=1000, n_max=64000, seed=4000):
    """(offsets, positions (total, 2), targets (total,)) of the batch."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max), nprob))).astype(np.int64)
    off = np.zeros(nprob + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    pos = rng.random((int(off[-1]), 2))
    f = np.sin(2 * np.pi * pos[:, 0]) * np.cos(2 * np.pi * pos[:, 1])
    f += rng.normal(0.0, 1e-3, f.shape)
    return off, pos, f


def jittered_grid(nx, ny, jitter=0.3, seed=5000):
    """nx x ny points on the cell centres of [0,1)^2, each moved by
    U[-jitter h, jitter h] per axis (h = 1/nx, 1/ny)."""
    rng = np.random.default_rng(seed)
    hx, hy = 1.0 / nx, 1.0 / ny
    gx = (np.arange(nx) + 0.5) * hx
    gy = (np.arange(ny) + 0.5) * hy
    x, y = np.meshgrid(gx, gy, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel()], 1)
    pos[:, 0] += rng.uniform(-jitter * hx, jitter * hx, pos.shape[0])
    pos[:, 1] += rng.uniform(-jitter * hy, jitter * hy, pos.shape[0])
    return pos
