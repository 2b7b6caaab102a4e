"""Seeded synthetic inputs for the particle hot-path workloads.

numpy's PCG64 generator is deterministic across machines, so this container
and the GPU host regenerate identical inputs from the seed alone.

  cells_uniform   3D, N=100,000 i.i.d. uniform in the unit cube, non-periodic,
                  cell side = radius = 0.04 (n_particles_in_leaf = 6 -> 25^3)
  cells_mixture   3D, N=4,000,000 periodic unit cube, mixture of 64 isotropic
                  Gaussians (sigma 0.02, wrapped), radius 0.01
                  (n_particles_in_leaf = 3.95 -> 100^3), Wendland C2, 8 RHS
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
