"""Restarted GMRES(m) restated as what the mathematics fixes, for the GPU
solver tests: per cycle of k steps from x, the update is the d in the Krylov
space K_k(A M^-1, r), r = b - A x, that minimises ||r - A M^-1 d||, and x
moves by M^-1 d (M^-1 = I without a preconditioner).  An orthonormal basis
of the Krylov space (Gram-Schmidt twice per vector) and a dense
least-squares solve: no Arnoldi recurrence or Givens rotation in common with
the solver under test."""
import numpy as np


def krylov_step(A, r, k):
    """argmin over d in span{r, A r, ..., A^(k-1) r} of ||r - A d||."""
    Q = np.zeros((len(r), k))
    v = r / np.linalg.norm(r)
    for j in range(k):
        for _ in range(2):
            v = v - Q[:, :j] @ (Q[:, :j].T @ v)
        Q[:, j] = v / np.linalg.norm(v)
        v = A @ Q[:, j]
    y, *_ = np.linalg.lstsq(A @ Q, r, rcond=None)
    return Q @ y


def gmres_iterate(A, b, restart, max_iters, x0=None, minv=None):
    """The GMRES(restart) iterate after max_iters steps from x0 (cycles of
    `restart` steps and a shorter last one)."""
    x = np.zeros(len(b)) if x0 is None else np.array(x0, dtype=np.float64)
    op = A if minv is None else A @ minv
    left = max_iters
    while left > 0:
        d = krylov_step(op, b - A @ x, min(restart, left))
        x += d if minv is None else minv @ d
        left -= restart
    return x
