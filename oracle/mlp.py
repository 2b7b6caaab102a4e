"""FP64 numpy restatement of the learned preconditioner's network.

TEST INFRASTRUCTURE ONLY: the checker of the GPU forward / backward kernels.
Network: n_in -> hidden (SiLU) ... -> n_out (linear).  Parameters are one
flat array, layer by layer: weight matrix (out x in, row-major), then bias.
Loss: mean over rows and outputs of (net(x) - t)^2.
"""
import numpy as np


def layout(dims):
    """[(w_offset, b_offset, n_in, n_out)] per layer and the parameter count."""
    out, off = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        out.append((off, off + a * b, a, b))
        off += a * b + b
    return out, off


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(dims, params, x, keep=False):
    lay, _ = layout(dims)
    p = np.asarray(params, np.float64)
    a = np.asarray(x, np.float64)
    acts, zs = [a], []
    for li, (wo, bo, ni, no) in enumerate(lay):
        W = p[wo:wo + ni * no].reshape(no, ni)
        z = a @ W.T + p[bo:bo + no]
        if li < len(lay) - 1:
            zs.append(z)
            a = silu(z)
            acts.append(a)
        else:
            a = z
    return (a, acts, zs) if keep else a


def loss_grad(dims, params, x, t):
    """(loss, gradient) of the mean squared error, by hand-written backward."""
    lay, npar = layout(dims)
    p = np.asarray(params, np.float64)
    y, acts, zs = forward(dims, p, x, keep=True)
    t = np.asarray(t, np.float64)
    m = y.size
    loss = float(np.sum((y - t) ** 2) / m)
    g = np.zeros(npar)
    d = 2.0 * (y - t) / m
    for li in range(len(lay) - 1, -1, -1):
        wo, bo, ni, no = lay[li]
        if li < len(lay) - 1:
            d = d * silu_grad(zs[li])
        g[wo:wo + ni * no] = (d.T @ acts[li]).reshape(-1)
        g[bo:bo + no] = d.sum(axis=0)
        if li > 0:
            W = p[wo:wo + ni * no].reshape(no, ni)
            d = d @ W
    return loss, g


def loss(dims, params, x, t):
    y = forward(dims, params, x)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def loss_grad_residual(dims, params, x, mats):
    """(loss, gradient) of the patch-residual loss: mean over rows of
    |M_r y_r - e_1|^2 with y = net(x) and M_r (n_out x n_out) per row."""
    lay, npar = layout(dims)
    p = np.asarray(params, np.float64)
    y, acts, zs = forward(dims, p, x, keep=True)
    M = np.asarray(mats, np.float64)
    rows = y.shape[0]
    res = np.einsum("rab,rb->ra", M, y)
    res[:, 0] -= 1.0
    loss = float(np.sum(res ** 2) / rows)
    g = np.zeros(npar)
    d = 2.0 * np.einsum("rab,ra->rb", M, res) / rows
    for li in range(len(lay) - 1, -1, -1):
        wo, bo, ni, no = lay[li]
        if li < len(lay) - 1:
            d = d * silu_grad(zs[li])
        g[wo:wo + ni * no] = (d.T @ acts[li]).reshape(-1)
        g[bo:bo + no] = d.sum(axis=0)
        if li > 0:
            W = p[wo:wo + ni * no].reshape(no, ni)
            d = d @ W
    return loss, g


def loss_residual(dims, params, x, mats):
    y = forward(dims, params, x)
    res = np.einsum("rab,rb->ra", np.asarray(mats, np.float64), y)
    res[:, 0] -= 1.0
    return float(np.sum(res ** 2) / y.shape[0])
