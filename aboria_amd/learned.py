"""Training pipeline of the learned preconditioner.

The network maps the relative positions of a particle's k nearest
neighbours (scaled by the kernel width) to its approximate-inverse row on
that pattern.  Training rows come from seeded point sets: for each, the cell
list is built on the GPU and the features are gathered on the GPU; the
network is then trained by the library's own forward / backward / Adam
kernels.  Two targets:

* patch (training_set_patch / train_residual): the local solves of every
  particle's patch -- its k nearest neighbours, the particle's row of
  (K_P + shift I)^{-1} -- which the network stands in for.  The loss is the
  patch residual |A_P w - e_1|^2, so the network is fitted where the patch
  operator weighs its error; a fit of the inverse's entries themselves
  (which alternate in sign and cancel) leaves errors the operator amplifies.
  The patch matrix is a function of the features alone.
* lsq (training_set / train): the least-squares approximate-inverse rows
  of A = K + shift I on the pattern (abh_ainv_create_lsq), fitted entry by
  entry (mean squared error).
"""
import numpy as np
import torch

from . import ApproxInverse, CellListOrdered, Mlp


def training_set(ctx, point_sets, kernel, radius, shift, n_leaf, k=32, feature_scale=1.0,
                 max_rows=None, seed=0, low=None, high=None):
    """(features (m, k*dim) FP32, targets (m, k) FP64) from a list of (n, dim)
    numpy point sets, subsampled to max_rows rows (seeded)."""
    xs, ts = [], []
    for pts in point_sets:
        pts = np.asarray(pts, np.float64)
        dim = pts.shape[1]
        lo = low if low is not None else [0.0] * dim
        hi = high if high is not None else [1.0] * dim
        cells = CellListOrdered(ctx, lo, hi, [False] * dim, n_leaf)
        p = torch.tensor(pts, device=ctx.device)
        cells.update_positions(p)
        ai = ApproxInverse.lsq(ctx, cells, kernel, radius, shift, k)
        xs.append(ai.features(feature_scale))
        ts.append(ai.weights())
        ai.close()
        cells.close()
    X = torch.cat(xs)
    T = torch.cat(ts)
    if max_rows is not None and X.shape[0] > max_rows:
        g = torch.Generator().manual_seed(int(seed))
        sel = torch.randperm(X.shape[0], generator=g)[:max_rows].to(X.device)
        X, T = X[sel].contiguous(), T[sel].contiguous()
    return X, T


def train(ctx, X, T, hidden=128, n_hidden=3, epochs=30, batch=1024, lr=2e-3, lr_decay=0.93,
          seed=0, graph=False, fused=False):
    """Trains a network on (X, T); targets are divided by their RMS (returned
    as out_scale, the factor the network's outputs are multiplied by)."""
    out_scale = float(torch.sqrt(torch.mean(T * T)).item()) or 1.0
    Tn = (T / out_scale).to(torch.float32).contiguous()
    net = Mlp(ctx, X.shape[1], hidden, n_hidden, T.shape[1], seed=seed)
    losses = net.train(X.contiguous(), Tn, epochs=epochs, batch=batch, lr=lr, lr_decay=lr_decay,
                       seed=seed, graph=graph, fused=fused)
    return net, out_scale, losses


def training_set_patch(ctx, point_sets, kernel, radius, shift, n_leaf, k=32, feature_scale=1.0,
                       max_rows=None, seed=0, low=None, high=None):
    """(features (m, k*dim) FP32, patch matrices (m, k, k) FP32 times
    out_scale, exact patch rows (m, k) FP64, out_scale) from a list of
    (n, dim) numpy point sets, subsampled to max_rows rows (seeded);
    out_scale is the RMS of the exact rows, the factor the network's outputs
    are multiplied by."""
    xs, ms, ws = [], [], []
    for pts in point_sets:
        pts = np.asarray(pts, np.float64)
        dim = pts.shape[1]
        lo = low if low is not None else [0.0] * dim
        hi = high if high is not None else [1.0] * dim
        cells = CellListOrdered(ctx, lo, hi, [False] * dim, n_leaf)
        p = torch.tensor(pts, device=ctx.device)
        cells.update_positions(p)
        ai = ApproxInverse.patch(ctx, cells, kernel, radius, shift, k)
        xs.append(ai.features(feature_scale))
        ms.append(ai.patch_matrices(1.0))
        ws.append(ai.weights())
        ai.close()
        cells.close()
    X, Mt, W = torch.cat(xs), torch.cat(ms), torch.cat(ws)
    if max_rows is not None and X.shape[0] > max_rows:
        g = torch.Generator().manual_seed(int(seed))
        sel = torch.randperm(X.shape[0], generator=g)[:max_rows].to(X.device)
        X, Mt, W = X[sel].contiguous(), Mt[sel].contiguous(), W[sel].contiguous()
    out_scale = float(torch.sqrt(torch.mean(W * W)).item()) or 1.0
    Mt.mul_(out_scale)
    return X, Mt, W, out_scale


def train_residual(ctx, X, mats, hidden=128, n_hidden=3, epochs=30, batch=512, lr=2e-3,
                   lr_decay=0.93, seed=0, graph=False, fused=False):
    """Trains a network on the patch-residual loss (mats already scaled by
    out_scale); returns (net, per-epoch losses)."""
    k = mats.shape[1]
    net = Mlp(ctx, X.shape[1], hidden, n_hidden, k, seed=seed)
    losses = net.train_residual(X.contiguous(), mats.contiguous(), epochs=epochs, batch=batch,
                                lr=lr, lr_decay=lr_decay, seed=seed, graph=graph,
                                fused=fused)
    return net, losses
