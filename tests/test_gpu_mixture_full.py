"""Neighbour lists of the full N=4,000,000 periodic Gaussian-mixture
workload (64 wrapped Gaussians, sigma 0.02, radius 0.01, 100^3 buckets, some
buckets holding hundreds of particles, centres with thousands of
neighbours) against the oracle at that size, so that the kernels' crowded
bucket paths (LDS ring staging of long rows, rows spanning many 64-byte
segments, the stencil widened for centres near a bucket face) are checked
where they run.

Bar (bit-exact, through the C ABI):
* the build: gather order, bucket indices, bucket begin/end, positions;
* row_ptr of every centre equal to the oracle's counts;
* every row's multiset of neighbour indices equal to the oracle's, checked
  by an order-independent digest computed on both sides (count, sum of
  indices, sum of squared indices, sum of a 64-bit mix of each index -- an
  extra, missing or substituted pair changes the last two);
* the 2,000 longest rows compared exactly after sorting;
* the single-traversal search writing the same row_ptr and lists.
"""
import numpy as np
import pytest
import torch

import aboria_amd as ab
import oracle
from aboria_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _s64(c):
    """an unsigned 64-bit constant as the signed int64 with the same bits"""
    return c - (1 << 64) if c >= (1 << 63) else c


def _lsr(x, k):
    """logical right shift of int64 bit patterns"""
    return (x >> k) & ((1 << (64 - k)) - 1)


def _mix(x):
    """SplitMix64 finaliser of x + golden gamma, wrapping int64 arithmetic
    (the oracle's digest_mix)"""
    x = x + _s64(0x9E3779B97F4A7C15)
    x = (x ^ _lsr(x, 30)) * _s64(0xBF58476D1CE4E5B9)
    x = (x ^ _lsr(x, 27)) * _s64(0x94D049BB133111EB)
    return x ^ _lsr(x, 31)


def gpu_digest(rp, col, chunk=1 << 27):
    """(nq, 4) int64 digest of the CSR lists held on the device, computed in
    row ranges of at most `chunk` entries by prefix sums (wrapping int64)."""
    nq = rp.numel() - 1
    dev = rp.device
    out = torch.zeros((nq, 4), dtype=torch.int64, device=dev)
    out[:, 0] = rp[1:] - rp[:-1]
    rp_h = rp.cpu().numpy()
    zero = torch.zeros(1, dtype=torch.int64, device=dev)
    r0 = 0
    while r0 < nq:
        r1 = int(np.searchsorted(rp_h, rp_h[r0] + chunk, side="right")) - 1
        r1 = min(max(r1, r0 + 1), nq)
        a, b = int(rp_h[r0]), int(rp_h[r1])
        off = rp[r0:r1 + 1] - a
        v = col[a:b].to(torch.int64)
        for c, vals in ((1, v), (2, v * v), (3, _mix(v))):
            cs = torch.cat([zero, torch.cumsum(vals, 0)])
            out[r0:r1, c] = cs[off[1:]] - cs[off[:-1]]
        del v
        r0 = r1
    return out


def test_mixture_full_density_lists(ctx):
    c = wl.MIXTURE
    n, r = c["n"], c["radius"]
    print("mixture N=%d r=%g: build" % (n, r), flush=True)
    pos = wl.mixture_positions(n, c["n_centres"], c["sigma"], seed=c["seed"])
    gp = torch.tensor(pos, device="cuda")
    cl = ab.CellListOrdered(ctx, c["low"], c["high"], c["periodic"], c["n_leaf"])
    assert cl.update_positions(gp) == n
    oc = oracle.Cells(c["low"], c["high"], c["periodic"], c["n_leaf"])
    assert oc.update(pos.copy()) == n
    assert cl.grid == oc.grid == (100, 100, 100)
    assert np.array_equal(cl.order().cpu().numpy(), oc.order)
    assert np.array_equal(cl.bucket_index().cpu().numpy().view(np.uint32), oc.keys)
    assert np.array_equal(cl.bucket_begin().cpu().numpy().view(np.uint32), oc.begin)
    assert np.array_equal(cl.bucket_end().cpu().numpy().view(np.uint32), oc.end)
    assert np.array_equal(cl.positions().cpu().numpy(), oc.positions)
    occupancy = (oc.end.astype(np.int64) - oc.begin).max()
    print("fullest bucket %d particles" % occupancy, flush=True)
    assert occupancy > 400  # the crowded-bucket regime of this workload

    print("search (count + fill)", flush=True)
    sp = cl.positions()
    s = ab.Search(cl, sp, r)
    total = s.total
    rp_g = s.row_ptr().clone()
    col = torch.empty(total, dtype=torch.int32, device="cuda")
    s.fill(col)
    torch.cuda.synchronize()

    print("oracle digest of %d rows" % n, flush=True)
    dg_o = oc.search_digest(oc.positions, r)
    rp_o = np.zeros(n + 1, np.int64)
    np.cumsum(dg_o[:, 0], out=rp_o[1:])
    assert total == int(rp_o[-1])
    assert np.array_equal(rp_g.cpu().numpy(), rp_o)
    longest = int(dg_o[:, 0].max())
    print("pairs %d, longest row %d" % (total, longest), flush=True)
    assert longest > 2000

    dg_g = gpu_digest(rp_g, col).cpu().numpy()
    bad = np.nonzero((dg_g != dg_o).any(axis=1))[0]
    assert bad.size == 0, "rows differ: %s" % bad[:10]

    print("2,000 longest rows sorted", flush=True)
    rows = np.argsort(-dg_o[:, 0], kind="stable")[:2000]
    rp_s, col_s = oc.search(oc.positions[rows], r)
    for k, i in enumerate(rows):
        g = col[int(rp_o[i]):int(rp_o[i + 1])].cpu().numpy()
        o = col_s[rp_s[k]:rp_s[k + 1]]
        assert np.array_equal(np.sort(g), np.sort(o)), int(i)

    print("single traversal", flush=True)
    col2 = torch.full((total + 16,), -7, dtype=torch.int32, device="cuda")
    t2 = s.run_fused(cl, sp, r, col2)
    assert t2 == total
    assert torch.equal(s.row_ptr(), rp_g)
    assert torch.equal(col2[:total], col)
    assert torch.all(col2[total:] == -7)
    del col, col2
    s.close()
    torch.cuda.empty_cache()
