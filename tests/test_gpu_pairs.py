"""GPU bucket-pair traversal (get_neighbouring_buckets / bucket_pair_iterator)
against the oracle's restatement of the reference's iterator, through the C
ABI.  Bar: the pair sequence (bucket i, bucket j, periodic image) identical
item by item, and the pair-loop neighbour counts bit-exact, on the
reference's fast-bucket-search configurations (tests/neighbours.h:1326-1370)
and on the 4,000,000-particle periodic mixture grid (100^3 buckets)."""
import numpy as np
import pytest
import torch

import aboria_amd as ab
import oracle
from aboria_amd import workloads as wl

pytestmark = pytest.mark.gpu

CASES = [(1, 14, 0.1, False), (1, 14, 0.1, True), (1, 1000, 0.1, True), (1, 1000, 0.1, False),
         (2, 1000, 0.1, True), (2, 1000, 0.1, False), (2, 1000, 0.5, True), (2, 1000, 0.5, False),
         (2, 1000, 0.2, True), (3, 1000, 0.2, True), (3, 1000, 0.2, False),
         (3, 20000, 0.15, True)]


@pytest.mark.parametrize("case", CASES, ids=["d%d_n%d_r%g_%s" % c for c in CASES])
def test_bucket_pairs_and_counts(ctx, case):
    dim, n, r, periodic = case
    print("pairs case", case, flush=True)
    rng = np.random.default_rng(n + dim + int(periodic))
    pos = rng.uniform(-1, 1, (n, dim))
    n_leaf = n * r ** dim / 2.0 ** dim
    cl = ab.CellListOrdered(ctx, [-1.0] * dim, [1.0] * dim, [periodic] * dim, n_leaf)
    cl.update_positions(torch.tensor(pos, device="cuda"))
    oc = oracle.Cells([-1.0] * dim, [1.0] * dim, [periodic] * dim, n_leaf)
    oc.update(pos.copy())
    assert cl.grid == oc.grid
    bi, bj, im = cl.bucket_pairs()
    obi, obj, oq = oc.bucket_pairs()
    assert np.array_equal(bi.cpu().numpy(), obi)
    assert np.array_equal(bj.cpu().numpy(), obj)
    assert np.array_equal(im.cpu().numpy().astype(np.int32), oq)
    got = cl.pair_counts(r).cpu().numpy()
    assert np.array_equal(got, oc.pair_counts(r))


def test_pair_counts_mixture_grid(ctx):
    """The periodic mixture workload's grid (100^3 buckets, r = 0.01 = the
    bucket side) at 400,000 particles: pair sequence and counts."""
    n = 400_000
    pos = wl.mixture_positions(n, seed=7)
    n_leaf = 3.95 * n / 4_000_000
    cl = ab.CellListOrdered(ctx, [0] * 3, [1] * 3, [True] * 3, n_leaf)
    cl.update_positions(torch.tensor(pos, device="cuda"))
    oc = oracle.Cells([0] * 3, [1] * 3, [True] * 3, n_leaf)
    oc.update(pos.copy())
    assert cl.grid == (100, 100, 100)
    bi, bj, im = cl.bucket_pairs()
    obi, obj, oq = oc.bucket_pairs()
    assert bi.numel() == len(obi)
    assert np.array_equal(bi.cpu().numpy(), obi) and np.array_equal(bj.cpu().numpy(), obj)
    assert np.array_equal(im.cpu().numpy().astype(np.int32), oq)
    assert np.array_equal(cl.pair_counts(0.01).cpu().numpy(), oc.pair_counts(0.01))


def test_pair_counts_rejects_small_buckets(ctx):
    pos = torch.rand((1000, 2), dtype=torch.float64, device="cuda")
    cl = ab.CellListOrdered(ctx, [0, 0], [1, 1], [False, False], 10.0)
    cl.update_positions(pos)
    with pytest.raises(ab.AbhError):
        cl.pair_counts(0.5)
