"""GPU black-box Chebyshev FMM against the CPU oracle (same tree, same
traversal) and against an FP64 direct sum.

Bar: the FP64 GPU FMM equals the oracle's FMM to 1e-11 relative in 1-3
dimensions (2-norm;
summation order and the translated M2L operators differ only by rounding);
the mixed form (BF16x3 MFMA M2L with FP32 accumulation; P2M, L2P and P2P
in FP64 on the matrix cores as in the FP64 form) is within 5e-5 relative of
the FP64 form; the order-3 multiquadric case meets the reference's own accuracy test
(tests/fmm.h:107-109: sum (y - y_direct)^2 / sum y_direct^2 < 1e-2).
"""
import numpy as np
import pytest
import torch

import aboria_amd as ab
import oracle

pytestmark = pytest.mark.gpu

KM, KI = 3, 4


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("order", [3, 4, 6, 7, 8])
def test_fmm_matches_oracle(ctx, order):
    print("fmm oracle order", order, flush=True)
    # orders 7 and 8: fewer, fuller leaves (three levels) keep the CPU
    # oracle to seconds
    n, leaf = (3000, 16) if order <= 6 else (2000, 40)
    rng = np.random.default_rng(40 + order)
    pos = rng.random((n, 3))
    q = rng.uniform(-1, 1, n)
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, leaf, order,
                       ab.Kernel.smooth_inv_r(1e-3))
    y = (f @ torch.tensor(q, device="cuda")).cpu().numpy()
    t = oracle.Octree([0, 0, 0], [1, 1, 1], pos, leaf)
    assert np.array_equal(f.order().cpu().numpy(), t.order)
    y_o_sorted = t.fmm_matvec(order, (KI, 1e-3), q[t.order])
    y_o = np.empty(n)
    y_o[t.order] = y_o_sorted
    assert rel(y, y_o) < 1e-11
    info = f.info
    assert info["n_leaves"] > 0 and info["m2l_pairs"] > 0


def test_fmm_reference_accuracy_order3(ctx):
    # tests/fmm.h:165-215 (3D): cosine source, sqrt(|dx|^2 + 0.01), leaf 10
    n = 1000
    rng = np.random.default_rng(7)
    pos = rng.random((n, 3))
    q = np.prod(np.cos(2 * np.pi * pos), axis=1) / n
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, 10, 3,
                       ab.Kernel.multiquadric(0.1))
    y = (f @ torch.tensor(q, device="cuda")).cpu().numpy()
    ref = oracle.dense_matvec(pos, pos, (KM, 0.1), q)
    assert np.sum((y - ref) ** 2) / np.sum(ref ** 2) < 1e-2


def test_fmm_multi_rhs_and_mixed(ctx):
    n = 20000
    rng = np.random.default_rng(9)
    pos = rng.random((n, 3))
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, 64, 6,
                       ab.Kernel.smooth_inv_r(1e-3))
    nrhs = 20  # two passes of 16
    x = rng.uniform(-1, 1, (nrhs, n))
    xt = torch.tensor(x, device="cuda")
    y64 = f.apply_colmajor(xt, torch.zeros_like(xt), ab.FP64).cpu().numpy()
    # each column equals a single-RHS apply
    for r in (0, 17):
        y1 = (f @ xt[r].contiguous()).cpu().numpy()
        assert rel(y64[r], y1) < 1e-13
    ymx = f.apply_colmajor(xt, torch.zeros_like(xt), ab.MIXED).cpu().numpy()
    err_mixed = rel(ymx, y64)
    print("mixed vs fp64 relative difference %.3e" % err_mixed, flush=True)
    assert err_mixed < 5e-5
    # accuracy against the direct sum on 500 targets
    idx = rng.choice(n, 500, replace=False)
    ref = oracle.dense_matvec(pos[idx], pos, (KI, 1e-3), x[0])
    assert rel(y64[0][idx], ref) < 1e-4


@pytest.mark.parametrize("order", [6, 7, 8])
def test_fmm_high_order_accuracy(ctx, order):
    """Error against the FP64 direct sum on 1,000 targets of 40,000 uniform
    points (the 1/(r + 1e-3) kernel, leaves of at most 64): it falls with the
    order, and order 8 is below 1e-6 -- the accuracy the FMM throughput is
    quoted at.  MIXED at orders above 6 keeps the M2L in FP64 (its BF16 split
    would floor the error near 1e-6)."""
    n = 40_000
    rng = np.random.default_rng(11)
    pos = rng.random((n, 3))
    x = rng.uniform(-1, 1, n)
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, 64, order,
                       ab.Kernel.smooth_inv_r(1e-3))
    xt = torch.tensor(x, device="cuda")
    y = (f @ xt).cpu().numpy()
    idx = rng.choice(n, 1000, replace=False)
    ref = oracle.dense_matvec(pos[idx], pos, (KI, 1e-3), x)
    err = rel(y[idx], ref)
    ym = f.apply_colmajor(xt[None, :].contiguous(), torch.zeros((1, n), dtype=torch.float64,
                                                               device="cuda"), ab.MIXED)
    errm = rel(ym.cpu().numpy()[0][idx], ref)
    print("order %d: fp64 err %.3e mixed err %.3e" % (order, err, errm), flush=True)
    bound = {6: 2e-5, 7: 5e-6, 8: 1e-6}[order]
    assert err < bound
    if order > 6:
        assert errm < bound


@pytest.mark.parametrize("order", [3, 4])
def test_fmm_compressed_full_rank_equals_uncompressed(ctx, order):
    """abh_fmm_compress with a tolerance at the rounding of the operators
    (1e-15) keeps every singular vector at orders 3 and 4 (rank = order^3,
    U U^T = I): the compressed apply (W^ = U^T W, C_o = U^T K_o U,
    G += U G^) equals the uncompressed one to 1e-11 relative -- the layouts
    and transposes of the three transforms are exact.  (At higher orders the
    stacked operators' singular values fall below what the FP64 Gram matrix
    resolves, about 1e-8 of the largest; the accuracy test below covers
    them at working tolerances.)"""
    print("fmm compressed full rank order", order, flush=True)
    n, leaf = 6000, 16
    rng = np.random.default_rng(60 + order)
    pos = rng.random((n, 3))
    x = rng.uniform(-1, 1, (3, n))
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, leaf, order,
                       ab.Kernel.smooth_inv_r(1e-3))
    xt = torch.tensor(x, device="cuda")
    y0 = f.apply_colmajor(xt, torch.zeros_like(xt)).cpu().numpy()
    kp, ranks = f.compress(1e-15)
    print("padded rank", kp, "ranks", ranks, flush=True)
    assert ranks and max(ranks) <= order ** 3 and kp >= max(ranks)
    assert min(ranks) == order ** 3  # every direction the operators use
    y1 = f.apply_colmajor(xt, torch.zeros_like(xt)).cpu().numpy()
    assert rel(y1, y0) < 1e-11
    kp0, _ = f.compress(0.0)
    assert kp0 == 0
    y2 = f.apply_colmajor(xt, torch.zeros_like(xt)).cpu().numpy()
    assert np.array_equal(y2, y0)


@pytest.mark.parametrize("order,eps,bound", [(6, 1e-7, 2e-5), (8, 1e-7, 1e-6)])
def test_fmm_compressed_accuracy(ctx, order, eps, bound):
    """The compressed operators at tolerance eps (ranks well below order^3)
    against the FP64 direct sum on 1,000 targets of 40,000 uniform points:
    within the uncompressed FMM's bound for the order, and within 20 eps
    (relative, 2-norm) of the uncompressed apply."""
    print("fmm compressed order %d eps %g" % (order, eps), flush=True)
    n = 40_000
    rng = np.random.default_rng(11)
    pos = rng.random((n, 3))
    x = rng.uniform(-1, 1, n)
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, 64, order,
                       ab.Kernel.smooth_inv_r(1e-3))
    xt = torch.tensor(x, device="cuda")
    y = (f @ xt).cpu().numpy()
    kp, ranks = f.compress(eps)
    yc = (f @ xt).cpu().numpy()
    idx = rng.choice(n, 1000, replace=False)
    ref = oracle.dense_matvec(pos[idx], pos, (KI, 1e-3), x)
    err, errc = rel(y[idx], ref), rel(yc[idx], ref)
    print("order %d eps %g: ranks %s (padded %d), err %.3e compressed %.3e, vs uncompressed "
          "%.3e" % (order, eps, ranks, kp, err, errc, rel(yc, y)), flush=True)
    assert kp < order ** 3
    assert errc < bound
    assert rel(yc, y) < 20 * eps


@pytest.mark.parametrize("order,eps", [(6, 1e-7), (8, 2e-7)])
def test_fmm_compressed_mixed_fp16(ctx, order, eps):
    """MIXED on compressed operators: the rank-k M2L on FP16 matrix cores
    (two-term split, power-of-two scaling per level, FP32 accumulation), 16
    right-hand sides on 40,000 uniform points:
    within 5e-7 (relative, 2-norm) of the FP64 compressed apply and within
    1e-6 of the FP64 direct sum at order 8; the M2L autotuner's two tile
    shapes (1 and 2 sibling tiles per block) give identical results."""
    print("fmm compressed mixed order %d eps %g" % (order, eps), flush=True)
    n = 40_000
    rng = np.random.default_rng(12)
    pos = rng.random((n, 3))
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, 64, order,
                       ab.Kernel.smooth_inv_r(1e-3))
    kp, ranks = f.compress(eps)
    x = rng.uniform(-1, 1, (16, n))
    xt = torch.tensor(x, device="cuda")
    y64 = f.apply_colmajor(xt, torch.zeros_like(xt), ab.FP64).cpu().numpy()
    ymx = f.apply_colmajor(xt, torch.zeros_like(xt), ab.MIXED).cpu().numpy()
    idx = rng.choice(n, 500, replace=False)
    ref = oracle.dense_matvec(pos[idx], pos, (KI, 1e-3), x[3])
    d = rel(ymx, y64)
    e64, emx = rel(y64[3][idx], ref), rel(ymx[3][idx], ref)
    print("order %d eps %g (padded rank %d): mixed vs fp64 compressed %.3e; vs direct: fp64 "
          "%.3e mixed %.3e" % (order, eps, kp, d, e64, emx), flush=True)
    assert kp <= 192  # the FP16 path's ranks
    assert d < 5e-7
    if order == 8:
        assert emx < 1e-6
    # the autotuner timed both tile shapes on the first MIXED apply; either
    # shape gives the same bits
    tt, ms = f.m2l_tuning()
    assert tt in (1, 2) and ms[0] > 0 and ms[1] > 0
    outs = []
    for shape in (1, 2):
        f.set_m2l_tuning(shape)
        outs.append(f.apply_colmajor(xt, torch.zeros_like(xt), ab.MIXED))
    assert torch.equal(outs[0], outs[1])
    assert np.array_equal(outs[tt - 1].cpu().numpy(), ymx)


def test_fmm_tree_leaf_rule(ctx):
    """The device-built tree keeps the inclusive leaf rule: on the 10-point
    known answer of tests/test_oracle_fmm.py (an octant of exactly
    n_leaf = 4 particles is one leaf, one of 5 is split) the GPU tree has the
    oracle's order and number of non-empty leaves, and its FMM equals the
    direct sum to 1e-11 (every pair is near: P2P only)."""
    octant0 = [[0.1, 0.1, 0.1], [0.2, 0.3, 0.1], [0.4, 0.2, 0.3], [0.3, 0.4, 0.4]]
    octant7 = [[0.6, 0.6, 0.6], [0.9, 0.9, 0.9], [0.6, 0.9, 0.6], [0.9, 0.6, 0.9],
               [0.7, 0.7, 0.8]]
    pos = np.array(octant0 + octant7 + [[0.1, 0.1, 0.9]], dtype=np.float64)
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0] * 3, [1] * 3, 4, 3,
                       ab.Kernel.smooth_inv_r(1e-3))
    t = oracle.Octree([0, 0, 0], [1, 1, 1], pos, 4)
    assert np.array_equal(f.order().cpu().numpy(), t.order)
    nonempty = int(((t.leaves[:, 1] - t.leaves[:, 0]) > 0).sum())
    assert f.info["n_leaves"] == nonempty
    q = np.linspace(-1.0, 1.0, len(pos))
    y = (f @ torch.tensor(q, device="cuda")).cpu().numpy()
    assert rel(y, oracle.dense_matvec(pos, pos, (KI, 1e-3), q)) < 1e-11


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("order", [3, 5, 8])
def test_fmm_lower_dims_match_oracle(ctx, dim, order):
    """1D and 2D (tests/fmm.h runs helper_fast_methods<D> for D = 1, 2, 3):
    the device tree's order and pair counts are the oracle's, the FP64 FMM
    equals the oracle's FMM to 1e-11 on two columns of a 20-column apply
    (crossing the 16-column pass), MIXED equals FP64 bit for bit (both
    compute in FP64 here) and the M2L compression is refused.  In 1D a
    cluster of 200 particles within 1e-6 drives leaves to about level 24."""
    print("fmm lower dims", dim, order, flush=True)
    n, leaf = (4000, 12) if dim == 1 else (3000, 16)
    rng = np.random.default_rng(60 + 10 * dim + order)
    pos = rng.random((n, dim))
    if dim == 1:
        pos[:200, 0] = 0.3 + 1e-6 * rng.random(200)
    X = rng.uniform(-1, 1, (20, n))
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0.0] * dim, [1.0] * dim, leaf,
                       order, ab.Kernel.smooth_inv_r(1e-3))
    xt = torch.tensor(X, device="cuda")
    y = f.apply_colmajor(xt, torch.zeros_like(xt)).cpu().numpy()
    ym = f.apply_colmajor(xt, torch.zeros_like(xt), ab.MIXED).cpu().numpy()
    assert np.array_equal(y, ym)
    t = oracle.Octree([0.0] * dim, [1.0] * dim, pos, leaf)
    assert np.array_equal(f.order().cpu().numpy(), t.order)
    L = t.interaction_lists()
    info = f.info
    assert info["levels"] == t.levels
    assert info["m2l_pairs"] == len(L["m2l_src"]) and info["p2p_pairs"] == len(L["p2p_src"])
    if dim == 1:
        assert t.levels >= 20
    for r in (0, 17):
        yo = np.empty(n)
        yo[t.order] = t.fmm_matvec(order, (KI, 1e-3), X[r][t.order])
        assert rel(y[r], yo) < 1e-11, (r, rel(y[r], yo))
    with pytest.raises(Exception):
        f.compress(1e-6)
    f.close()


@pytest.mark.parametrize("dim", [1, 2])
def test_fmm_lower_dims_reference_accuracy(ctx, dim):
    """tests/fmm.h:165-215 in 1D and 2D: N = 1000, 10 per leaf, cosine
    source, sqrt(|dx|^2 + 0.01), order 3: sum (y - y_direct)^2 /
    sum y_direct^2 < 1e-2 against the FP64 direct sum."""
    n = 1000
    rng = np.random.default_rng(70 + dim)
    pos = rng.random((n, dim))
    q = np.prod(np.cos(2 * np.pi * pos), axis=1) / n
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda"), [0.0] * dim, [1.0] * dim, 10, 3,
                       ab.Kernel.multiquadric(0.1))
    y = (f @ torch.tensor(q, device="cuda")).cpu().numpy()
    ref = oracle.dense_matvec(pos, pos, (KM, 0.1), q)
    assert np.sum((y - ref) ** 2) / np.sum(ref ** 2) < 1e-2
    f.close()


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("n", [0, 1, 5])
def test_fmm_tiny_inputs(ctx, dim, n):
    """Empty and tiny particle sets (at most 8 per leaf: every pair is near,
    so the FMM is its P2P): y += K x leaves y untouched for n = 0 and equals
    the FP64 direct sum (self term included) otherwise; the tree's order is
    the oracle's."""
    rng = np.random.default_rng(90 + 10 * dim + n)
    pos = rng.random((n, dim))
    f = ab.FmmOperator(ctx, torch.tensor(pos, device="cuda").reshape(n, dim), [0.0] * dim,
                       [1.0] * dim, 8, 4, ab.Kernel.smooth_inv_r(1e-3))
    x = rng.uniform(-1, 1, (3, n))
    xt = torch.tensor(x, device="cuda")
    y = f.apply_colmajor(xt, torch.full_like(xt, 2.0)).cpu().numpy()
    if n == 0:
        assert y.size == 0
    else:
        t = oracle.Octree([0.0] * dim, [1.0] * dim, pos, 8)
        assert np.array_equal(f.order().cpu().numpy(), t.order)
        for r in range(3):
            ref = 2.0 + oracle.dense_matvec(pos, pos, (KI, 1e-3), x[r])
            assert rel(y[r], ref) < 1e-12
    f.close()


def test_fmm_c3_scale(ctx):
    """The C3 workload itself (N = 2,000,000 uniform, leaves of at most 64:
    6 levels, 107,417 leaves, the leaf-level M2L tiles): FP64 order 6 within
    2e-5 of the FP64 direct sum on 300 seeded targets (the CPU oracle's),
    MIXED within 5e-5 of FP64, and order 8 compressed at 2e-7 within 1e-6."""
    from aboria_amd import workloads as wl
    c = wl.FMM
    print("fmm C3 scale", flush=True)
    pos = wl.fmm_positions(c["n"], c["seed"])
    x = wl.fmm_rhs(c["n"], 2, c["seed"] + 1)
    gp = torch.tensor(pos, device="cuda")
    xt = torch.tensor(x, device="cuda")
    idx = np.sort(np.random.default_rng(5).choice(c["n"], 300, replace=False))
    ref = np.stack([oracle.dense_matvec(pos[idx], pos, (KI, c["eps"]), x[r]) for r in range(2)])
    f = ab.FmmOperator(ctx, gp, [0] * 3, [1] * 3, c["n_leaf"], 6, ab.Kernel.smooth_inv_r(c["eps"]))
    info = f.info
    assert info["levels"] == 6 and info["n_leaves"] == 107417
    y = f.apply_colmajor(xt, torch.zeros_like(xt)).cpu().numpy()
    ym = f.apply_colmajor(xt, torch.zeros_like(xt), ab.MIXED).cpu().numpy()
    e64, emix = rel(y[:, idx], ref), rel(ym, y)
    f.close()
    f8 = ab.FmmOperator(ctx, gp, [0] * 3, [1] * 3, c["n_leaf"], 8,
                        ab.Kernel.smooth_inv_r(c["eps"]))
    f8.compress(2e-7)
    y8 = f8.apply_colmajor(xt, torch.zeros_like(xt)).cpu().numpy()
    e8 = rel(y8[:, idx], ref)
    f8.close()
    print("C3 scale: fp64 order 6 %.2e, mixed vs fp64 %.2e, order 8 compressed %.2e" % (
        e64, emix, e8), flush=True)
    assert e64 < 2e-5 and emix < 5e-5 and e8 < 1e-6
