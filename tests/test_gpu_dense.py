"""GPU batched dense operators and batched GMRES.

Bar: FP64 batched matvec within 1e-13 * sum_j |f_ij x_j| of the oracle's
KernelDense::evaluate restatement for every problem; FP32-compensated within
2e-6 of the same scale; GMRES reaches ||b - A x|| <= 1e-8 ||b|| with the
residual recomputed independently by the oracle; the GMRES(m) iterate after
a fixed number of steps (full or restarted, eager or graph-replayed) has the
residual of the Krylov-space least-squares minimiser within 1e-7 relative
(1e-6 with the batched Schwarz preconditioner on the right).
"""
import numpy as np
import pytest
import torch

import aboria_amd as ab
import oracle
from krylov_ref import gmres_iterate

pytestmark = pytest.mark.gpu

KM = 3


def make_batch(sizes, seed):
    rng = np.random.default_rng(seed)
    pos = [rng.random((n, 2)) for n in sizes]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return pos, off


@pytest.mark.parametrize("precision", [ab.FP64, ab.FP32_COMPENSATED])
def test_dense_batch_matches_oracle(ctx, precision):
    sizes = [1, 17, 256, 300, 1000, 2500]
    pos, off = make_batch(sizes, 3)
    allpos = torch.tensor(np.concatenate(pos), device="cuda")
    A = ab.DenseBatchOperator(ctx, off, allpos, ab.Kernel.multiquadric(0.05))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(off[-1])
    y = torch.full((int(off[-1]),), 0.5, dtype=torch.float64, device="cuda")
    A.apply(torch.tensor(x, device="cuda"), y, precision)
    y = y.cpu().numpy() - 0.5
    tol = 1e-13 if precision == ab.FP64 else 2e-6
    for p in range(len(sizes)):
        a, b = off[p], off[p + 1]
        ref = oracle.dense_matvec(pos[p], pos[p], (KM, 0.05), x[a:b])
        scale = oracle.dense_matvec(pos[p], pos[p], (KM, 0.05), np.abs(x[a:b]))
        assert np.all(np.abs(y[a:b] - ref) <= tol * scale + 1e-300), p


@pytest.mark.parametrize("precision", [ab.FP64, ab.FP32_COMPENSATED])
def test_batched_gmres_residual(ctx, precision):
    # Gaussian RBF systems (well conditioned enough for a short test)
    sizes = [200, 500, 800]
    pos, off = make_batch(sizes, 5)
    allpos = torch.tensor(np.concatenate(pos), device="cuda")
    A = ab.DenseBatchOperator(ctx, off, allpos, ab.Kernel.gaussian(0.02))
    rng = np.random.default_rng(6)
    b = rng.standard_normal(off[-1])
    x, stats = A.gmres(torch.tensor(b, device="cuda"), restart=50, max_iters=2000, tol=1e-8,
                       precision=precision)
    x = x.cpu().numpy()
    for p, st in enumerate(stats):
        a, e = off[p], off[p + 1]
        r = b[a:e] - oracle.dense_matvec(pos[p], pos[p], (1, 0.02), x[a:e])
        res = np.linalg.norm(r) / np.linalg.norm(b[a:e])
        assert st["converged"], st
        assert res <= 1.01e-8, (p, res, st)


@pytest.mark.parametrize("restart,max_iters,graph", [
    (50, 1, False), (50, 3, False), (50, 12, False), (4, 12, False), (7, 12, True)])
def test_gmres_iterates_are_krylov_minimisers(ctx, restart, max_iters, graph):
    # GMRES(m) (what Eigen::GMRES computes with restart m) is fixed by the
    # mathematics: each cycle of k steps from x adds the d in the Krylov
    # space K_k(A, b - A x) that minimises ||b - A (x + d)||.  With a
    # tolerance no residual reaches (1e-300) every problem runs exactly
    # max_iters steps, cycles of `restart` steps and a shorter last one;
    # the residual after each cycle is pinned here against the dense
    # least-squares restatement of krylov_ref (a dropped term, wrong sign
    # or index in the Arnoldi or Givens updates leaves a larger residual).
    sizes = [150, 300, 600]
    pos, off = make_batch(sizes, 11)
    allpos = torch.tensor(np.concatenate(pos), device="cuda")
    A = ab.DenseBatchOperator(ctx, off, allpos, ab.Kernel.gaussian(0.05))
    b = np.random.default_rng(12).standard_normal(off[-1])
    x, stats = A.gmres(torch.tensor(b, device="cuda"), restart=restart, max_iters=max_iters,
                       tol=1e-300, graph=graph)
    x = x.cpu().numpy()
    for p, st in enumerate(stats):
        a, e = off[p], off[p + 1]
        K = oracle.dense_matvec(pos[p], pos[p], (1, 0.05), np.eye(sizes[p]))
        bp = b[a:e]
        xr = gmres_iterate(K, bp, restart, max_iters)
        want = np.linalg.norm(bp - K @ xr) / np.linalg.norm(bp)
        got = np.linalg.norm(bp - K @ x[a:e]) / np.linalg.norm(bp)
        assert st["iterations"] == max_iters and not st["converged"], st
        assert 1e-3 < want < 0.9, want
        assert abs(got - want) <= 1e-7 * want, (p, got, want)
        assert abs(st["residual"] - got) <= 1e-9 * got, (p, st, got)


@pytest.mark.parametrize("restart,max_iters,graph", [
    (50, 1, False), (50, 2, False), (2, 4, False), (2, 4, True)])
def test_preconditioned_gmres_iterates_are_krylov_minimisers(ctx, restart, max_iters, graph):
    # Right-preconditioned GMRES(m): per cycle x += M^-1 u with u in
    # K_k(A M^-1, r) minimising ||r - A M^-1 u||.  M^-1 is the batched
    # Schwarz apply (pinned against the oracle by the test above), taken
    # column by column as a matrix.
    sizes = [300, 600, 900]
    pos, off = make_batch(sizes, 31)
    allp = np.concatenate(pos)
    A = ab.DenseBatchOperator(ctx, off, torch.tensor(allp, device="cuda"),
                              ab.Kernel.gaussian(0.05))
    sw = ab.SchwartzPreconditioner.for_dense_batch(A, [0, 0], [1, 1], 10.0, 1e-6)
    minv = [np.zeros((n, n)) for n in sizes]
    for j in range(max(sizes)):
        e = torch.zeros(int(off[-1]), dtype=torch.float64, device="cuda")
        for p in range(len(sizes)):
            if j < sizes[p]:
                e[int(off[p]) + j] = 1.0
        y = sw.solve(e).cpu().numpy()
        for p in range(len(sizes)):
            if j < sizes[p]:
                minv[p][:, j] = y[off[p]:off[p + 1]]
    b = np.random.default_rng(32).standard_normal(off[-1])
    x, stats = A.gmres(torch.tensor(b, device="cuda"), restart=restart, max_iters=max_iters,
                       tol=1e-300, preconditioner=sw, graph=graph)
    x = x.cpu().numpy()
    for p, st in enumerate(stats):
        a, e = off[p], off[p + 1]
        K = oracle.dense_matvec(pos[p], pos[p], (1, 0.05), np.eye(sizes[p]))
        bp = b[a:e]
        xr = gmres_iterate(K, bp, restart, max_iters, minv=minv[p])
        want = np.linalg.norm(bp - K @ xr) / np.linalg.norm(bp)
        got = np.linalg.norm(bp - K @ x[a:e]) / np.linalg.norm(bp)
        print("preconditioned GMRES", restart, max_iters, p, "%.3e %.3e" % (got, want),
              flush=True)
        assert st["iterations"] == max_iters and not st["converged"], st
        assert 1e-9 < want < 1.0 - 1e-6, want
        assert abs(got - want) <= 1e-6 * want, (p, got, want)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_rectangular_dense_apply_matches_oracle(ctx, dim):
    rng = np.random.default_rng(50 + dim)
    a = rng.random((777, dim))
    b = rng.random((1500, dim))
    x = rng.normal(size=(20, 1500))  # 20 right-hand sides: two passes
    kern = (KM, 0.05)
    ref = oracle.dense_matvec(a, b, kern, x.T).T
    y0 = rng.normal(size=(20, 777))
    at, bt = torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda")
    y = ab.dense_apply(ctx, at, bt, ab.Kernel.multiquadric(0.05), torch.tensor(x, device="cuda"),
                       torch.tensor(y0, device="cuda")).cpu().numpy()
    scale = np.abs(oracle.dense_matvec(a, b, kern, np.abs(x.T)).T)
    assert np.max(np.abs(y - y0 - ref) / scale) < 1e-13
    y32 = ab.dense_apply(ctx, at, bt, ab.Kernel.multiquadric(0.05), torch.tensor(x, device="cuda"),
                         precision=ab.FP32_COMPENSATED).cpu().numpy()
    assert np.max(np.abs(y32 - ref) / scale) < 2e-6


def test_rectangular_dense_apply_column_split(ctx):
    """Few rows, many columns: the columns are split over blocks and the
    partial sums added in range order (deterministic, within the FP64 bar)."""
    rng = np.random.default_rng(77)
    a = rng.random((300, 3))
    b = rng.random((20000, 3))
    x = rng.normal(size=(20, 20000))
    kern = (KM, 0.05)
    ref = oracle.dense_matvec(a, b, kern, x.T).T
    scale = np.abs(oracle.dense_matvec(a, b, kern, np.abs(x.T)).T)
    at, bt = torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda")
    xt = torch.tensor(x, device="cuda")
    y = ab.dense_apply(ctx, at, bt, ab.Kernel.multiquadric(0.05), xt).cpu().numpy()
    assert np.max(np.abs(y - ref) / scale) < 1e-13
    y2 = ab.dense_apply(ctx, at, bt, ab.Kernel.multiquadric(0.05), xt).cpu().numpy()
    assert np.array_equal(y, y2)
    y32 = ab.dense_apply(ctx, at, bt, ab.Kernel.multiquadric(0.05), xt,
                         precision=ab.FP32_COMPENSATED).cpu().numpy()
    assert np.max(np.abs(y32 - ref) / scale) < 2e-6


def test_dense_tile_variants_agree_and_tuner_picks_one(ctx):
    """Every P2P tile shape gives the same FP64 rows (same summation order
    per row) and the FP32c bar; the autotuner returns one of them."""
    rng = np.random.default_rng(77)
    sizes = [700, 1500, 2600, 90]
    off = np.concatenate([[0], np.cumsum(sizes)])
    pos = rng.random((off[-1], 2))
    x = rng.normal(size=off[-1])
    op = ab.DenseBatchOperator(ctx, off, torch.tensor(pos, device="cuda"),
                               ab.Kernel.multiquadric(0.05))
    xt = torch.tensor(x, device="cuda")
    outs = {}
    for rpt in (1, 2, 4):
        op.tune(ab.FP64, force=rpt)
        outs[rpt] = (op @ xt).cpu().numpy()
    assert np.array_equal(outs[1], outs[2]) and np.array_equal(outs[1], outs[4])
    ref = np.concatenate([oracle.dense_matvec(pos[a:b], pos[a:b], (KM, 0.05), x[a:b])
                          for a, b in zip(off[:-1], off[1:])])
    assert np.linalg.norm(outs[1] - ref) / np.linalg.norm(ref) < 1e-13
    for rpt in (1, 2, 4):
        op.tune(ab.FP32_COMPENSATED, force=rpt)
        y32 = op.apply(xt, torch.zeros_like(xt), ab.FP32_COMPENSATED).cpu().numpy()
        assert np.linalg.norm(y32 - ref) / np.linalg.norm(ref) < 2e-6
    # a fresh operator tunes itself
    op2 = ab.DenseBatchOperator(ctx, off, torch.tensor(pos, device="cuda"),
                                ab.Kernel.multiquadric(0.05))
    rpt, ms = op2.tune(ab.FP64)
    print("tuned rows per thread", rpt, ms, flush=True)
    assert rpt in (1, 2, 4) and all(v > 0 for v in ms.values())


def test_batch_schwarz_domains_match_oracle(ctx):
    """Domain lists of the batched Schwarz preconditioner, bit-exact."""
    from oracle import batch_schwarz as bs
    sizes = [1500, 300, 4000, 1, 20]
    pos, off = make_batch(sizes, 21)
    allp = np.concatenate(pos)
    A = ab.DenseBatchOperator(ctx, off, torch.tensor(allp, device="cuda"),
                              ab.Kernel.multiquadric(0.05))
    sw = ab.SchwartzPreconditioner.for_dense_batch(A, [0, 0], [1, 1], 32.0)
    o, i, n = sw.domains()
    oo, oi, on = bs.domains(off, allp, [0, 0], [1, 1], 32.0)
    assert np.array_equal(o, oo) and np.array_equal(i, oi) and np.array_equal(n, on)
    assert sw.info["failed_domains"] == 0 and sw.info["n"] == off[-1]


@pytest.mark.parametrize("kernel,shift,tol", [((1, 0.05), 1e-6, 1e-9), ((KM, 0.05), 0.0, 1e-5)])
def test_batch_schwarz_apply_matches_oracle(ctx, kernel, shift, tol):
    """Signed-Cholesky local inverses against the oracle's LU local solves:
    1e-9 relative for the positive definite Gaussian blocks, 1e-5 for the
    indefinite multiquadric blocks (condition numbers up to ~1e9)."""
    from oracle import batch_schwarz as bs
    sizes = [1200, 250, 2000]
    pos, off = make_batch(sizes, 22)
    allp = np.concatenate(pos)
    k = ab.Kernel.gaussian(kernel[1]) if kernel[0] == 1 else ab.Kernel.multiquadric(kernel[1])
    A = ab.DenseBatchOperator(ctx, off, torch.tensor(allp, device="cuda"), k)
    sw = ab.SchwartzPreconditioner.for_dense_batch(A, [0, 0], [1, 1], 40.0, shift)
    b = np.random.default_rng(23).normal(size=int(off[-1]))
    x = sw.solve(torch.tensor(b, device="cuda")).cpu().numpy()
    ref = bs.apply(allp, bs.domains(off, allp, [0, 0], [1, 1], 40.0), kernel, shift, b)
    for p in range(len(sizes)):
        a, e = off[p], off[p + 1]
        err = np.linalg.norm(x[a:e] - ref[a:e]) / np.linalg.norm(ref[a:e])
        print("batch schwarz apply", kernel, p, "rel err %.2e" % err, flush=True)
        assert err < tol
    assert sw.info["failed_domains"] == 0


def test_batch_gmres_with_schwarz_converges(ctx):
    """Multiquadric c = 0.05 interpolation systems of the batched RBF
    workload's form at sizes whose FP64 residual floor lies below 1e-8: the
    preconditioned batch converges (true residual recomputed by the oracle),
    the unpreconditioned one stalls on the same iteration budget."""
    sizes = [1000, 1600, 2200]
    pos, off = make_batch(sizes, 24)
    allp = np.concatenate(pos)
    f = np.sin(2 * np.pi * allp[:, 0]) * np.cos(2 * np.pi * allp[:, 1])
    f += np.random.default_rng(25).normal(0.0, 1e-3, f.shape)
    A = ab.DenseBatchOperator(ctx, off, torch.tensor(allp, device="cuda"),
                              ab.Kernel.multiquadric(0.05))
    sw = ab.SchwartzPreconditioner.for_dense_batch(A, [0, 0], [1, 1], 64.0)
    bt = torch.tensor(f, device="cuda")
    x, st = A.gmres(bt, restart=50, max_iters=400, preconditioner=sw)
    _, st0 = A.gmres(bt, restart=50, max_iters=400)
    print("batch gmres schwarz", st, "none", st0, flush=True)
    x = x.cpu().numpy()
    for p in range(len(sizes)):
        a, e = off[p], off[p + 1]
        assert st[p]["converged"]
        r = oracle.dense_matvec(pos[p], pos[p], (KM, 0.05), x[a:e]) - f[a:e]
        # recomputed in another summation order: rounding of |K| |x| allowed
        assert np.linalg.norm(r) / np.linalg.norm(f[a:e]) <= 2e-8
        assert st[p]["iterations"] < st0[p]["iterations"] or not st0[p]["converged"]


@pytest.mark.parametrize("precision", [ab.FP64, ab.FP32_COMPENSATED])
def test_batch_gmres_graph_equals_eager(ctx, precision):
    """graph=True captures a restart cycle's Arnoldi steps (operator,
    Schwarz apply with its device copy, CGS2, norms, Givens kernels) into
    HIP graphs of 5 steps replayed every cycle, with the host's checks and
    synchronisations between replays.  Restart 4 gives several replays; every
    problem's iterations, flags and residuals and the solution are those of
    the eager run bit for bit (the same kernels in the same order; steps
    after every problem has left a cycle are masked no-ops), and in FP64 the
    true residual recomputed by the oracle meets the tolerance."""
    sizes = [300, 900, 1400, 60]
    pos, off = make_batch(sizes, 31)
    allp = np.concatenate(pos)
    f = np.sin(2 * np.pi * allp[:, 0]) * np.cos(2 * np.pi * allp[:, 1])
    A = ab.DenseBatchOperator(ctx, off, torch.tensor(allp, device="cuda"),
                              ab.Kernel.multiquadric(0.05))
    sw = ab.SchwartzPreconditioner.for_dense_batch(A, [0, 0], [1, 1], 64.0)
    bt = torch.tensor(f, device="cuda")
    xe, se = A.gmres(bt, restart=4, max_iters=400, preconditioner=sw, precision=precision)
    xg, sg = A.gmres(bt, restart=4, max_iters=400, preconditioner=sw, precision=precision,
                     graph=True)
    print("graph", sg, "eager", se, flush=True)
    assert se == sg
    assert torch.equal(xe, xg)
    assert max(s["iterations"] for s in sg) > 8  # three replays or more
    x = xg.cpu().numpy()
    for p in range(len(sizes) if precision == ab.FP64 else 0):
        a, e = off[p], off[p + 1]
        assert sg[p]["converged"], (p, sg[p])
        r = oracle.dense_matvec(pos[p], pos[p], (KM, 0.05), x[a:e]) - f[a:e]
        assert np.linalg.norm(r) / np.linalg.norm(f[a:e]) <= 2e-8
    # unpreconditioned too (no device copy in the graph)
    xe, se = A.gmres(bt, restart=10, max_iters=60, precision=precision)
    xg, sg = A.gmres(bt, restart=10, max_iters=60, precision=precision, graph=True)
    assert se == sg and torch.equal(xe, xg)


def test_batch_gmres_graph_ragged_batch(ctx):
    """Graph-captured GMRES on a ragged batch with an empty problem, a
    one-point problem and a problem whose right-hand side is zero: equal to
    the eager run bit for bit; the one-point problem is solved in one step
    (x = b / K(0)), the zero right-hand side returns x = 0 converged."""
    sizes = [0, 1, 400, 3, 0, 250]
    pos, off = make_batch(sizes, 41)
    allp = np.concatenate(pos)
    b = np.random.default_rng(42).normal(size=int(off[-1]))
    b[off[3]:off[4]] = 0.0
    A = ab.DenseBatchOperator(ctx, off, torch.tensor(allp, device="cuda"),
                              ab.Kernel.gaussian(0.02))
    bt = torch.tensor(b, device="cuda")
    xe, se = A.gmres(bt, restart=20, max_iters=1000, tol=1e-8)
    xg, sg = A.gmres(bt, restart=20, max_iters=1000, tol=1e-8, graph=True)
    assert se == sg and torch.equal(xe, xg)
    x = xg.cpu().numpy()
    assert sg[1]["converged"] and abs(x[off[1]] - b[off[1]]) <= 1e-15 * abs(b[off[1]])
    assert sg[3]["converged"] and np.all(x[off[3]:off[4]] == 0.0)
    for p in (2, 5):
        a, e = off[p], off[p + 1]
        r = b[a:e] - oracle.dense_matvec(pos[p], pos[p], (1, 0.02), x[a:e])
        assert np.linalg.norm(r) / np.linalg.norm(b[a:e]) <= 1.01e-8, (p, sg[p])


def test_batch_gmres_stops_stagnated_problems(ctx):
    """stall_restarts: a problem whose residual cannot reach the tolerance
    (multiquadric at N = 6000: FP64 rounding of K x sits far above 1e-8 ||b||)
    stops after two restarts without a 10% residual cut; a solvable one in
    the same batch converges as without the option."""
    sizes = [1200, 6000]
    pos, off = make_batch(sizes, 26)
    allp = np.concatenate(pos)
    f = np.sin(2 * np.pi * allp[:, 0]) * np.cos(2 * np.pi * allp[:, 1])
    A = ab.DenseBatchOperator(ctx, off, torch.tensor(allp, device="cuda"),
                              ab.Kernel.multiquadric(0.05))
    sw = ab.SchwartzPreconditioner.for_dense_batch(A, [0, 0], [1, 1], 64.0)
    bt = torch.tensor(f, device="cuda")
    _, st = A.gmres(bt, restart=50, max_iters=3000, preconditioner=sw, stall_restarts=2)
    _, st0 = A.gmres(bt, restart=50, max_iters=3000, preconditioner=sw)
    print("stall", st, "no stall", st0, flush=True)
    assert st[0]["converged"] and st[0]["iterations"] == st0[0]["iterations"]
    assert not st[1]["converged"] and st[1]["iterations"] < 3000
    assert st0[1]["iterations"] >= st[1]["iterations"]
    # never returned worse than its best restart: residual below the start
    assert st[1]["residual"] <= 1.0 and st0[1]["residual"] <= 1.0


def _gpu_blocks(op, off):
    """The FP64 matrix blocks the batched operator applies, column by column
    (y = 0, then y += K e_j: each entry is its kernel value times 1 added to
    zeros, so it is exactly the value the apply uses)."""
    n = int(off[-1])
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(e)
    blocks = [np.zeros((int(off[p + 1] - off[p]),) * 2) for p in range(len(off) - 1)]
    for p in range(len(off) - 1):
        for j in range(int(off[p]), int(off[p + 1])):
            e.zero_()
            e[j] = 1.0
            y.zero_()  # the apply adds K e into y
            op.apply(e, y, ab.FP64)
            blocks[p][:, j - int(off[p])] = y[int(off[p]):int(off[p + 1])].cpu().numpy()
    return blocks


def test_dense_residual_dd_is_exact(ctx):
    """The double-double residual equals the exact residual of the FP64
    matrix (Fractions) within its error bound -- the final rounding (2 ulp)
    plus 2 n eps^2 sum_j |K_ij x_j| for the accumulation -- at an x that
    solves the systems to FP64 accuracy, where the plain FP64 residual is
    mostly rounding noise."""
    from fractions import Fraction
    pos, off = make_batch([90, 130], 21)
    allpos = torch.tensor(np.concatenate(pos), device="cuda")
    op = ab.DenseBatchOperator(ctx, off, allpos, ab.Kernel.multiquadric(0.05))
    blocks = _gpu_blocks(op, off)
    rng = np.random.default_rng(22)
    b = rng.normal(size=int(off[-1]))
    x = np.concatenate([np.linalg.solve(blocks[p], b[off[p]:off[p + 1]])
                        for p in range(len(blocks))])
    r = op.residual_dd(torch.tensor(b, device="cuda"), torch.tensor(x, device="cuda"))
    r = r.cpu().numpy()
    r64 = b.copy()
    for p, K in enumerate(blocks):
        sl = slice(int(off[p]), int(off[p + 1]))
        r64[sl] = b[sl] - K @ x[sl]
        for i in range(K.shape[0]):
            exact = Fraction(b[sl][i]) - sum(Fraction(K[i, j]) * Fraction(x[sl][j])
                                             for j in range(K.shape[1]))
            want = float(exact)
            scale = float(np.sum(np.abs(K[i] * x[sl])))
            bound = 2 * np.spacing(abs(want)) + 2 * K.shape[1] * 2.0 ** -106 * scale
            assert abs(r[sl][i] - want) <= bound
    # the FP64 residual differs from the exact one by far more than that
    assert np.max(np.abs(r64 - r)) > 1e3 * np.max(np.spacing(np.abs(r)))


def test_dense_refine_converges_where_gmres_stagnates(ctx):
    """Three systems of the batched RBF workload (multiquadric c = 0.05,
    2,401 / 2,532 / 3,291 points; the first and last stagnate above 1e-8 in
    FP64 GMRES(50) with Schwarz, 300 iterations): iterative refinement with
    the double-double residual and Schwarz-preconditioned GMRES corrections
    to 1e-3 brings all three below 1e-8, checked by an extended-precision
    residual of the operator's own FP64 matrix."""
    from aboria_amd import workloads as wl
    c = wl.RBF_BATCH
    off0, pos0, f0 = wl.rbf_batch_problems(c["nprob"], c["n_min"], c["n_max"], c["seed"])
    sel = [4, 2, 5]
    sizes = [int(off0[i + 1] - off0[i]) for i in sel]
    assert sizes == [2401, 2532, 3291]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pos = np.concatenate([pos0[off0[i]:off0[i + 1]] for i in sel])
    f = np.concatenate([f0[off0[i]:off0[i + 1]] for i in sel])
    op = ab.DenseBatchOperator(ctx, off, torch.tensor(pos, device="cuda"),
                               ab.Kernel.multiquadric(c["c"]))
    sw = ab.SchwartzPreconditioner.for_dense_batch(op, [0.0, 0.0], [1.0, 1.0],
                                                   c["schwarz_n_leaf"])
    b = torch.tensor(f, device="cuda")
    _, st0 = op.gmres(b, restart=50, max_iters=300, tol=1e-8, preconditioner=sw)
    x, st = op.refine(b, tol=1e-8, max_outer=8, inner_tol=1e-3, inner_iters=100,
                      preconditioner=sw)
    print("gmres", [(s["iterations"], s["converged"], "%.2e" % s["residual"]) for s in st0],
          "refine", [(s["iterations"], s["converged"], "%.2e" % s["residual"]) for s in st],
          flush=True)
    assert not st0[0]["converged"] and not st0[2]["converged"]
    assert all(s["converged"] for s in st)
    blocks = _gpu_blocks(op, off)
    xh = x.cpu().numpy()
    for p, K in enumerate(blocks):
        sl = slice(int(off[p]), int(off[p + 1]))
        rr = f[sl].astype(np.longdouble) - K.astype(np.longdouble) @ xh[sl].astype(np.longdouble)
        rel = float(np.sqrt((rr * rr).sum()) / np.linalg.norm(f[sl]))
        assert rel <= 1.01e-8, (p, rel)
