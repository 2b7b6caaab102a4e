"""The general sparse operator (a radius per row, block-valued kernels) and
the device query header used from a kernel outside the library, through
the C ABI, against the oracle.

Bars: the reference's 2-vector KAT exactly (tests/operators.h:904-933);
per-row neighbour counts, the neighbour index sums of the external kernel
and the CSR columns bit-exact (the oracle's search per row); FP64 applies
within 1e-13 of sum_j |f_ij x_j| (kernel evaluation differs from libm by a
few ulp only)."""
import ctypes
import os

import numpy as np
import pytest
import torch

import aboria_amd as ab
import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _build(ctx, pos, periodic, n_leaf):
    dim = pos.shape[1]
    cl = ab.CellListOrdered(ctx, [0.0] * dim, [1.0] * dim, [periodic] * dim, n_leaf)
    cl.update_positions(torch.tensor(pos, device="cuda"))
    oc = oracle.Cells([0.0] * dim, [1.0] * dim, [periodic] * dim, n_leaf)
    oc.update(pos.copy())
    assert np.array_equal(cl.positions().cpu().numpy(), oc.positions)
    return cl, oc


def test_block_kernel_kat(ctx):
    """C2 = [s1_a + s2_b, s1_a - s2_b] within 0.1 of three particles on a
    line, s1 = 1, s2 = 2: the matrix written out at tests/operators.h:904-909
    exactly, and C2 (1, 2, 3) = (9, -3, 18, -6, 15, -5).

    The reference's check vector (tests/operators.h:929) holds -7 in place
    of -6, but its loop compares only the first n = 3 entries
    (tests/operators.h:930-932); the matrix it states gives
    -1 - 2 - 3 = -6 for that entry."""
    d = 0.1
    pos = torch.tensor([[0, 0, 0], [d * 0.9, 0, 0], [d * 1.8, 0, 0]], dtype=torch.float64,
                       device="cuda")
    parts = ab.Particles(ctx, pos, s1=torch.full((3,), 1.0, dtype=torch.float64),
                         s2=torch.full((3,), 2.0, dtype=torch.float64))
    parts.init_neighbour_search([-1] * 3, [1] * 3, [False] * 3)
    k = ab.BlockKernel(ab.Kernel.constant(), w0=[[0.0], [0.0]], w_row=[[1.0], [1.0]],
                       w_col=[[1.0], [-1.0]])
    C2 = ab.create_sparse_operator(parts, parts, d, k, row_scalar="s1", col_scalar="s2")
    ids = parts.id.cpu().numpy()
    v = torch.tensor(np.array([1.0, 2.0, 3.0])[ids], device="cuda")
    ans = (C2 @ v).cpu().numpy().reshape(3, 2)
    check = np.array([9.0, -3.0, 18.0, -6.0, 15.0, -5.0]).reshape(3, 2)
    assert np.array_equal(ans, check[ids])
    # assemble: the dense form times v equals the apply (tests/operators.h:935-945)
    rp, col, val = C2.assemble()
    A = np.zeros((6, 3))
    rpn, coln, valn = rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
    for r in range(6):
        for e in range(rpn[r], rpn[r + 1]):
            A[r, coln[e]] += valn[e]
    assert np.array_equal(A @ v.cpu().numpy(), ans.reshape(-1))
    M = np.array([[3, 3, 0], [-1, -1, 0], [3, 3, 3], [-1, -1, -1], [0, 3, 3], [0, -1, -1]],
                 dtype=np.float64)
    for k in range(3):
        for b in range(2):
            assert np.array_equal(A[2 * k + b], M[2 * ids[k] + b, ids])


@pytest.mark.parametrize("dim,periodic", [(2, False), (3, True), (1, True)])
def test_variable_radius_matches_oracle(ctx, dim, periodic):
    """create_sparse_operator(rows, cols, radius_function, f): Gaussian
    operator with a radius per row in [0.5, 1.5] r0."""
    n = {1: 2000, 2: 3000, 3: 3000}[dim]
    r0 = {1: 0.004, 2: 0.05, 3: 0.1}[dim]
    rng = np.random.default_rng(dim * 10 + periodic)
    pos = rng.random((n, dim))
    cl, oc = _build(ctx, pos, periodic, 6.0)
    P = oc.positions
    rr = r0 * (0.5 + rng.random(n))
    x = rng.normal(size=n)
    op = ab.SparseBlockOperator(ctx, cl, cl.positions(), 0.0, ab.Kernel.gaussian(r0),
                                row_radius=torch.tensor(rr, device="cuda"))
    y = (op @ torch.tensor(x, device="cuda")).cpu().numpy()
    rp, col, val = op.assemble()
    rpn, coln = rp.cpu().numpy(), col.cpu().numpy()
    y_o = np.zeros(n)
    scale = np.zeros(n)
    total = 0
    for i in range(n):
        prow, crow = oc.search(P[i:i + 1], rr[i])
        assert np.array_equal(coln[rpn[i]:rpn[i + 1]], crow), i  # visit order kept
        dx = P[crow] - P[i]
        if periodic:
            dx -= np.round(dx)
        f = np.exp(-(dx * dx).sum(1) / r0 ** 2)
        y_o[i] = (f * x[crow]).sum()
        scale[i] = (f * np.abs(x[crow])).sum()
        total += len(crow)
    assert op.pairs == total
    assert np.all(np.abs(y - y_o) <= 1e-13 * scale + 1e-300)


def test_block_operator_multi_entry(ctx):
    """A 3 x 2 block Wendland kernel with row and column scalars: apply equals
    the assembled CSR times x, and the CSR equals entries computed from the
    oracle's neighbour lists (columns exactly, values within 1e-13 of the
    weights' magnitude)."""
    n, r = 2500, 0.06
    rng = np.random.default_rng(5)
    pos = rng.random((n, 2))
    cl, oc = _build(ctx, pos, False, 5.0)
    a, b = rng.random(n), rng.random(n)
    w0, wr, wc = rng.normal(size=(3, 3, 2))
    k = ab.BlockKernel(ab.Kernel.wendland_c2(r), w0, wr, wc)
    op = ab.SparseBlockOperator(ctx, cl, cl.positions(), r, k,
                                row_scalar=torch.tensor(a, device="cuda"),
                                col_scalar=torch.tensor(b, device="cuda"))
    x = rng.normal(size=2 * n)
    y = (op @ torch.tensor(x, device="cuda")).cpu().numpy()
    rp, col, val = (t.cpu().numpy() for t in op.assemble())
    ys = np.add.reduceat(val * x[col], rp[:-1])
    assert np.max(np.abs(ys - y)) <= 1e-13 * np.max(np.abs(y))
    P = oc.positions
    rp_o, col_o = oc.search(P, r)
    for i in rng.choice(n, 50, replace=False):
        nb = col_o[rp_o[i]:rp_o[i + 1]]
        q = np.linalg.norm(P[nb] - P[i], axis=1) / r
        phi = np.where(q < 1, (1 - q) ** 4 * (4 * q + 1), 0.0)
        for ii in range(3):
            seg = slice(rp[3 * i + ii], rp[3 * i + ii + 1])
            assert np.array_equal(col[seg], (2 * nb[:, None] + np.arange(2)).ravel())
            want = (phi[:, None] * (w0[ii] + wr[ii] * a[i] + wc[ii] * b[nb][:, None])).ravel()
            # phi <= 1 is computed from sqrt and (1 - q)^4 on both sides,
            # which agree to a few ulp of 1 (not of phi near q = 1), and the
            # weights can cancel: the bound scales with the weights' size
            mag = (np.abs(w0[ii]) + np.abs(wr[ii] * a[i])
                   + np.abs(wc[ii] * b[nb][:, None])).ravel()
            assert np.all(np.abs(val[seg] - want) <= 1e-13 * mag)


@pytest.mark.parametrize("dim,periodic", [(3, True), (2, False)])
def test_device_query_header_from_external_kernel(ctx, dim, periodic):
    """tests/cpp/libdevice_query_example.so (built with contraction on)
    counts and sums neighbour indices through abh::for_each_neighbour on the
    view of a cell list: bit-exact against the oracle search."""
    lib = ctypes.CDLL(os.path.join(HERE, "cpp", "libdevice_query_example.so"))
    lib.example_count_neighbours.restype = ctypes.c_int
    n, r = 20000, (0.06 if dim == 3 else 0.02)
    rng = np.random.default_rng(3 + dim)
    pos = rng.random((n, dim))
    cl, oc = _build(ctx, pos, periodic, 8.0)
    centres = rng.uniform(-0.1, 1.1, (4000, dim))
    centres[5, 0] = np.nan
    ct = torch.tensor(centres, device="cuda")
    cnt = torch.zeros(4000, dtype=torch.int64, device="cuda")
    sm = torch.zeros(4000, dtype=torch.int64, device="cuda")
    v = cl.query_view()
    torch.cuda.synchronize()
    assert lib.example_count_neighbours(ctypes.byref(v), ctypes.c_void_p(ct.data_ptr()),
                                        ctypes.c_int64(4000), ctypes.c_double(r),
                                        ctypes.c_void_p(cnt.data_ptr()),
                                        ctypes.c_void_p(sm.data_ptr())) == 0
    dg = oc.search_digest(centres, r)
    assert np.array_equal(cnt.cpu().numpy(), dg[:, 0])
    assert np.array_equal(sm.cpu().numpy(), dg[:, 1])


def test_particles_vtu_round_trip(ctx, tmp_path):
    """A device particle set after its neighbour-search reorder, written as a
    VTK unstructured grid (Float64 arrays) and read back into a new device
    set: positions, ids and columns bit-exact, and the new set's cell list
    identical (aboria_amd/io.py, src/Particles.h:586-684)."""
    from aboria_amd import io as pio
    rng = np.random.default_rng(8)
    pos = torch.tensor(rng.random((3000, 3)), device="cuda")
    parts = ab.Particles(ctx, pos, density=torch.tensor(rng.random(3000)))
    parts.init_neighbour_search([0] * 3, [1] * 3, [True] * 3, 10.0)
    h = parts.to_host()
    f = str(tmp_path / "parts.vtu")
    pio.write_vtu(f, h["position"], id=h["id"], alive=h["alive"], columns=h["columns"],
                  float64=True)
    r = pio.read_vtu(f)
    back = ab.Particles.from_host(ctx, r["position"], id=r["id"], alive=r["alive"],
                                  columns={"density": r["density"]})
    assert torch.equal(back.position, parts.position)
    assert torch.equal(back.id, parts.id)
    assert torch.equal(back.columns["density"], parts.columns["density"])
    back.init_neighbour_search([0] * 3, [1] * 3, [True] * 3, 10.0)
    # already in bucket order: the rebuild keeps the order
    assert torch.equal(back.position, parts.position)
    assert torch.equal(back.id, parts.id)
