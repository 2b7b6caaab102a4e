"""abh_dense_apply_noalloc (the allocation-free variant of abh_dense_apply,
column-split partial sums in a static device array) against abh_dense_apply,
through the C ABI.

Bar: bit-identical y (the same kernels, column split and range-ordered
reduction run on both paths), on two calls in a row with the second smaller,
in FP64 and FP32-compensated; a shape whose split needs more than the static
array is refused with ABH_ERR_CAPACITY before any launch (y untouched)."""
import ctypes

import numpy as np
import pytest
import torch

import aboria_amd as ab
from aboria_amd import _capi as C

pytestmark = pytest.mark.gpu

ABH_ERR_CAPACITY = 4


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _call(fn, ctx, rows, cols, kernel, x, y, precision):
    kd = kernel.to_c()
    return fn(ctx.handle, rows.shape[1], rows.shape[0], _ptr(rows), cols.shape[0], _ptr(cols),
              ctypes.byref(kd), x.shape[0], _ptr(x), x.shape[1], _ptr(y), y.shape[1],
              int(precision))


@pytest.mark.parametrize("precision", [ab.FP64, ab.FP32_COMPENSATED])
def test_noalloc_matches_dense_apply(ctx, precision):
    lib = C.lib()
    rng = np.random.default_rng(71)
    kernel = ab.Kernel.multiquadric(0.05)
    # (rows, columns, right-hand sides): the first two split the columns
    # over many ranges (few rows), the second call is the smaller one; the
    # last has rows enough for a single range
    for nrows, ncols, nrhs in ((1000, 200_000, 3), (300, 50_000, 2), (600_000, 3000, 1)):
        print("noalloc case rows=%d cols=%d rhs=%d precision=%d" % (nrows, ncols, nrhs,
                                                                     int(precision)), flush=True)
        rows = torch.tensor(rng.random((nrows, 3)), device="cuda")
        cols = torch.tensor(rng.random((ncols, 3)), device="cuda")
        x = torch.tensor(rng.standard_normal((nrhs, ncols)), device="cuda")
        y0 = torch.tensor(rng.standard_normal((nrhs, nrows)), device="cuda")
        y_ref, y_na = y0.clone(), y0.clone()
        assert _call(lib.abh_dense_apply, ctx, rows, cols, kernel, x, y_ref, precision) == 0
        assert _call(lib.abh_dense_apply_noalloc, ctx, rows, cols, kernel, x, y_na,
                     precision) == 0
        torch.cuda.synchronize()
        assert torch.equal(y_na, y_ref)
        assert not torch.equal(y_na, y0)


def test_noalloc_refuses_oversized_split(ctx):
    lib = C.lib()
    rng = np.random.default_rng(72)
    # 1,500 row blocks of 256 on a 256-CU device: the columns split in two
    # ranges, 2 x 16 x 384,000 partial sums (98 MB) exceed the 32 MiB array
    nrows, ncols, nrhs = 384_000, 8192, 16
    rows = torch.tensor(rng.random((nrows, 3)), device="cuda")
    cols = torch.tensor(rng.random((ncols, 3)), device="cuda")
    x = torch.tensor(rng.standard_normal((nrhs, ncols)), device="cuda")
    y = torch.zeros((nrhs, nrows), dtype=torch.float64, device="cuda")
    props = torch.cuda.get_device_properties(0)
    if props.multi_processor_count < 128:
        pytest.skip("the split of this shape depends on the CU count")
    st = _call(lib.abh_dense_apply_noalloc, ctx, rows, cols, ab.Kernel.multiquadric(0.05), x, y,
               ab.FP64)
    assert st == ABH_ERR_CAPACITY
    assert b"static scratch" in lib.abh_last_error(ctx.handle)
    torch.cuda.synchronize()
    assert torch.count_nonzero(y) == 0
