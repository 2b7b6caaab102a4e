"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

P = ctypes.POINTER
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class KernelC(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("reserved", ctypes.c_int32), ("param", ctypes.c_double),
                ("w0", ctypes.c_double), ("w_row", ctypes.c_double), ("w_col", ctypes.c_double)]


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        l.orc_cells_create.restype = vp
        l.orc_cells_create.argtypes = [ctypes.c_int, _dp, _dp, _u8p, ctypes.c_double]
        l.orc_cells_destroy.argtypes = [vp]
        l.orc_cells_update.restype = ctypes.c_int64
        l.orc_cells_update.argtypes = [vp, ctypes.c_int64, _dp, _u8p]
        l.orc_cells_grid.restype = ctypes.c_int64
        l.orc_cells_grid.argtypes = [vp, _u32p, _dp]
        l.orc_cells_get.argtypes = [vp, _i32p, _u32p, _u32p, _u32p, _dp]
        l.orc_search_count.restype = ctypes.c_int64
        l.orc_search_count.argtypes = [vp, ctypes.c_int64, _dp, ctypes.c_double, _i64p, ctypes.c_int]
        l.orc_search_fill.argtypes = [vp, ctypes.c_int64, _dp, ctypes.c_double, _i64p, _i32p,
                                      ctypes.c_int]
        l.orc_bucket_pairs.restype = ctypes.c_int64
        l.orc_bucket_pairs.argtypes = [vp, vp, vp, vp]
        l.orc_pair_counts.argtypes = [vp, ctypes.c_double, _i32p]
        l.orc_search_digest.argtypes = [vp, ctypes.c_int64, _dp, ctypes.c_double, _i64p,
                                        ctypes.c_int]
        l.orc_sparse_matvec.argtypes = [vp, ctypes.c_int64, _dp, ctypes.c_double, P(KernelC),
                                        vp, vp, ctypes.c_int, _dp, ctypes.c_int64, _dp,
                                        ctypes.c_int64, ctypes.c_int]
        l.orc_dense_matvec.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_int64, _dp,
                                       P(KernelC), ctypes.c_int, _dp, ctypes.c_int64, _dp,
                                       ctypes.c_int64, ctypes.c_int]
        l.orc_max_threads.restype = ctypes.c_int
        l.orc_octree_create.restype = vp
        l.orc_octree_create.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _dp,
                                        ctypes.c_double]
        l.orc_octree_lists.restype = ctypes.c_int
        l.orc_octree_lists.argtypes = [vp, _i64p, vp, vp, vp, vp]
        l.orc_octree_destroy.argtypes = [vp]
        l.orc_octree_info.argtypes = [vp, _i64p]
        l.orc_octree_get.argtypes = [vp, _i32p, _i32p, _i32p, _i32p]
        l.orc_fmm_matvec.restype = ctypes.c_int
        l.orc_fmm_matvec.argtypes = [vp, ctypes.c_int, P(KernelC), _dp, _dp]
        l.orc_cheb_node.restype = ctypes.c_double
        l.orc_cheb_node.argtypes = [ctypes.c_uint, ctypes.c_uint]
        l.orc_cheb_Sn.restype = ctypes.c_double
        l.orc_cheb_Sn.argtypes = [ctypes.c_double, ctypes.c_uint, ctypes.c_uint]
        l.orc_schwarz_create.restype = vp
        l.orc_schwarz_create.argtypes = [vp, P(KernelC), ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double]
        l.orc_schwarz_create_coarse.restype = vp
        l.orc_schwarz_create_coarse.argtypes = [vp, P(KernelC), ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_int, ctypes.c_uint64]
        l.orc_schwarz_analyse.restype = vp
        l.orc_schwarz_analyse.argtypes = [vp, ctypes.c_double]
        l.orc_schwarz_destroy.argtypes = [vp]
        l.orc_schwarz_info.argtypes = [vp, _i64p]
        l.orc_schwarz_domains.argtypes = [vp, _i64p, _i32p, _i32p]
        l.orc_schwarz_apply.argtypes = [vp, _dp, _dp]
        l.orc_knn.argtypes = [vp, ctypes.c_int, _i32p, _dp]
        l.orc_lsq_rows.argtypes = [vp, P(KernelC), ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   _i32p, _dp]
        _lib = l
    return _lib


def nthreads_default():
    return lib().orc_max_threads()


class Cells:
    """CellListOrdered restated on the CPU."""

    def __init__(self, low, high, periodic, n_leaf=10.0):
        self.dim = len(low)
        self.low = np.asarray(low, np.float64)
        self.high = np.asarray(high, np.float64)
        self.h = lib().orc_cells_create(self.dim, np.asarray(low, np.float64),
                                        np.asarray(high, np.float64),
                                        np.asarray(periodic, np.uint8), float(n_leaf))
        self.n_alive = 0

    def update(self, pos, alive=None):
        """pos (n, dim) float64 and alive (n,) uint8 are updated in place."""
        pos = np.ascontiguousarray(pos, np.float64)
        n = pos.shape[0]
        if alive is None:
            alive = np.ones(n, np.uint8)
        self.pos_in = pos
        self.alive_in = alive
        self.n_alive = lib().orc_cells_update(self.h, n, pos.reshape(-1), alive)
        grid = np.zeros(3, np.uint32)
        side = np.zeros(3, np.float64)
        self.nb = lib().orc_cells_grid(self.h, grid, side)
        self.grid = tuple(int(g) for g in grid[:self.dim])
        self.side = side[:self.dim].copy()
        na, nb = self.n_alive, self.nb
        self.order = np.zeros(na, np.int32)
        self.keys = np.zeros(na, np.uint32)
        self.begin = np.zeros(nb, np.uint32)
        self.end = np.zeros(nb, np.uint32)
        self.positions = np.zeros((na, self.dim), np.float64)
        lib().orc_cells_get(self.h, self.order, self.keys, self.begin, self.end,
                            self.positions.reshape(-1))
        return self.n_alive

    def search(self, centres, radius, nthreads=None):
        """Neighbour lists in the reference's iteration order (CSR)."""
        c = np.ascontiguousarray(centres, np.float64)
        nq = c.shape[0]
        nt = nthreads or nthreads_default()
        counts = np.zeros(nq, np.int64)
        total = lib().orc_search_count(self.h, nq, c.reshape(-1), float(radius), counts, nt)
        row_ptr = np.zeros(nq + 1, np.int64)
        np.cumsum(counts, out=row_ptr[1:])
        col = np.zeros(max(total, 1), np.int32)
        lib().orc_search_fill(self.h, nq, c.reshape(-1), float(radius), row_ptr, col, nt)
        return row_ptr, col[:total]

    def bucket_pairs(self):
        """The bucket_pair_iterator sequence: (bi, bj, q) with q (npairs, dim)
        the periodic quadrant of each pair (offset q * (high - low))."""
        n = lib().orc_bucket_pairs(self.h, None, None, None)
        bi = np.zeros(max(n, 1), np.int32)
        bj = np.zeros(max(n, 1), np.int32)
        q = np.zeros((max(n, 1), 3), np.int32)
        lib().orc_bucket_pairs(self.h, bi.ctypes.data, bj.ctypes.data, q.ctypes.data)
        return bi[:n], bj[:n], q[:n, :self.dim]

    def pair_counts(self, radius):
        """neighbour counts (strict radius, self included) by bucket-pair
        traversal, in container order"""
        out = np.zeros(self.n_alive, np.int32)
        lib().orc_pair_counts(self.h, float(radius), out)
        return out

    def search_digest(self, centres, radius, nthreads=None):
        """(nq, 4) int64 per-centre digest of the neighbour lists: count, sum
        of indices, sum of squared indices, sum of a 64-bit mix of each
        index (wrapping).  Order-independent: equal for equal multisets."""
        c = np.ascontiguousarray(centres, np.float64)
        nq = c.shape[0]
        out = np.zeros((nq, 4), np.int64)
        lib().orc_search_digest(self.h, nq, c.reshape(-1), float(radius), out.reshape(-1),
                                nthreads or nthreads_default())
        return out

    def sparse_matvec(self, centres, radius, kernel, x, y=None, row_s=None, col_s=None,
                      nthreads=None):
        """y += K x, x (ncols,) or (ncols, nrhs)."""
        c = np.ascontiguousarray(centres, np.float64)
        nq = c.shape[0]
        x2 = np.asarray(x, np.float64)
        if x2.ndim == 1:
            x2 = x2[:, None]
        nrhs = x2.shape[1]
        xt = np.ascontiguousarray(x2.T)
        yt = np.zeros((nrhs, nq)) if y is None else np.ascontiguousarray(np.asarray(y).reshape(nq, -1).T)
        kd = KernelC(kernel[0], 0, kernel[1], *(kernel[2:] if len(kernel) > 2 else (0.0, 0.0, 0.0)))
        rs = np.ascontiguousarray(row_s, np.float64) if row_s is not None else None
        cs = np.ascontiguousarray(col_s, np.float64) if col_s is not None else None
        lib().orc_sparse_matvec(self.h, nq, c.reshape(-1), float(radius), ctypes.byref(kd),
                                rs.ctypes.data if rs is not None else None,
                                cs.ctypes.data if cs is not None else None,
                                nrhs, xt.reshape(-1), xt.shape[1], yt.reshape(-1), yt.shape[1],
                                nthreads or nthreads_default())
        out = yt.T.copy()
        return out[:, 0] if np.asarray(x).ndim == 1 else out

    def schwarz(self, kernel, radius, shift, buffer=1e5 * np.finfo(float).eps,
                coarse_grid_n=0, seed=0):
        """Restricted additive Schwarz preconditioner of A = K + shift I
        (coarse_grid_n > 0: plus the coarse-grid domain, solved last)."""
        return Schwarz(self, kernel, radius, shift, buffer, coarse_grid_n, seed)

    def schwarz_domains(self, buffer=1e5 * np.finfo(float).eps):
        """(offsets, indices, n_interior) of the Schwarz domains
        (analyzePattern only, no factorisation)."""
        h = lib().orc_schwarz_analyse(self.h, float(buffer))
        try:
            info = np.zeros(3, np.int64)
            lib().orc_schwarz_info(h, info)
            nd, total = int(info[0]), int(info[1])
            off = np.zeros(nd + 1, np.int64)
            idx = np.zeros(max(total, 1), np.int32)
            nint = np.zeros(max(nd, 1), np.int32)
            lib().orc_schwarz_domains(h, off, idx, nint)
        finally:
            lib().orc_schwarz_destroy(h)
        return off, idx[:total], nint[:nd]

    def knn(self, k):
        """(n, k) neighbour indices (self first) and squared distances."""
        nbr = np.zeros((self.n_alive, k), np.int32)
        d2 = np.zeros((self.n_alive, k), np.float64)
        lib().orc_knn(self.h, int(k), nbr.reshape(-1), d2.reshape(-1))
        return nbr, d2

    def lsq_rows(self, kernel, radius, shift, nbr):
        """Least-squares approximate-inverse rows on the pattern nbr (n, k)."""
        nbr = np.ascontiguousarray(nbr, np.int32)
        w = np.zeros(nbr.shape, np.float64)
        kd = KernelC(kernel[0], 0, kernel[1], 0.0, 0.0, 0.0)
        lib().orc_lsq_rows(self.h, ctypes.byref(kd), float(radius), float(shift), nbr.shape[1],
                           nbr.reshape(-1), w.reshape(-1))
        return w

    def __del__(self):
        try:
            lib().orc_cells_destroy(self.h)
        except Exception:
            pass


class Schwarz:
    """SchwartzPreconditioner<LLT> restated on the CPU over a Cells object."""

    def __init__(self, cells, kernel, radius, shift, buffer, coarse_grid_n=0, seed=0):
        self.cells = cells
        kd = KernelC(kernel[0], 0, kernel[1], 0.0, 0.0, 0.0)
        self.h = lib().orc_schwarz_create_coarse(cells.h, ctypes.byref(kd), float(radius),
                                                 float(shift), float(buffer),
                                                 int(coarse_grid_n), ctypes.c_uint64(seed))
        info = np.zeros(3, np.int64)
        lib().orc_schwarz_info(self.h, info)
        self.n_domains, total, self.failed = (int(v) for v in info)
        self.offsets = np.zeros(self.n_domains + 1, np.int64)
        self.indices = np.zeros(max(total, 1), np.int32)
        self.n_interior = np.zeros(max(self.n_domains, 1), np.int32)
        lib().orc_schwarz_domains(self.h, self.offsets, self.indices, self.n_interior)
        self.indices = self.indices[:total]
        self.n_interior = self.n_interior[:self.n_domains]

    def apply(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        lib().orc_schwarz_apply(self.h, b, x)
        return x

    def __del__(self):
        try:
            lib().orc_schwarz_destroy(self.h)
        except Exception:
            pass


def dense_matvec(a, b, kernel, x, nthreads=None):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x2 = np.asarray(x, np.float64)
    if x2.ndim == 1:
        x2 = x2[:, None]
    xt = np.ascontiguousarray(x2.T)
    yt = np.zeros((x2.shape[1], a.shape[0]))
    kd = KernelC(kernel[0], 0, kernel[1], 0.0, 0.0, 0.0)
    lib().orc_dense_matvec(a.shape[1], a.shape[0], a.reshape(-1), b.shape[0], b.reshape(-1),
                           ctypes.byref(kd), xt.shape[0], xt.reshape(-1), xt.shape[1],
                           yt.reshape(-1), yt.shape[1], nthreads or nthreads_default())
    out = yt.T.copy()
    return out[:, 0] if np.asarray(x).ndim == 1 else out


class Octree:
    """HyperOctree restated on the CPU (dimension 1, 2 or 3, from pos)."""

    def __init__(self, low, high, pos, n_leaf):
        pos = np.ascontiguousarray(pos, np.float64)
        self.n, self.dim = pos.shape
        self.h = lib().orc_octree_create(self.dim, np.asarray(low, np.float64),
                                         np.asarray(high, np.float64), self.n,
                                         pos.reshape(-1), float(n_leaf))
        info = np.zeros(3, np.int64)
        lib().orc_octree_info(self.h, info)
        self.n_nodes, self.n_leaves, self.levels = (int(v) for v in info)
        self.order = np.zeros(self.n, np.int32)
        self.tags = np.zeros(self.n, np.int32)
        self.nodes = np.zeros(self.n_nodes, np.int32)
        self.leaves = np.zeros((self.n_leaves, 2), np.int32)
        lib().orc_octree_get(self.h, self.order, self.tags, self.nodes, self.leaves.reshape(-1))
        self.positions = pos[self.order]

    def interaction_lists(self):
        """The dual traversal's lists: dict with p2p_ptr, p2p_src (per
        leaf, source leaves), m2l_ptr, m2l_src (per node, source nodes) and
        the P2P particle-pair count."""
        counts = np.zeros(3, np.int64)
        if lib().orc_octree_lists(self.h, counts, None, None, None, None) != 0:
            raise ValueError("interaction lists failed")
        out = {"p2p_ptr": np.zeros(self.n_leaves + 1, np.int32),
               "p2p_src": np.zeros(int(counts[0]), np.int32),
               "m2l_ptr": np.zeros(self.n_nodes + 1, np.int32),
               "m2l_src": np.zeros(int(counts[1]), np.int32)}
        p = [out[k].ctypes.data_as(ctypes.c_void_p) for k in
             ("p2p_ptr", "p2p_src", "m2l_ptr", "m2l_src")]
        lib().orc_octree_lists(self.h, counts, *p)
        out["interactions"] = int(counts[2])
        return out

    def fmm_matvec(self, order, kernel, q):
        """y = K q (q in tree order), one right-hand side."""
        q = np.ascontiguousarray(q, np.float64)
        y = np.zeros_like(q)
        kd = KernelC(kernel[0], 0, kernel[1], 0.0, 0.0, 0.0)
        rc = lib().orc_fmm_matvec(self.h, int(order), ctypes.byref(kd), q, y)
        if rc != 0:
            raise ValueError("unsupported Chebyshev order %d" % order)
        return y

    def __del__(self):
        try:
            lib().orc_octree_destroy(self.h)
        except Exception:
            pass


def cheb_node(i, n):
    return lib().orc_cheb_node(i, n)


def cheb_Sn(x, i, n):
    return lib().orc_cheb_Sn(float(x), i, n)
