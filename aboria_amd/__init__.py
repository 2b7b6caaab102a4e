"""aboria_amd -- MI355X-native particle hot path behind Aboria's API names.

Python mirror of the reference's user-facing interface for the hot path:
``Particles.init_neighbour_search`` / ``update_positions`` / ``get_query``
(src/Particles.h:445-540), bulk ``euclidean_search`` (src/Search.h:839-845)
and ``create_sparse_operator`` (src/Operators.h:472-516).  Device memory and
streams are torch CUDA tensors; all compute goes through libaboria_hip.so.
"""
import ctypes

import numpy as np
import torch

from . import _capi as C
from ._capi import AbhError, check

__all__ = [
    "Context", "CellListOrdered", "Particles", "SearchResult", "euclidean_search",
    "Kernel", "SparseOperator", "create_sparse_operator", "AbhError",
    "FP64", "FP32_COMPENSATED", "MIXED", "FmmOperator", "create_fmm_operator",
    "DenseBatchOperator", "SchwartzPreconditioner", "knn", "Mlp", "ApproxInverse",
    "gmres_with_fallback", "dense_apply", "IdMap", "BlockKernel", "SparseBlockOperator",
    "HyperOctree",
]

FP64 = C.FP64
FP32_COMPENSATED = C.FP32_COMPENSATED
MIXED = C.MIXED


class _CAI:
    """__cuda_array_interface__ wrapper of a raw device pointer."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {
            "shape": tuple(shape), "typestr": typestr, "data": (int(ptr), False),
            "version": 2, "strides": None,
        }


_TYPESTR = {torch.float64: "<f8", torch.int32: "<i4", torch.int64: "<i8", torch.uint8: "|u1",
            torch.float32: "<f4"}
_UINT32 = "<u4"


def _view(ptr, shape, dtype, device):
    """Zero-copy torch view of library-owned device memory."""
    n = 1
    for s in shape:
        n *= s
    if n == 0 or not ptr:
        return torch.empty(shape, dtype=dtype, device=device)
    ts = _UINT32 if dtype == "uint32" else _TYPESTR[dtype]
    t = torch.as_tensor(_CAI(ptr, shape, ts), device=device)
    return t


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _require_cuda(t, dtype, name):
    if not (t.is_cuda and t.dtype == dtype and t.is_contiguous()):
        raise ValueError("%s must be a contiguous %s CUDA tensor" % (name, dtype))


class Context:
    """One HIP device + stream (torch's current stream by default)."""

    def __init__(self, device=0, stream=None):
        self.device = torch.device("cuda", device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.stream = stream
        h = ctypes.c_void_p()
        check(C.lib().abh_context_create(device, ctypes.c_void_p(stream.cuda_stream),
                                         ctypes.byref(h)))
        self.handle = h

    def sync(self):
        check(C.lib().abh_context_sync(self.handle), self.handle)

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CellListOrdered:
    """Sorted cell list (src/CellListOrdered.h).  Doubles as its own query
    object: bucket arrays and sorted positions live on the device."""

    def __init__(self, ctx, low, high, periodic, n_particles_in_leaf=10.0):
        self.ctx = ctx
        self.dim = len(low)
        lo = (ctypes.c_double * self.dim)(*[float(v) for v in low])
        hi = (ctypes.c_double * self.dim)(*[float(v) for v in high])
        per = (ctypes.c_uint8 * self.dim)(*[1 if p else 0 for p in periodic])
        h = ctypes.c_void_p()
        check(C.lib().abh_cells_create(ctx.handle, self.dim, lo, hi, per,
                                       float(n_particles_in_leaf), ctypes.byref(h)), ctx.handle)
        self.handle = h
        self.n_alive = 0
        self.id_map = None

    def find(self, ids):
        """CellListOrderedQuery::find: container indices of the particles with
        these ids (n_alive for a missing id); needs init_id_search on the
        particle set that owns this cell list."""
        if self.id_map is None:
            raise RuntimeError("no id map: call init_id_search on the particle set")
        return self.id_map.find(ids)

    def bucket_pairs(self):
        """get_neighbouring_buckets(query): the bucket_pair_iterator sequence
        as device tensors (bi, bj, image); the position offset of pair p is
        image[p] * (high - low)."""
        n = ctypes.c_int64()
        check(C.lib().abh_cells_bucket_pairs(self.handle, 0, None, None, None, ctypes.byref(n)),
              self.ctx.handle)
        dev = self.ctx.device
        bi = torch.empty(max(n.value, 1), dtype=torch.int32, device=dev)
        bj = torch.empty(max(n.value, 1), dtype=torch.int32, device=dev)
        im = torch.empty((max(n.value, 1), self.dim), dtype=torch.int8, device=dev)
        check(C.lib().abh_cells_bucket_pairs(self.handle, n.value, _ptr(bi), _ptr(bj), _ptr(im),
                                             ctypes.byref(n)), self.ctx.handle)
        return bi[:n.value], bj[:n.value], im[:n.value]

    def pair_counts(self, radius):
        """Neighbours of every particle (strict radius, itself included) by
        the symmetric bucket-pair traversal; needs bucket sides >= radius."""
        out = torch.empty(self.n_alive, dtype=torch.int32, device=self.ctx.device)
        check(C.lib().abh_cells_pair_counts(self.handle, float(radius), _ptr(out)),
              self.ctx.handle)
        return out

    def query_view(self):
        """The abh_query_view POD (ctypes) to pass to HIP kernels that use
        include/aboria_hip_device.hpp."""
        v = C.QueryView()
        check(C.lib().abh_cells_query_view(self.handle, ctypes.byref(v)))
        return v

    def update_positions(self, positions, alive=None):
        """Wraps/kills in place, sorts by bucket.  Returns the alive count."""
        _require_cuda(positions, torch.float64, "positions")
        if positions.dim() != 2 or positions.shape[1] != self.dim:
            raise ValueError("positions must be (n, %d)" % self.dim)
        if alive is not None:
            _require_cuda(alive, torch.uint8, "alive")
        na = ctypes.c_int64()
        check(C.lib().abh_cells_update(self.handle, positions.shape[0], _ptr(positions),
                                       _ptr(alive), ctypes.byref(na)), self.ctx.handle)
        self.n_alive = na.value
        return na.value

    @property
    def info(self):
        i = C.CellsInfo()
        check(C.lib().abh_cells_get_info(self.handle, ctypes.byref(i)))
        return i

    @property
    def grid(self):
        i = self.info
        return tuple(i.grid[d] for d in range(self.dim))

    @property
    def n_buckets(self):
        return self.info.n_buckets

    def order(self):
        return _view(C.lib().abh_cells_order(self.handle), (self.n_alive,), torch.int32,
                     self.ctx.device)

    def bucket_index(self):
        return _view(C.lib().abh_cells_bucket_index(self.handle), (self.n_alive,), "uint32",
                     self.ctx.device)

    def bucket_begin(self):
        return _view(C.lib().abh_cells_bucket_begin(self.handle), (self.n_buckets,), "uint32",
                     self.ctx.device)

    def bucket_end(self):
        return _view(C.lib().abh_cells_bucket_end(self.handle), (self.n_buckets,), "uint32",
                     self.ctx.device)

    def positions(self):
        return _view(C.lib().abh_cells_positions(self.handle), (self.n_alive, self.dim),
                     torch.float64, self.ctx.device)

    def gather(self, src, out=None):
        """out[k] = src[order[k]] for every alive particle."""
        if not (src.is_cuda and src.is_contiguous()):
            raise ValueError("src must be a contiguous CUDA tensor")
        shape = (self.n_alive,) + tuple(src.shape[1:])
        if out is None:
            out = torch.empty(shape, dtype=src.dtype, device=src.device)
        eb = src.element_size() * (src[0].numel() if src.dim() > 1 else 1)
        check(C.lib().abh_gather(self.ctx.handle, ctypes.c_void_p(
            C.lib().abh_cells_order(self.handle)), self.n_alive, _ptr(src), _ptr(out), eb),
            self.ctx.handle)
        return out

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_cells_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HyperOctree:
    """HyperOctree (src/OctTree.h:102-911) built on the device: tags, stable
    sort and the level-by-level node / leaf structure.  Doubles as its own
    query object (query_view() gives the HyperOctreeQuery POD)."""

    def __init__(self, ctx, low, high, periodic=None, n_particles_in_leaf=10.0):
        self.ctx = ctx
        self.dim = len(low)
        lo = (ctypes.c_double * self.dim)(*[float(v) for v in low])
        hi = (ctypes.c_double * self.dim)(*[float(v) for v in high])
        per = (ctypes.c_uint8 * self.dim)(*[1 if p else 0 for p in (periodic or [0] * self.dim)])
        h = ctypes.c_void_p()
        check(C.lib().abh_octree_create(ctx.handle, self.dim, lo, hi, per,
                                        float(n_particles_in_leaf), ctypes.byref(h)), ctx.handle)
        self.handle = h
        self.n_alive = 0

    def update_positions(self, positions, alive=None):
        """Wraps/kills in place like the cell list, then tags, sorts and
        builds the tree.  Returns the alive count."""
        _require_cuda(positions, torch.float64, "positions")
        if positions.dim() != 2 or positions.shape[1] != self.dim:
            raise ValueError("positions must be (n, %d)" % self.dim)
        if alive is not None:
            _require_cuda(alive, torch.uint8, "alive")
        na = ctypes.c_int64()
        check(C.lib().abh_octree_update(self.handle, positions.shape[0], _ptr(positions),
                                        _ptr(alive), ctypes.byref(na)), self.ctx.handle)
        self.n_alive = na.value
        return na.value

    @property
    def info(self):
        i = C.OctreeInfo()
        check(C.lib().abh_octree_get_info(self.handle, ctypes.byref(i)))
        return i

    def order(self):
        return _view(C.lib().abh_octree_order(self.handle), (self.n_alive,), torch.int32,
                     self.ctx.device)

    def positions(self):
        return _view(C.lib().abh_octree_positions(self.handle), (self.n_alive, self.dim),
                     torch.float64, self.ctx.device)

    def tags(self):
        return _view(C.lib().abh_octree_tags(self.handle), (self.n_alive,), "uint32",
                     self.ctx.device)

    def nodes(self):
        return _view(C.lib().abh_octree_nodes(self.handle), (self.info.n_nodes,), torch.int32,
                     self.ctx.device)

    def leaves(self):
        return _view(C.lib().abh_octree_leaves(self.handle), (self.info.n_leaves, 2),
                     torch.int32, self.ctx.device)

    def query_view(self):
        """The abh_octree_view POD (ctypes) for HIP kernels that use
        include/aboria_hip_device.hpp."""
        v = C.OctreeView()
        check(C.lib().abh_octree_query_view(self.handle, ctypes.byref(v)))
        return v

    def interaction_lists(self):
        """The dual traversal's lists: dict of device int32 tensors
        p2p_ptr, p2p_src (per non-empty leaf, its source leaves), m2l_ptr,
        m2l_src (per node, its M2L source nodes) and the P2P particle-pair
        count."""
        counts = (ctypes.c_int64 * 3)()
        check(C.lib().abh_octree_interaction_lists(self.handle, None, None, None, None, counts),
              self.ctx.handle)
        info = self.info
        dev = self.ctx.device
        out = {"p2p_ptr": torch.empty(info.n_leaves + 1, dtype=torch.int32, device=dev),
               "p2p_src": torch.empty(max(counts[0], 1), dtype=torch.int32, device=dev),
               "m2l_ptr": torch.empty(info.n_nodes + 1, dtype=torch.int32, device=dev),
               "m2l_src": torch.empty(max(counts[1], 1), dtype=torch.int32, device=dev)}
        check(C.lib().abh_octree_interaction_lists(
            self.handle, _ptr(out["p2p_ptr"]), _ptr(out["p2p_src"]), _ptr(out["m2l_ptr"]),
            _ptr(out["m2l_src"]), counts), self.ctx.handle)
        out["p2p_src"] = out["p2p_src"][:counts[0]]
        out["m2l_src"] = out["m2l_src"][:counts[1]]
        out["interactions"] = int(counts[2])
        return out

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_octree_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Particles:
    """SoA particle container (src/Particles.h) holding ``position``, ``id``,
    ``alive`` and user columns as CUDA tensors."""

    def __init__(self, ctx, positions, **columns):
        self.ctx = ctx
        self.position = positions.to(device=ctx.device, dtype=torch.float64).contiguous()
        n = self.position.shape[0]
        self.id = torch.arange(n, dtype=torch.int64, device=ctx.device)
        self.alive = torch.ones(n, dtype=torch.uint8, device=ctx.device)
        self.columns = {k: v.to(ctx.device).contiguous() for k, v in columns.items()}
        self.search = None
        self.id_map = None

    @property
    def dimension(self):
        return self.position.shape[1]

    def size(self):
        return self.position.shape[0]

    def init_neighbour_search(self, low, high, periodic, n_particles_in_leaf=10.0):
        self.search = CellListOrdered(self.ctx, low, high, periodic, n_particles_in_leaf)
        self.update_positions()

    def update_positions(self):
        """update_positions + reorder (src/Particles.h:526-540, 694-732)."""
        s = self.search
        s.update_positions(self.position, self.alive)
        self.position = s.positions().clone()
        self.id = s.gather(self.id)
        self.alive = torch.ones(s.n_alive, dtype=torch.uint8, device=self.ctx.device)
        self.columns = {k: s.gather(v) for k, v in self.columns.items()}
        if self.id_map is not None:
            self.id_map.update(self.id)
        s.id_map = self.id_map

    def init_id_search(self):
        """Particles::init_id_search (src/Particles.h:460-465): keeps a
        find-by-id map of this set, rebuilt by every update_positions."""
        if self.id_map is None:
            self.id_map = IdMap(self.ctx, self.id)
        else:
            self.id_map.update(self.id)
        if self.search is not None:
            self.search.id_map = self.id_map

    def find(self, ids):
        """Container index of the particle with each id, or size() when no
        particle has it (CellListOrderedQuery::find)."""
        if self.id_map is None:
            raise RuntimeError("init_id_search not called on this particle set")
        return self.id_map.find(ids)

    def get_query(self):
        if self.search is None:
            raise RuntimeError("init_neighbour_search not called on this particle set")
        return self.search

    def to_host(self):
        """position, id, alive and the user columns as numpy arrays (the
        input of aboria_amd.io's writers)."""
        return {"position": self.position.cpu().numpy(), "id": self.id.cpu().numpy(),
                "alive": self.alive.cpu().numpy(),
                "columns": {k: v.cpu().numpy() for k, v in self.columns.items()}}

    @classmethod
    def from_host(cls, ctx, position, id=None, alive=None, columns=None):
        """A particle set on the context's device from host arrays (what
        aboria_amd.io's readers return); ids are kept when given.  The arrays
        are copied (readers may hand over read-only buffers)."""
        p = cls(ctx, torch.from_numpy(np.array(position, dtype=np.float64, order="C")),
                **{k: torch.from_numpy(np.array(v, order="C")) for k, v in (columns or {}).items()})
        if id is not None:
            p.id = torch.from_numpy(np.array(id, dtype=np.int64)).to(ctx.device)
        if alive is not None:
            p.alive = torch.from_numpy(np.array(alive, dtype=np.uint8)).to(ctx.device)
        return p


class IdMap:
    """Find-by-id map: the (id, index) pairs of a particle set sorted by id
    (neighbour_search_base's m_id_map_key / m_id_map_value)."""

    def __init__(self, ctx, ids):
        _require_cuda(ids, torch.int64, "ids")
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(C.lib().abh_id_map_create(ctx.handle, ids.numel(), _ptr(ids), ctypes.byref(h)),
              ctx.handle)
        self.handle = h
        self.n = ids.numel()

    def update(self, ids):
        _require_cuda(ids, torch.int64, "ids")
        check(C.lib().abh_id_map_update(self.handle, ids.numel(), _ptr(ids)), self.ctx.handle)
        self.n = ids.numel()

    def keys(self):
        return _view(C.lib().abh_id_map_keys(self.handle), (self.n,), torch.int64,
                     self.ctx.device)

    def values(self):
        return _view(C.lib().abh_id_map_values(self.handle), (self.n,), torch.int32,
                     self.ctx.device)

    def find(self, ids):
        """(nq,) int32 container indices of the given ids; n for a missing id."""
        ids = torch.as_tensor(ids, dtype=torch.int64, device=self.ctx.device).contiguous()
        out = torch.empty(ids.shape, dtype=torch.int32, device=self.ctx.device)
        check(C.lib().abh_id_map_find(self.handle, ids.numel(), _ptr(ids), _ptr(out)),
              self.ctx.handle)
        return out

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_id_map_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SearchResult:
    """Neighbour lists in CSR form: row i is col[row_ptr[i]:row_ptr[i+1]]."""

    def __init__(self, row_ptr, col):
        self.row_ptr = row_ptr
        self.col = col

    @property
    def total(self):
        return self.col.numel()


class Search:
    """Reusable bulk search handle: ``run`` counts, ``fill`` writes indices
    into a caller buffer.  Device buffers persist across runs."""

    def __init__(self, query, centres, radius):
        _require_cuda(centres, torch.float64, "centres")
        self.ctx = query.ctx
        h = ctypes.c_void_p()
        check(C.lib().abh_search_create(self.ctx.handle, query.handle, centres.shape[0],
                                        _ptr(centres), float(radius), ctypes.byref(h)),
              self.ctx.handle)
        self.handle = h
        self.nq = centres.shape[0]

    def run(self, query, centres, radius):
        _require_cuda(centres, torch.float64, "centres")
        check(C.lib().abh_search_update(self.handle, query.handle, centres.shape[0],
                                        _ptr(centres), float(radius)), self.ctx.handle)
        self.nq = centres.shape[0]
        return self.total

    @property
    def total(self):
        t = ctypes.c_int64()
        check(C.lib().abh_search_total(self.handle, ctypes.byref(t)))
        return t.value

    def row_ptr(self):
        return _view(C.lib().abh_search_row_ptr(self.handle), (self.nq + 1,), torch.int64,
                     self.ctx.device)

    def fill(self, col):
        _require_cuda(col, torch.int32, "col")
        check(C.lib().abh_search_fill(self.handle, _ptr(col), col.numel()), self.ctx.handle)
        return col

    def run_fused(self, query, centres, radius, col):
        """Counts and fills in one traversal of the candidates (rows staged at
        upper-bound offsets in a handle-owned buffer, then packed into col).
        Returns the number of pairs; row_ptr() is valid afterwards."""
        _require_cuda(centres, torch.float64, "centres")
        _require_cuda(col, torch.int32, "col")
        t = ctypes.c_int64()
        check(C.lib().abh_search_run_fused(self.handle, query.handle, centres.shape[0],
                                           _ptr(centres), float(radius), _ptr(col), col.numel(),
                                           ctypes.byref(t)), self.ctx.handle)
        self.nq = centres.shape[0]
        return t.value

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_search_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def euclidean_search(query, centres, radius):
    """Bulk euclidean_search(query, centre, radius) for every row of centres
    (a (nq, dim) float64 CUDA tensor)."""
    _require_cuda(centres, torch.float64, "centres")
    ctx = query.ctx
    h = ctypes.c_void_p()
    check(C.lib().abh_search_create(ctx.handle, query.handle, centres.shape[0], _ptr(centres),
                                    float(radius), ctypes.byref(h)), ctx.handle)
    try:
        total = ctypes.c_int64()
        check(C.lib().abh_search_total(h, ctypes.byref(total)))
        rp = _view(C.lib().abh_search_row_ptr(h), (centres.shape[0] + 1,), torch.int64,
                   ctx.device).clone()
        col = torch.empty(total.value, dtype=torch.int32, device=ctx.device)
        check(C.lib().abh_search_fill(h, _ptr(col), total.value), ctx.handle)
        ctx.sync()
    finally:
        C.lib().abh_search_destroy(h)
    return SearchResult(rp, col)


class Kernel:
    """Kernel descriptor: phi(|dx|) * (w0 + w_row*a_i + w_col*b_j)."""

    def __init__(self, kind, param=0.0, w0=0.0, w_row=0.0, w_col=0.0):
        self.kind = kind
        self.param = float(param)
        self.w0, self.w_row, self.w_col = float(w0), float(w_row), float(w_col)

    def to_c(self):
        return C.Kernel(self.kind, 0, self.param, self.w0, self.w_row, self.w_col)

    @staticmethod
    def constant(**w):
        return Kernel(C.KERNEL_CONSTANT, 0.0, **w)

    @staticmethod
    def gaussian(scale, **w):
        return Kernel(C.KERNEL_GAUSSIAN, scale, **w)

    @staticmethod
    def wendland_c2(support, **w):
        return Kernel(C.KERNEL_WENDLAND_C2, support, **w)

    @staticmethod
    def multiquadric(c, **w):
        return Kernel(C.KERNEL_MULTIQUADRIC, c, **w)

    @staticmethod
    def smooth_inv_r(eps, **w):
        return Kernel(C.KERNEL_SMOOTH_INV_R, eps, **w)


class SparseOperator:
    """Matrix-free sparse kernel operator (KernelSparseConst)."""

    def __init__(self, ctx, cols, row_positions, radius, kernel, row_scalar=None,
                 col_scalar=None):
        _require_cuda(row_positions, torch.float64, "row_positions")
        self.ctx = ctx
        self.cols = cols
        self.nrows = row_positions.shape[0]
        self.ncols = cols.n_alive
        self._keep = (row_positions, row_scalar, col_scalar)
        kd = kernel.to_c()
        h = ctypes.c_void_p()
        check(C.lib().abh_sparse_op_create(ctx.handle, cols.handle, self.nrows,
                                           _ptr(row_positions), float(radius), ctypes.byref(kd),
                                           _ptr(row_scalar), _ptr(col_scalar), ctypes.byref(h)),
              ctx.handle)
        self.handle = h

    def apply(self, x, y, precision=FP64):
        """y += K x.  x: (ncols,) or (ncols, nrhs); y likewise with nrows."""
        x2 = x if x.dim() == 2 else x[:, None]
        y2 = y if y.dim() == 2 else y[:, None]
        # column-major storage: transpose views of (nrhs, n) contiguous tensors
        xt = x2.t().contiguous()
        yt = y2.t().contiguous()
        nrhs = xt.shape[0]
        check(C.lib().abh_sparse_op_apply(self.handle, nrhs, _ptr(xt), xt.shape[1], _ptr(yt),
                                          yt.shape[1], precision), self.ctx.handle)
        y2.copy_(yt.t())
        return y

    @property
    def nnz(self):
        n = ctypes.c_int64()
        check(C.lib().abh_sparse_op_nnz(self.handle, ctypes.byref(n)), self.ctx.handle)
        return n.value

    def cache(self, precision=FP64, enable=True):
        """Keep (or drop) the operator's CSR form for `precision` on the
        device; applies in that precision then stream it."""
        check(C.lib().abh_sparse_op_cache(self.handle, int(precision), int(bool(enable))),
              self.ctx.handle)
        return self

    def assemble(self):
        """The operator as CSR (row_ptr int64, col int32, val float64), the
        triplets of KernelSparse::assemble."""
        dev = self.ctx.device
        nnz = self.nnz
        rp = torch.empty(self.nrows + 1, dtype=torch.int64, device=dev)
        col = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        val = torch.empty(max(nnz, 1), dtype=torch.float64, device=dev)
        check(C.lib().abh_sparse_op_assemble(self.handle, _ptr(rp), _ptr(col), _ptr(val), nnz),
              self.ctx.handle)
        return rp, col[:nnz], val[:nnz]

    def apply_colmajor(self, xt, yt, precision=FP64):
        """y^T += (K x)^T on (nrhs, n) contiguous tensors, no copies."""
        check(C.lib().abh_sparse_op_apply(self.handle, xt.shape[0], _ptr(xt), xt.shape[1],
                                          _ptr(yt), yt.shape[1], precision), self.ctx.handle)
        return yt

    def __matmul__(self, x):
        shape = (self.nrows,) + tuple(x.shape[1:])
        y = torch.zeros(shape, dtype=torch.float64, device=x.device)
        return self.apply(x, y)

    def gmres(self, b, x=None, diag_shift=0.0, preconditioner=None, restart=50, max_iters=1000,
              tol=1e-8, precision=FP64, stall_restarts=0):
        """Solves (K + diag_shift I) x = b with restarted GMRES, right
        preconditioned by a SchwartzPreconditioner or an ApproxInverse (the
        reference's Eigen::GMRES<MatrixReplacement, SchwartzPreconditioner>).
        x (if given) is the initial guess; stall_restarts > 0 stops after that
        many restarts in a row that fail to cut the residual by 10%.  Returns
        (x, stats)."""
        if x is None:
            x = torch.zeros_like(b)
        sw = preconditioner.handle if isinstance(preconditioner, SchwartzPreconditioner) else None
        ai = preconditioner.handle if isinstance(preconditioner, ApproxInverse) else None
        if preconditioner is not None and sw is None and ai is None:
            raise TypeError("preconditioner must be a SchwartzPreconditioner or an ApproxInverse")
        opts = C.GmresOptions(int(restart), int(max_iters), float(tol), int(precision),
                              int(stall_restarts))
        st = C.GmresStats()
        check(C.lib().abh_sparse_op_gmres(self.handle, float(diag_shift), sw, ai, _ptr(b), _ptr(x),
                                          ctypes.byref(opts), ctypes.byref(st)), self.ctx.handle)
        return x, {"iterations": st.iterations, "converged": bool(st.converged),
                   "residual": st.residual}

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_sparse_op_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BlockKernel:
    """Matrix-valued kernel: entry (ii, jj) of the block of pair (i, j) is
    phi(|dx|) * (w0[ii][jj] + w_row[ii][jj] a_i + w_col[ii][jj] b_j), with
    phi the profile of a Kernel (its weights are ignored) and a, b the row and
    column scalars (zero when not given)."""

    def __init__(self, profile, w0, w_row=None, w_col=None):
        import numpy as np
        w0 = np.atleast_2d(np.asarray(w0, np.float64))
        self.block_rows, self.block_cols = w0.shape
        z = np.zeros_like(w0)
        self.profile = profile
        self.w = [w0, np.atleast_2d(np.asarray(w_row if w_row is not None else z, np.float64)),
                  np.atleast_2d(np.asarray(w_col if w_col is not None else z, np.float64))]
        for w in self.w:
            if w.shape != w0.shape:
                raise ValueError("weight blocks must have the same shape")

    def to_c(self):
        k = C.BlockKernel()
        k.type, k.block_rows, k.block_cols = self.profile.kind, self.block_rows, self.block_cols
        k.param = self.profile.param
        for arr, w in zip((k.w0, k.w_row, k.w_col), self.w):
            for e, v in enumerate(w.ravel()):
                arr[e] = float(v)
        return k


class SparseBlockOperator:
    """General sparse kernel operator: a radius per row (radius_function of
    create_sparse_operator) and/or a block-valued kernel."""

    def __init__(self, ctx, cols, row_positions, radius, kernel, row_radius=None,
                 row_scalar=None, col_scalar=None):
        _require_cuda(row_positions, torch.float64, "row_positions")
        if row_radius is not None:
            _require_cuda(row_radius, torch.float64, "row_radius")
        if not isinstance(kernel, BlockKernel):
            kernel = BlockKernel(kernel, [[1.0]])
        self.ctx = ctx
        self.kernel = kernel
        self.nrows = row_positions.shape[0]
        self.ncols = cols.n_alive
        self._keep = (row_positions, row_radius, row_scalar, col_scalar)
        kd = kernel.to_c()
        h = ctypes.c_void_p()
        check(C.lib().abh_sparse_block_op_create(
            ctx.handle, cols.handle, self.nrows, _ptr(row_positions), float(radius),
            _ptr(row_radius), ctypes.byref(kd), _ptr(row_scalar), _ptr(col_scalar),
            ctypes.byref(h)), ctx.handle)
        self.handle = h

    @property
    def pairs(self):
        n = ctypes.c_int64()
        check(C.lib().abh_sparse_block_op_pairs(self.handle, ctypes.byref(n)), self.ctx.handle)
        return n.value

    def apply(self, x, y):
        """y += K x; x (ncols * block_cols,), y (nrows * block_rows,)"""
        _require_cuda(x, torch.float64, "x")
        _require_cuda(y, torch.float64, "y")
        check(C.lib().abh_sparse_block_op_apply(self.handle, _ptr(x), _ptr(y)), self.ctx.handle)
        return y

    def __matmul__(self, x):
        y = torch.zeros(self.nrows * self.kernel.block_rows, dtype=torch.float64,
                        device=x.device)
        return self.apply(x.contiguous(), y)

    def assemble(self):
        """CSR (row_ptr int64, col int32, val float64) of the scalar matrix."""
        dev = self.ctx.device
        ne = self.pairs * self.kernel.block_rows * self.kernel.block_cols
        rp = torch.empty(self.nrows * self.kernel.block_rows + 1, dtype=torch.int64, device=dev)
        col = torch.empty(max(ne, 1), dtype=torch.int32, device=dev)
        val = torch.empty(max(ne, 1), dtype=torch.float64, device=dev)
        check(C.lib().abh_sparse_block_op_assemble(self.handle, _ptr(rp), _ptr(col), _ptr(val),
                                                   ne), self.ctx.handle)
        return rp, col[:ne], val[:ne]

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_sparse_block_op_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_sparse_operator(row_particles, col_particles, radius, kernel, row_scalar=None,
                           col_scalar=None):
    """create_sparse_operator(rows, cols, radius, f) -> SparseOperator."""
    rs = row_particles.columns[row_scalar] if isinstance(row_scalar, str) else row_scalar
    cs = col_particles.columns[col_scalar] if isinstance(col_scalar, str) else col_scalar
    if isinstance(kernel, BlockKernel) or callable(radius) or isinstance(radius, torch.Tensor):
        # a radius per row (a tensor, or a function of the row particle set)
        # and/or a block kernel: the general operator
        rr = radius(row_particles) if callable(radius) else radius
        if isinstance(rr, torch.Tensor):
            return SparseBlockOperator(col_particles.ctx, col_particles.get_query(),
                                       row_particles.position, 0.0, kernel,
                                       row_radius=rr.to(torch.float64).contiguous(),
                                       row_scalar=rs, col_scalar=cs)
        return SparseBlockOperator(col_particles.ctx, col_particles.get_query(),
                                   row_particles.position, float(rr), kernel, row_scalar=rs,
                                   col_scalar=cs)
    return SparseOperator(col_particles.ctx, col_particles.get_query(), row_particles.position,
                          radius, kernel, rs, cs)


def _train_flags(graph, fused):
    return (1 if graph else 0) | (2 if fused else 0)


class FmmOperator:
    """Black-box Chebyshev FMM operator over a HyperOctree in 1-3
    dimensions (positions (n, dim) float64; the M2L compression and the
    mixed-precision apply are 3D, other dimensions apply in FP64)."""

    def __init__(self, ctx, positions, low, high, n_particles_in_leaf, order, kernel):
        _require_cuda(positions, torch.float64, "positions")
        self.ctx = ctx
        self.n = positions.shape[0]
        self.dim = positions.shape[1] if positions.dim() == 2 else 1
        self._keep = positions
        lo = (ctypes.c_double * 3)(*[float(v) for v in low])
        hi = (ctypes.c_double * 3)(*[float(v) for v in high])
        kd = kernel.to_c()
        h = ctypes.c_void_p()
        check(C.lib().abh_fmm_create(ctx.handle, self.dim, lo, hi, self.n, _ptr(positions),
                                     float(n_particles_in_leaf), int(order), ctypes.byref(kd),
                                     ctypes.byref(h)), ctx.handle)
        self.handle = h

    @property
    def info(self):
        i = C.FmmInfo()
        check(C.lib().abh_fmm_get_info(self.handle, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in C.FmmInfo._fields_}

    def order(self):
        return _view(C.lib().abh_fmm_order(self.handle), (self.n,), torch.int32, self.ctx.device)

    def compress(self, eps):
        """Low-rank M2L operators (KernelH2::compress(tol)): singular values
        of each level's stacked operators below eps times the largest are
        dropped; eps <= 0 restores the full operators.  Returns (padded rank,
        [rank per M2L level])."""
        levels = self.info["levels"]
        kp = ctypes.c_int()
        ranks = (ctypes.c_int * max(1, levels - 1))()
        check(C.lib().abh_fmm_compress(self.handle, float(eps), ctypes.byref(kp), ranks),
              self.ctx.handle)
        return kp.value, [ranks[i] for i in range(max(0, levels - 1))] if kp.value else []

    def m2l_tuning(self):
        """(tiles per block chosen for the FP16 compressed M2L, [ms with 1,
        ms with 2]); 0 before the first MIXED apply after compress."""
        tt = ctypes.c_int()
        ms = (ctypes.c_float * 2)()
        check(C.lib().abh_fmm_m2l_tuning(self.handle, ctypes.byref(tt), ms), self.ctx.handle)
        return tt.value, [ms[0], ms[1]]

    def set_m2l_tuning(self, tiles_per_block):
        """Fix the FP16 compressed M2L's tiles per block (1, 2; 0 = retune)."""
        check(C.lib().abh_fmm_set_m2l_tuning(self.handle, int(tiles_per_block)),
              self.ctx.handle)

    def apply_colmajor(self, xt, yt, precision=FP64):
        """yt += (K x)^T for (nrhs, n) contiguous xt, yt."""
        check(C.lib().abh_fmm_apply(self.handle, xt.shape[0], _ptr(xt), xt.shape[1], _ptr(yt),
                                    yt.shape[1], precision), self.ctx.handle)
        return yt

    def __matmul__(self, x):
        x2 = x if x.dim() == 2 else x[:, None]
        xt = x2.t().contiguous()
        yt = torch.zeros_like(xt)
        self.apply_colmajor(xt, yt)
        y = yt.t()
        return y if x.dim() == 2 else y[:, 0].contiguous()

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_fmm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_fmm_operator(particles, order, kernel, low, high, n_particles_in_leaf=10.0):
    """create_fmm_operator<N>(particles, particles, f, f) on a HyperOctree."""
    return FmmOperator(particles.ctx, particles.position, low, high, n_particles_in_leaf, order,
                       kernel)


def dense_apply(ctx, row_positions, col_positions, kernel, x, y=None, precision=FP64):
    """y += K x for the rectangular dense operator create_dense_operator(rows,
    cols, f); x (ncols,) or (nrhs, ncols) (rows of x are right-hand sides)."""
    _require_cuda(row_positions, torch.float64, "row_positions")
    _require_cuda(col_positions, torch.float64, "col_positions")
    xt = x if x.dim() == 2 else x[None, :]
    xt = xt.contiguous()
    nrows = row_positions.shape[0]
    if y is None:
        y = torch.zeros((xt.shape[0], nrows), dtype=torch.float64, device=x.device)
    yt = y if y.dim() == 2 else y[None, :]
    kd = kernel.to_c()
    check(C.lib().abh_dense_apply(ctx.handle, row_positions.shape[1], nrows, _ptr(row_positions),
                                  col_positions.shape[0], _ptr(col_positions), ctypes.byref(kd),
                                  xt.shape[0], _ptr(xt), xt.shape[1], _ptr(yt), yt.shape[1],
                                  precision), ctx.handle)
    return y


class DenseBatchOperator:
    """Batch of independent dense kernel operators (create_dense_operator per
    particle set), evaluated and solved together."""

    def __init__(self, ctx, offsets, positions, kernel):
        _require_cuda(positions, torch.float64, "positions")
        self.ctx = ctx
        self.offsets = [int(v) for v in offsets]
        self.nprob = len(self.offsets) - 1
        self.dim = positions.shape[1]
        self._keep = positions
        off = (ctypes.c_int64 * len(self.offsets))(*self.offsets)
        kd = kernel.to_c()
        h = ctypes.c_void_p()
        check(C.lib().abh_dense_batch_create(ctx.handle, self.dim, self.nprob, off,
                                             _ptr(positions), ctypes.byref(kd), ctypes.byref(h)),
              ctx.handle)
        self.handle = h

    def apply(self, x, y, precision=FP64):
        """y += K x (concatenated vectors)."""
        check(C.lib().abh_dense_batch_apply(self.handle, _ptr(x), _ptr(y), precision),
              self.ctx.handle)
        return y

    def __matmul__(self, x):
        y = torch.zeros_like(x)
        return self.apply(x, y)

    def tune(self, precision=FP64, force=0):
        """Runs the tile-shape autotuner (or sets rows-per-thread `force`);
        returns (rows_per_thread, {rpt: ms})."""
        rpt = ctypes.c_int(int(force))
        ms = (ctypes.c_float * 3)()
        check(C.lib().abh_dense_batch_tune(self.handle, precision, ctypes.byref(rpt), ms),
              self.ctx.handle)
        return rpt.value, {1: ms[0], 2: ms[1], 4: ms[2]}

    def gmres(self, b, x=None, restart=50, max_iters=1000, tol=1e-8, precision=FP64,
              preconditioner=None, stall_restarts=0, graph=False):
        """Solves K x = b for every problem, right-preconditioned by
        `preconditioner` (SchwartzPreconditioner.for_dense_batch) if given; a
        problem stops after stall_restarts (> 0) restarts in a row that fail
        to cut its residual by 10%.  graph: each restart cycle's Arnoldi
        steps run as one replayed HIP graph.  Returns (x, stats)."""
        if x is None:
            x = torch.zeros_like(b)
        opts = C.GmresOptions(int(restart), int(max_iters), float(tol), int(precision),
                              int(stall_restarts), int(bool(graph)))
        stats = (C.GmresStats * self.nprob)()
        ph = preconditioner.handle if preconditioner is not None else None
        check(C.lib().abh_dense_batch_gmres_precond(self.handle, ph, _ptr(b), _ptr(x),
                                                    ctypes.byref(opts), stats), self.ctx.handle)
        return x, [{"iterations": s.iterations, "converged": bool(s.converged),
                    "residual": s.residual} for s in stats]

    def residual_dd(self, b, x, r=None):
        """r = b - K x in double-double arithmetic, rounded once to FP64."""
        if r is None:
            r = torch.empty_like(b)
        check(C.lib().abh_dense_batch_residual_dd(self.handle, _ptr(b), _ptr(x), _ptr(r)),
              self.ctx.handle)
        return r

    def refine(self, b, x=None, tol=1e-8, max_outer=8, inner_tol=1e-3, restart=50,
               inner_iters=100, precision=FP64, preconditioner=None, stall_restarts=0,
               graph=False):
        """Iterative refinement: the residual in double-double arithmetic,
        corrections from preconditioned GMRES (inner_tol, inner_iters), at
        most max_outer residual evaluations.  Returns (x, stats) with the
        inner iterations summed and the double-double relative residual."""
        if x is None:
            x = torch.zeros_like(b)
        opts = C.GmresOptions(int(restart), int(inner_iters), float(inner_tol), int(precision),
                              int(stall_restarts), int(bool(graph)))
        stats = (C.GmresStats * self.nprob)()
        ph = preconditioner.handle if preconditioner is not None else None
        check(C.lib().abh_dense_batch_refine(self.handle, ph, _ptr(b), _ptr(x),
                                             ctypes.byref(opts), float(tol), int(max_outer),
                                             stats), self.ctx.handle)
        return x, [{"iterations": s.iterations, "converged": bool(s.converged),
                    "residual": s.residual} for s in stats]

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_dense_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SchwartzPreconditioner:
    """Restricted additive Schwarz preconditioner of A = K + diag_shift I over
    a cell list (SchwartzPreconditioner<Eigen::LLT<MatrixXd>>,
    src/Preconditioners.h:1121-1537): analyzePattern + factorize happen at
    construction, solve(b) is _solve_impl.  The kernel is restricted to
    |dx| <= radius (radius <= 0: every pair).  coarse_grid_n > 0 adds the
    coarse-grid domain of set_coarse_grid_n (particles sampled near an
    (n+1)^dim lattice with a seeded shuffle, solved last)."""

    def __init__(self, ctx, cells, kernel, radius, diag_shift=0.0, coarse_grid_n=0, seed=0):
        self.ctx = ctx
        self.cells = cells
        self.n = cells.n_alive
        kd = kernel.to_c()
        h = ctypes.c_void_p()
        check(C.lib().abh_schwarz_create_coarse(ctx.handle, cells.handle, ctypes.byref(kd),
                                                float(radius), float(diag_shift),
                                                int(coarse_grid_n), ctypes.c_uint64(seed),
                                                ctypes.byref(h)),
              ctx.handle)
        self.handle = h

    def factorize(self, kernel, radius, diag_shift=0.0):
        """factorize() alone on the analysed domains: new kernel values
        (kernel, radius, diagonal shift) for the same particles."""
        if self.cells is None:
            raise ValueError("refactorisation is for preconditioners built on a cell list")
        kd = kernel.to_c()
        check(C.lib().abh_schwarz_factorize(self.handle, self.cells.handle, ctypes.byref(kd),
                                            float(radius), float(diag_shift)), self.ctx.handle)

    @classmethod
    def for_dense_batch(cls, op, low, high, n_particles_in_leaf, diag_shift=0.0):
        """One restricted additive Schwarz preconditioner per problem of a
        DenseBatchOperator (K + diag_shift I over all pairs), with the bucket
        grid a cell list over [low, high] would size for each problem and a
        signed local factorisation (indefinite kernels allowed)."""
        self = cls.__new__(cls)
        self.ctx = op.ctx
        self.cells = None
        self.op = op
        self.n = op.offsets[-1]
        lo = (ctypes.c_double * op.dim)(*[float(v) for v in low])
        hi = (ctypes.c_double * op.dim)(*[float(v) for v in high])
        h = ctypes.c_void_p()
        check(C.lib().abh_dense_batch_schwarz_create(op.handle, lo, hi,
                                                     float(n_particles_in_leaf),
                                                     float(diag_shift), ctypes.byref(h)),
              op.ctx.handle)
        self.handle = h
        return self

    @property
    def info(self):
        i = C.SchwarzInfo()
        check(C.lib().abh_schwarz_get_info(self.handle, ctypes.byref(i)), self.ctx.handle)
        return {f: getattr(i, f) for f, _ in C.SchwarzInfo._fields_}

    def domains(self):
        """(offsets, indices, n_interior) as host int arrays."""
        import numpy as np
        inf = self.info
        nd = inf["n_domains"]
        off = np.zeros(nd + 1, np.int64)
        check(C.lib().abh_schwarz_get_domains(self.handle, off.ctypes.data, None, None),
              self.ctx.handle)
        idx = np.zeros(max(int(off[-1]), 1), np.int32)
        nint = np.zeros(max(nd, 1), np.int32)
        check(C.lib().abh_schwarz_get_domains(self.handle, None, idx.ctypes.data,
                                              nint.ctypes.data), self.ctx.handle)
        return off, idx[:int(off[-1])], nint[:nd]

    def apply(self, b, x):
        check(C.lib().abh_schwarz_apply(self.handle, _ptr(b), _ptr(x)), self.ctx.handle)
        return x

    def solve(self, b):
        return self.apply(b, torch.empty_like(b))

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_schwarz_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def knn(cells, k, with_dist2=False):
    """k nearest particles of every particle of the cell list (itself first),
    as an (n, k) int32 tensor (and the squared distances)."""
    ctx = cells.ctx
    n = cells.n_alive
    nbr = torch.empty((n, k), dtype=torch.int32, device=ctx.device)
    d2 = torch.empty((n, k), dtype=torch.float64, device=ctx.device) if with_dist2 else None
    check(C.lib().abh_knn(ctx.handle, cells.handle, int(k), _ptr(nbr), _ptr(d2)), ctx.handle)
    return (nbr, d2) if with_dist2 else nbr


class Mlp:
    """n_in -> hidden x n_hidden (SiLU) -> n_out network with hand-written
    FP32 training kernels and a BF16 MFMA inference kernel."""

    def __init__(self, ctx, n_in, hidden=128, n_hidden=3, n_out=32, seed=0):
        self.ctx = ctx
        self.dims = [n_in] + [hidden] * n_hidden + [n_out]
        h = ctypes.c_void_p()
        check(C.lib().abh_mlp_create(ctx.handle, int(n_in), int(hidden), int(n_hidden), int(n_out),
                                     int(seed), ctypes.byref(h)), ctx.handle)
        self.handle = h
        n = ctypes.c_int64()
        check(C.lib().abh_mlp_num_params(h, ctypes.byref(n)), ctx.handle)
        self.num_params = n.value

    def get_params(self):
        import numpy as np
        out = np.zeros(self.num_params, np.float32)
        check(C.lib().abh_mlp_get_params(self.handle, out.ctypes.data), self.ctx.handle)
        return out

    def set_params(self, values):
        import numpy as np
        v = np.ascontiguousarray(values, np.float32)
        if v.size != self.num_params:
            raise ValueError("expected %d parameters" % self.num_params)
        check(C.lib().abh_mlp_set_params(self.handle, v.ctypes.data), self.ctx.handle)

    def forward(self, x, bf16=False):
        _require_cuda(x, torch.float32, "x")
        y = torch.empty((x.shape[0], self.dims[-1]), dtype=torch.float32, device=x.device)
        check(C.lib().abh_mlp_forward(self.handle, x.shape[0], _ptr(x), _ptr(y), int(bool(bf16))),
              self.ctx.handle)
        return y

    def loss_grad(self, x, t):
        """(mean squared error, gradient tensor of num_params FP32)."""
        _require_cuda(x, torch.float32, "x")
        _require_cuda(t, torch.float32, "t")
        g = torch.empty(self.num_params, dtype=torch.float32, device=x.device)
        loss = ctypes.c_double()
        check(C.lib().abh_mlp_loss_grad(self.handle, x.shape[0], _ptr(x), _ptr(t),
                                        ctypes.byref(loss), _ptr(g)), self.ctx.handle)
        return loss.value, g

    def train(self, x, t, epochs=20, batch=1024, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8,
              lr_decay=1.0, seed=0, graph=False, fused=False):
        """Adam on the mean squared error; returns the per-epoch mean loss.
        graph: the steps replay one captured HIP graph; fused: a step runs as
        one kernel per 8 rows through the whole network plus one update."""
        import numpy as np
        _require_cuda(x, torch.float32, "x")
        _require_cuda(t, torch.float32, "t")
        o = C.MlpTrainOptions(int(epochs), int(batch), float(lr), float(beta1), float(beta2),
                              float(eps), float(lr_decay), _train_flags(graph, fused),
                              int(seed))
        losses = np.zeros(max(int(epochs), 1), np.float32)
        check(C.lib().abh_mlp_train(self.handle, x.shape[0], _ptr(x), _ptr(t), ctypes.byref(o),
                                    losses.ctypes.data), self.ctx.handle)
        return losses[:int(epochs)]

    def _check_mats(self, x, mats):
        _require_cuda(x, torch.float32, "x")
        _require_cuda(mats, torch.float32, "mats")
        k = self.dims[-1]
        if tuple(mats.shape) != (x.shape[0], k, k):
            raise ValueError("mats must be (n, %d, %d)" % (k, k))

    def loss_grad_residual(self, x, mats):
        """(mean patch residual |M y - e_1|^2, gradient of num_params FP32)."""
        self._check_mats(x, mats)
        g = torch.empty(self.num_params, dtype=torch.float32, device=x.device)
        loss = ctypes.c_double()
        check(C.lib().abh_mlp_loss_grad_residual(self.handle, x.shape[0], _ptr(x), _ptr(mats),
                                                 ctypes.byref(loss), _ptr(g)), self.ctx.handle)
        return loss.value, g

    def train_residual(self, x, mats, epochs=20, batch=1024, lr=1e-3, beta1=0.9, beta2=0.999,
                       eps=1e-8, lr_decay=1.0, seed=0, graph=False, fused=False):
        """Adam on the patch-residual loss; returns the per-epoch mean loss.
        graph / fused as for train()."""
        import numpy as np
        self._check_mats(x, mats)
        o = C.MlpTrainOptions(int(epochs), int(batch), float(lr), float(beta1), float(beta2),
                              float(eps), float(lr_decay), _train_flags(graph, fused),
                              int(seed))
        losses = np.zeros(max(int(epochs), 1), np.float32)
        check(C.lib().abh_mlp_train_residual(self.handle, x.shape[0], _ptr(x), _ptr(mats),
                                             ctypes.byref(o), losses.ctypes.data),
              self.ctx.handle)
        return losses[:int(epochs)]

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ApproxInverse:
    """Approximate-inverse rows on the k-nearest-neighbour pattern of a cell
    list: x_i = sum_k w_ik b_nbr_ik.  Built either from least squares on the
    operator (lsq) or by the network (learned)."""

    def __init__(self, ctx, handle, n, k, dim):
        self.ctx = ctx
        self.handle = handle
        self.n, self.k, self.dim = n, k, dim

    @classmethod
    def lsq(cls, ctx, cells, kernel, radius, diag_shift, k=32):
        kd = kernel.to_c()
        h = ctypes.c_void_p()
        check(C.lib().abh_ainv_create_lsq(ctx.handle, cells.handle, ctypes.byref(kd),
                                          float(radius), float(diag_shift), int(k),
                                          ctypes.byref(h)), ctx.handle)
        return cls(ctx, h, cells.n_alive, k, cells.info.dim)

    @classmethod
    def patch(cls, ctx, cells, kernel, radius, diag_shift, k=32):
        """Exact local solves of every particle's k-nearest-neighbour patch."""
        kd = kernel.to_c()
        h = ctypes.c_void_p()
        check(C.lib().abh_ainv_create_patch(ctx.handle, cells.handle, ctypes.byref(kd),
                                            float(radius), float(diag_shift), int(k),
                                            ctypes.byref(h)), ctx.handle)
        return cls(ctx, h, cells.n_alive, k, cells.info.dim)

    def patch_matrices(self, scale=1.0):
        """The (n, k, k) FP32 patch matrices times scale."""
        out = torch.empty((self.n, self.k, self.k), dtype=torch.float32, device=self.ctx.device)
        check(C.lib().abh_ainv_patch_matrices(self.handle, float(scale), _ptr(out)),
              self.ctx.handle)
        return out

    @classmethod
    def learned(cls, ctx, cells, mlp, k, feature_scale, out_scale, bf16=True):
        h = ctypes.c_void_p()
        check(C.lib().abh_ainv_create_learned(ctx.handle, cells.handle, mlp.handle, int(k),
                                              float(feature_scale), float(out_scale),
                                              int(bool(bf16)), ctypes.byref(h)), ctx.handle)
        return cls(ctx, h, cells.n_alive, k, cells.info.dim)

    def pattern(self):
        out = torch.empty((self.n, self.k), dtype=torch.int32, device=self.ctx.device)
        check(C.lib().abh_ainv_get(self.handle, _ptr(out), None), self.ctx.handle)
        return out

    def weights(self):
        out = torch.empty((self.n, self.k), dtype=torch.float64, device=self.ctx.device)
        check(C.lib().abh_ainv_get(self.handle, None, _ptr(out)), self.ctx.handle)
        return out

    def features(self, feature_scale):
        out = torch.empty((self.n, self.k * self.dim), dtype=torch.float32,
                          device=self.ctx.device)
        check(C.lib().abh_ainv_features(self.handle, float(feature_scale), _ptr(out)),
              self.ctx.handle)
        return out

    def apply(self, b, x):
        check(C.lib().abh_ainv_apply(self.handle, _ptr(b), _ptr(x)), self.ctx.handle)
        return x

    def solve(self, b):
        return self.apply(b, torch.empty_like(b))

    def close(self):
        if getattr(self, "handle", None):
            C.lib().abh_ainv_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gmres_with_fallback(op, b, diag_shift, learned, schwarz, restart=50, max_iters=2000,
                        learned_iters=100, tol=1e-8):
    """GMRES preconditioned by the learned approximate inverse for at most
    learned_iters iterations; if that has not converged, continues from its
    iterate with the plain Schwarz preconditioner (the fallback).  Returns
    (x, report)."""
    x = torch.zeros_like(b)
    rep = {"learned": None, "fallback": None}
    if learned is not None:
        x, st = op.gmres(b, x, diag_shift, learned, restart, learned_iters, tol)
        rep["learned"] = st
        if st["converged"]:
            return x, rep
    x, st = op.gmres(b, x, diag_shift, schwarz, restart, max_iters, tol)
    rep["fallback"] = st
    return x, rep
