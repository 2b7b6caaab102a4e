"""ctypes declarations of libaboria_hip.so (include/aboria_hip.h).

The library is loaded from this package directory only; there is no
fallback implementation -- a missing library raises at import of the
first object that needs it.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaboria_hip.so")

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t
P = ctypes.POINTER

ABH_OK = 0
STATUS_NAMES = {
    0: "ABH_OK",
    1: "ABH_ERR_INVALID_ARG",
    2: "ABH_ERR_HIP",
    3: "ABH_ERR_OUT_OF_MEMORY",
    4: "ABH_ERR_CAPACITY",
    5: "ABH_ERR_UNSUPPORTED",
    6: "ABH_ERR_DOMAIN",
    7: "ABH_ERR_NOT_CONVERGED",
    8: "ABH_ERR_NUMERICAL",
}

KERNEL_CONSTANT = 0
KERNEL_GAUSSIAN = 1
KERNEL_WENDLAND_C2 = 2
KERNEL_MULTIQUADRIC = 3
KERNEL_SMOOTH_INV_R = 4

FP64 = 0
FP32_COMPENSATED = 1


class CellsInfo(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("grid", ctypes.c_uint32 * 3),
        ("side", ctypes.c_double * 3),
        ("low", ctypes.c_double * 3),
        ("high", ctypes.c_double * 3),
        ("periodic", ctypes.c_uint8 * 3),
        ("n_buckets", ctypes.c_int64),
        ("n_alive", ctypes.c_int64),
    ]


class Kernel(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("param", ctypes.c_double),
        ("w0", ctypes.c_double),
        ("w_row", ctypes.c_double),
        ("w_col", ctypes.c_double),
    ]


class BlockKernel(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("block_rows", ctypes.c_int32),
                ("block_cols", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("param", ctypes.c_double), ("w0", ctypes.c_double * 16),
                ("w_row", ctypes.c_double * 16), ("w_col", ctypes.c_double * 16)]


class QueryView(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("periodic", ctypes.c_int32 * 3),
                ("size", ctypes.c_int32 * 3), ("end_bucket", ctypes.c_int32 * 3),
                ("low", ctypes.c_double * 3), ("high", ctypes.c_double * 3),
                ("side", ctypes.c_double * 3), ("inv_side", ctypes.c_double * 3),
                ("pos", ctypes.c_void_p), ("bucket_begin", ctypes.c_void_p),
                ("bucket_end", ctypes.c_void_p), ("n", ctypes.c_int64),
                ("n_buckets", ctypes.c_int64)]


class OctreeInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("levels", ctypes.c_int32), ("n_alive", ctypes.c_int64),
                ("n_nodes", ctypes.c_int64), ("n_leaves", ctypes.c_int64)]


class OctreeView(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("levels", ctypes.c_int32),
                ("periodic", ctypes.c_int32 * 3), ("reserved", ctypes.c_int32),
                ("low", ctypes.c_double * 3), ("high", ctypes.c_double * 3),
                ("nodes", ctypes.c_void_p), ("leaves", ctypes.c_void_p), ("pos", ctypes.c_void_p),
                ("order", ctypes.c_void_p), ("n", ctypes.c_int64), ("n_nodes", ctypes.c_int64),
                ("n_leaves", ctypes.c_int64)]


class FmmInfo(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int32),
        ("levels", ctypes.c_int32),
        ("n_nodes", ctypes.c_int64),
        ("n_leaves", ctypes.c_int64),
        ("m2l_pairs", ctypes.c_int64),
        ("p2p_pairs", ctypes.c_int64),
        ("p2p_interactions", ctypes.c_int64),
    ]


MIXED = 2


class GmresOptions(ctypes.Structure):
    _fields_ = [("restart", ctypes.c_int32), ("max_iters", ctypes.c_int32), ("tol", ctypes.c_double),
                ("precision", ctypes.c_int32), ("stall_restarts", ctypes.c_int32),
                ("graph", ctypes.c_int32)]


class GmresStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("residual", ctypes.c_double)]


class SchwarzInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("n_domains", ctypes.c_int64), ("min_size", ctypes.c_int64),
                ("max_size", ctypes.c_int64), ("min_interior", ctypes.c_int64),
                ("max_interior", ctypes.c_int64), ("factor_bytes", ctypes.c_int64),
                ("failed_domains", ctypes.c_int64)]


class MlpTrainOptions(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("lr", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("lr_decay", ctypes.c_float), ("flags", ctypes.c_int32),
                ("seed", ctypes.c_uint64)]


class AbhError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, str(status)), message))


# name -> (restype, argtypes)
SIGNATURES = {
    "abh_context_create": (c_int, [c_int, c_void_p, P(c_void_p)]),
    "abh_context_destroy": (c_int, [c_void_p]),
    "abh_context_sync": (c_int, [c_void_p]),
    "abh_status_string": (ctypes.c_char_p, [c_int]),
    "abh_last_error": (ctypes.c_char_p, [c_void_p]),
    "abh_malloc": (c_int, [c_void_p, P(c_void_p), c_size_t]),
    "abh_free": (c_int, [c_void_p, c_void_p]),
    "abh_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "abh_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "abh_memcpy_d2d": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "abh_memset": (c_int, [c_void_p, c_void_p, c_int, c_size_t]),
    "abh_cells_create": (c_int, [c_void_p, c_int, P(c_double), P(c_double), P(ctypes.c_uint8),
                                 c_double, P(c_void_p)]),
    "abh_cells_destroy": (c_int, [c_void_p]),
    "abh_cells_update": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, P(c_int64)]),
    "abh_cells_get_info": (c_int, [c_void_p, P(CellsInfo)]),
    "abh_cells_order": (c_void_p, [c_void_p]),
    "abh_cells_bucket_index": (c_void_p, [c_void_p]),
    "abh_cells_bucket_begin": (c_void_p, [c_void_p]),
    "abh_cells_bucket_end": (c_void_p, [c_void_p]),
    "abh_cells_positions": (c_void_p, [c_void_p]),
    "abh_gather": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_size_t]),
    "abh_search_create": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_double, P(c_void_p)]),
    "abh_search_update": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_double]),
    "abh_search_destroy": (c_int, [c_void_p]),
    "abh_search_total": (c_int, [c_void_p, P(c_int64)]),
    "abh_search_row_ptr": (c_void_p, [c_void_p]),
    "abh_search_fill": (c_int, [c_void_p, c_void_p, c_int64]),
    "abh_search_run_fused": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_double, c_void_p,
                                     c_int64, P(c_int64)]),
    "abh_sparse_op_create": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_double, P(Kernel),
                                     c_void_p, c_void_p, P(c_void_p)]),
    "abh_sparse_op_destroy": (c_int, [c_void_p]),
    "abh_sparse_op_nnz": (c_int, [c_void_p, P(c_int64)]),
    "abh_sparse_op_assemble": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64]),
    "abh_sparse_op_apply": (c_int, [c_void_p, c_int, c_void_p, c_int64, c_void_p, c_int64, c_int]),
    "abh_sparse_op_cache": (c_int, [c_void_p, c_int, c_int]),
    "abh_octree_create": (c_int, [c_void_p, c_int, P(c_double), P(c_double), P(ctypes.c_uint8),
                                  c_double, P(c_void_p)]),
    "abh_octree_destroy": (c_int, [c_void_p]),
    "abh_octree_update": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, P(c_int64)]),
    "abh_octree_get_info": (c_int, [c_void_p, P(OctreeInfo)]),
    "abh_octree_order": (c_void_p, [c_void_p]),
    "abh_octree_positions": (c_void_p, [c_void_p]),
    "abh_octree_tags": (c_void_p, [c_void_p]),
    "abh_octree_nodes": (c_void_p, [c_void_p]),
    "abh_octree_leaves": (c_void_p, [c_void_p]),
    "abh_octree_query_view": (c_int, [c_void_p, P(OctreeView)]),
    "abh_octree_interaction_lists": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                             P(c_int64)]),
    "abh_ainv_create_patch": (c_int, [c_void_p, c_void_p, P(Kernel), c_double, c_double, c_int,
                                      P(c_void_p)]),
    "abh_ainv_patch_matrices": (c_int, [c_void_p, c_double, c_void_p]),
    "abh_mlp_loss_grad_residual": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, P(c_double),
                                           c_void_p]),
    "abh_mlp_train_residual": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, P(MlpTrainOptions),
                                       c_void_p]),
    "abh_fmm_create": (c_int, [c_void_p, c_int, P(c_double), P(c_double), c_int64, c_void_p,
                               c_double, c_int, P(Kernel), P(c_void_p)]),
    "abh_fmm_destroy": (c_int, [c_void_p]),
    "abh_fmm_get_info": (c_int, [c_void_p, P(FmmInfo)]),
    "abh_fmm_order": (c_void_p, [c_void_p]),
    "abh_fmm_compress": (c_int, [c_void_p, c_double, P(c_int), c_void_p]),
    "abh_fmm_apply": (c_int, [c_void_p, c_int, c_void_p, c_int64, c_void_p, c_int64, c_int]),
    "abh_dense_batch_create": (c_int, [c_void_p, c_int, c_int, P(c_int64), c_void_p, P(Kernel),
                                       P(c_void_p)]),
    "abh_dense_batch_destroy": (c_int, [c_void_p]),
    "abh_dense_batch_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    "abh_dense_batch_residual_dd": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "abh_dense_batch_refine": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, P(GmresOptions),
                                       c_double, c_int, c_void_p]),
    "abh_dense_batch_gmres": (c_int, [c_void_p, c_void_p, c_void_p, P(GmresOptions),
                                      P(GmresStats)]),
    "abh_dense_apply": (c_int, [c_void_p, c_int, c_int64, c_void_p, c_int64, c_void_p, P(Kernel),
                                c_int, c_void_p, c_int64, c_void_p, c_int64, c_int]),
    "abh_dense_apply_noalloc": (c_int, [c_void_p, c_int, c_int64, c_void_p, c_int64, c_void_p,
                                        P(Kernel), c_int, c_void_p, c_int64, c_void_p, c_int64,
                                        c_int]),
    "abh_cells_bucket_pairs": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                       P(c_int64)]),
    "abh_cells_pair_counts": (c_int, [c_void_p, c_double, c_void_p]),
    "abh_cells_query_view": (c_int, [c_void_p, P(QueryView)]),
    "abh_sparse_block_op_create": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_double,
                                           c_void_p, P(BlockKernel), c_void_p, c_void_p,
                                           P(c_void_p)]),
    "abh_sparse_block_op_destroy": (c_int, [c_void_p]),
    "abh_sparse_block_op_pairs": (c_int, [c_void_p, P(c_int64)]),
    "abh_sparse_block_op_apply": (c_int, [c_void_p, c_void_p, c_void_p]),
    "abh_sparse_block_op_assemble": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64]),
    "abh_id_map_create": (c_int, [c_void_p, c_int64, c_void_p, P(c_void_p)]),
    "abh_id_map_update": (c_int, [c_void_p, c_int64, c_void_p]),
    "abh_id_map_destroy": (c_int, [c_void_p]),
    "abh_id_map_size": (c_int, [c_void_p, P(c_int64)]),
    "abh_id_map_keys": (c_void_p, [c_void_p]),
    "abh_id_map_values": (c_void_p, [c_void_p]),
    "abh_id_map_find": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
    "abh_dense_batch_tune": (c_int, [c_void_p, c_int, P(c_int), P(ctypes.c_float)]),
    "abh_schwarz_create": (c_int, [c_void_p, c_void_p, P(Kernel), c_double, c_double,
                                   P(c_void_p)]),
    "abh_fmm_m2l_tuning": (c_int, [c_void_p, P(c_int), P(ctypes.c_float)]),
    "abh_fmm_set_m2l_tuning": (c_int, [c_void_p, c_int]),
    "abh_schwarz_create_coarse": (c_int, [c_void_p, c_void_p, P(Kernel), c_double, c_double,
                                          c_int, ctypes.c_uint64, P(c_void_p)]),
    "abh_schwarz_factorize": (c_int, [c_void_p, c_void_p, P(Kernel), c_double, c_double]),
    "abh_schwarz_destroy": (c_int, [c_void_p]),
    "abh_schwarz_get_info": (c_int, [c_void_p, P(SchwarzInfo)]),
    "abh_schwarz_get_domains": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "abh_schwarz_apply": (c_int, [c_void_p, c_void_p, c_void_p]),
    "abh_dense_batch_schwarz_create": (c_int, [c_void_p, P(c_double), P(c_double), c_double,
                                               c_double, P(c_void_p)]),
    "abh_dense_batch_gmres_precond": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p,
                                              P(GmresOptions), P(GmresStats)]),
    "abh_knn": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "abh_mlp_create": (c_int, [c_void_p, c_int, c_int, c_int, c_int, ctypes.c_uint64,
                               P(c_void_p)]),
    "abh_mlp_destroy": (c_int, [c_void_p]),
    "abh_mlp_num_params": (c_int, [c_void_p, P(c_int64)]),
    "abh_mlp_get_params": (c_int, [c_void_p, c_void_p]),
    "abh_mlp_set_params": (c_int, [c_void_p, c_void_p]),
    "abh_mlp_forward": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int]),
    "abh_mlp_loss_grad": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, P(c_double), c_void_p]),
    "abh_mlp_train": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, P(MlpTrainOptions),
                              c_void_p]),
    "abh_ainv_create_lsq": (c_int, [c_void_p, c_void_p, P(Kernel), c_double, c_double, c_int,
                                    P(c_void_p)]),
    "abh_ainv_create_learned": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_double, c_double,
                                        c_int, P(c_void_p)]),
    "abh_ainv_destroy": (c_int, [c_void_p]),
    "abh_ainv_get": (c_int, [c_void_p, c_void_p, c_void_p]),
    "abh_ainv_features": (c_int, [c_void_p, c_double, c_void_p]),
    "abh_ainv_apply": (c_int, [c_void_p, c_void_p, c_void_p]),
    "abh_sparse_op_gmres": (c_int, [c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                                    P(GmresOptions), P(GmresStats)]),
}

_lib = None


def lib():
    """Loads libaboria_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libaboria_hip.so not found at %s: run `python -m aboria_amd.build` "
                "(or __graft_entry__.build()) first" % LIB_PATH)
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(status, ctx=None):
    if status != ABH_OK:
        msg = ""
        if ctx:
            m = lib().abh_last_error(ctx)
            msg = m.decode() if m else ""
        raise AbhError(status, msg)
