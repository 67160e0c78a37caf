"""ctypes binding of libpann.so (include/pann.h).  Loading fails loudly when
the library has not been built; there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PANN_LIB_PATH") or os.path.join(HERE, "libpann.so")

PANN_OK, PANN_ERR_ARG, PANN_ERR_DEVICE, PANN_ERR_OOM, PANN_ERR_CAPACITY = 0, 1, 2, 3, 4
PANN_ERR_SHORT_BEAM, PANN_ERR_UNSUPPORTED = 5, 6
DT_F32, DT_U8, DT_I8 = 0, 1, 2
METRIC_L2, METRIC_MIPS = 0, 1
STORE_NATIVE, STORE_BF16, STORE_FP8, STORE_Q8, STORE_FP8_RESIDUAL = 0, 1, 2, 3, 4
SUM_INORDER, SUM_TREE = 0, 1

# every symbol include/pann.h declares
EXPORTS = ["pann_open", "pann_build_store", "pann_get_fp8_scales", "pann_get_q8_params",
           "pann_build_fp8_residual", "pann_get_fp8_residual",
           "pann_search", "pann_search_noalloc",
           "pann_search_host", "pann_beam_search_host", "pann_beam_search_visited_host",
           "pann_device_alloc_count", "pann_build_eval_counts", "pann_search_status",
           "pann_last_kernel_ms",
           "pann_search_visited", "pann_robust_prune_batch", "pann_robust_prune_groups",
           "pann_update_neighbors",
           "pann_download_graph", "pann_close", "pann_last_error", "pann_device_info",
           "pann_hnsw_set_level", "pann_hnsw_search_layer", "pann_hnsw_search",
           "pann_hnsw_select_neighbors", "pann_hnsw_add_reverse", "pann_hnsw_set_neighbors",
           "pann_hnsw_download_level", "pann_bruteforce_candidates"]


class PannError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"pann status {status}: {msg}")
        self.status = status


class IndexDesc(C.Structure):
    _fields_ = [("n", C.c_uint32), ("dim", C.c_uint32), ("max_degree", C.c_uint32),
                ("dtype", C.c_int32), ("metric", C.c_int32), ("points", C.c_void_p),
                ("graph", C.c_void_p), ("device", C.c_int32)]


class QParams(C.Structure):
    _fields_ = [("k", C.c_int64), ("beam_size", C.c_int64), ("limit", C.c_int64),
                ("degree_limit", C.c_int64), ("rerank_factor", C.c_double),
                ("batch_factor", C.c_double)]


class SearchOpts(C.Structure):
    _fields_ = [("store", C.c_int32), ("summation", C.c_int32), ("rerank", C.c_int32),
                ("hash_slots_log2", C.c_int32), ("async_", C.c_int32),
                ("waves_per_cu", C.c_int32), ("exact_counts", C.c_int32),
                ("visits_per_round", C.c_int32), ("eval_cache_log2", C.c_int32),
                ("frontier_filter", C.c_int32), ("waves_per_query", C.c_int32),
                ("query_order", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m parlayann_amd._build` "
                          "(or __graft_entry__.build()) first -- there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp, st = C.c_void_p, C.c_int
    sig = {
        "pann_open": [C.POINTER(IndexDesc), C.POINTER(vp)],
        "pann_build_store": [vp, C.c_int32],
        "pann_get_fp8_scales": [vp, vp],
        "pann_build_fp8_residual": [vp, C.c_uint32],
        "pann_get_fp8_residual": [vp, vp, vp, vp, vp],
        "pann_get_q8_params": [vp, vp, vp, vp, vp],
        "pann_search": [vp, vp, C.c_uint32, C.POINTER(QParams), vp, vp, vp, C.c_uint32,
                        C.POINTER(SearchOpts), vp, vp, C.c_uint32, vp, vp, vp],
        "pann_search_noalloc": [vp, vp, C.c_uint32, C.POINTER(QParams), vp, vp, vp,
                                C.c_uint32, C.POINTER(SearchOpts), vp, vp, C.c_uint32, vp, vp,
                                vp],
        "pann_search_host": [vp, vp, C.c_uint32, C.POINTER(QParams), vp, C.c_uint32,
                             C.POINTER(SearchOpts), vp, vp, vp, vp],
        "pann_beam_search_host": [vp, vp, C.c_uint32, C.POINTER(QParams), vp, C.c_uint32,
                                  C.POINTER(SearchOpts), vp, vp, vp, vp, vp],
        "pann_beam_search_visited_host": [vp, vp, C.c_uint32, C.POINTER(QParams), vp,
                                          C.c_uint32, C.POINTER(SearchOpts), vp, vp, vp,
                                          C.c_uint32, vp, vp, vp, vp],
        "pann_device_alloc_count": [C.POINTER(C.c_uint64)],
        "pann_build_eval_counts": [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   C.c_int32],
        "pann_search_status": [vp, vp],
        "pann_last_kernel_ms": [vp, C.POINTER(C.c_float)],
        "pann_search_visited": [vp, vp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32,
                                C.c_uint32, vp, vp, vp, vp],
        "pann_robust_prune_batch": [vp, vp, C.c_uint32, vp, vp, vp, C.c_uint32, C.c_double,
                                    C.c_uint32, C.c_int32, vp, vp, vp],
        "pann_robust_prune_groups": [vp, vp, C.c_uint32, vp, vp, C.c_uint32, C.c_double,
                                     C.c_uint32, C.c_int32, vp, vp, vp],
        "pann_update_neighbors": [vp, vp, C.c_uint32, vp, vp, C.c_uint32, vp],
        "pann_download_graph": [vp, vp],
        "pann_hnsw_set_level": [vp, C.c_uint32, C.c_uint32, vp],
        "pann_hnsw_search_layer": [vp, C.c_uint32, vp, vp, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_int64, vp, vp, C.c_uint32, C.POINTER(SearchOpts), vp, vp,
                                   C.c_uint32, vp, vp, vp, vp],
        "pann_hnsw_search": [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, C.c_int64, vp,
                             C.c_uint32, C.c_uint32, vp, C.POINTER(SearchOpts), vp, vp,
                             C.c_uint32, vp, vp, vp, vp],
        "pann_hnsw_select_neighbors": [vp, C.c_uint32, vp, C.c_uint32, vp, vp, vp, C.c_uint32,
                                       C.c_float, C.c_uint32, vp, vp, vp],
        "pann_hnsw_add_reverse": [vp, C.c_uint32, vp, C.c_uint32, vp, vp, C.c_uint32,
                                  C.c_uint32, vp, vp, vp],
        "pann_hnsw_set_neighbors": [vp, C.c_uint32, vp, C.c_uint32, vp, vp, C.c_uint32, vp],
        "pann_hnsw_download_level": [vp, C.c_uint32, vp],
        "pann_bruteforce_candidates": [vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp],
        "pann_close": [vp],
        "pann_device_info": [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                             C.POINTER(C.c_int64)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = st
    L.pann_last_error.argtypes = []
    L.pann_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(status: int):
    if status != PANN_OK:
        raise PannError(status, lib().pann_last_error().decode())
