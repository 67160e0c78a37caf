"""ctypes front end of the oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, and only as the checker or the timed
host-CPU baseline.  The product package never imports this module.

The C side restates the reference's search path (beamSearch.h:33-646,
hashset.h:12-55, euclidian_point.h:45-90, mips_point.h:44-66,
vamana/index.h:63-319); see ``pann_oracle.h`` for the per-function citations.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

DT = {"float": 0, "float32": 0, "uint8": 1, "int8": 2}
MT = {"Euclidian": 0, "l2": 0, "mips": 1, "ip": 1}
NP_DT = {0: np.float32, 1: np.uint8, 2: np.int8}


class _Points(C.Structure):
    _fields_ = [("data", C.c_void_p), ("n", C.c_size_t), ("dim", C.c_uint),
                ("stride", C.c_size_t), ("dtype", C.c_int), ("metric", C.c_int)]


class _Graph(C.Structure):
    _fields_ = [("adj", C.c_void_p), ("n", C.c_size_t), ("max_degree", C.c_uint)]


class _HNSW(C.Structure):
    _fields_ = [("n", C.c_size_t), ("top", C.c_int), ("adj", C.c_void_p),
                ("max_degree", C.c_void_p)]


class Q8Params(C.Structure):
    _fields_ = [("slope", C.c_float), ("offset", C.c_int32), ("max_val", C.c_float),
                ("scale", C.c_float)]


class _QP(C.Structure):
    _fields_ = [("k", C.c_long), ("beam_size", C.c_long), ("limit", C.c_long),
                ("degree_limit", C.c_long), ("rerank_factor", C.c_double),
                ("batch_factor", C.c_double)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _configure(C.CDLL(_LIB_PATH))
    return _lib


@contextlib.contextmanager
def use_library(path: str):
    """Runs the functions of this module on another build of pann_oracle.c
    (a variant with a planted mistake, built by `make -C oracle mutant`)."""
    global _lib
    saved = lib()
    _lib = _configure(C.CDLL(path))
    try:
        yield
    finally:
        _lib = saved


def _configure(L):
    if True:
        vp, sz, u32p, f32p = C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p
        L.oracle_distance.restype = C.c_float
        L.oracle_distance.argtypes = [vp, vp, C.c_uint, C.c_int, C.c_int]
        L.oracle_hashset_new.restype = vp
        L.oracle_hashset_new.argtypes = [C.c_long]
        L.oracle_hashset_test.restype = C.c_int
        L.oracle_hashset_test.argtypes = [vp, C.c_uint32]
        L.oracle_hashset_num_entries.restype = C.c_long
        L.oracle_hashset_num_entries.argtypes = [vp]
        L.oracle_hashset_capacity.restype = C.c_long
        L.oracle_hashset_capacity.argtypes = [vp]
        L.oracle_hashset_grew.restype = C.c_int
        L.oracle_hashset_grew.argtypes = [vp]
        L.oracle_hashset_free.argtypes = [vp]
        L.oracle_beam_search.restype = C.c_int
        L.oracle_beam_search.argtypes = [
            C.POINTER(_Graph), C.POINTER(_Points), vp, C.c_long, u32p, C.c_int, C.POINTER(_QP),
            u32p, f32p, C.POINTER(C.c_int), u32p, f32p, C.POINTER(C.c_long), C.c_long,
            C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
        L.oracle_search_batch.restype = C.c_int
        L.oracle_search_batch.argtypes = [
            C.POINTER(_Graph), C.POINTER(_Points), C.POINTER(_Points), vp, vp, sz, C.c_uint32,
            C.POINTER(_QP), C.c_int, u32p, f32p, u32p, vp, vp]
        L.oracle_robust_prune.restype = C.c_int
        L.oracle_robust_prune.argtypes = [
            C.POINTER(_Graph), C.POINTER(_Points), C.c_uint32, u32p, f32p, C.c_int, C.c_double,
            C.c_int, C.c_int, u32p, C.POINTER(C.c_long)]
        L.oracle_build_vamana.restype = C.c_int
        L.oracle_build_vamana.argtypes = [
            C.POINTER(_Points), u32p, C.c_uint, C.c_uint, C.c_double, C.c_int, C.c_double, u32p,
            C.c_int, C.c_int]
        L.oracle_groundtruth_f64.restype = C.c_int
        L.oracle_groundtruth_f64.argtypes = [C.POINTER(_Points), vp, sz, C.c_int, C.c_int, u32p,
                                             vp]
        L.oracle_sort_pairs.argtypes = [u32p, f32p, C.c_long]
        L.oracle_q8_params_compute.restype = C.c_int
        L.oracle_q8_params_compute.argtypes = [C.POINTER(_Points), C.POINTER(Q8Params)]
        L.oracle_q8_translate.restype = None
        L.oracle_q8_translate.argtypes = [C.POINTER(Q8Params), C.c_int, vp, C.c_uint, vp]
        L.oracle_hnsw_search.restype = C.c_int
        L.oracle_hnsw_search.argtypes = [C.POINTER(_HNSW), C.POINTER(_Points), u32p, C.c_int, vp,
                                         sz, C.c_int, C.c_int, C.c_long, C.c_int, u32p, f32p, vp,
                                         vp, u32p]
        L.oracle_build_hnsw.restype = C.c_int
        L.oracle_build_hnsw.argtypes = [C.POINTER(_HNSW), C.POINTER(_Points), u32p, C.c_uint,
                                        C.c_float, C.c_float, u32p, C.c_int]
        L.oracle_vamana_batches.restype = C.c_long
        L.oracle_vamana_batches.argtypes = [sz, sz, vp, vp, C.c_long]
        L.oracle_hnsw_batches.restype = C.c_long
        L.oracle_hnsw_batches.argtypes = [C.c_uint32, C.c_float, vp, vp, C.c_long]
    return L


def vamana_batches(n, m=None):
    """batch_insert's [(floor, ceiling)] (vamana/index.h:202-234)."""
    m = n if m is None else m
    nb = lib().oracle_vamana_batches(n, m, None, None, 0)
    f = np.zeros(max(nb, 1), np.uint64)
    c = np.zeros(max(nb, 1), np.uint64)
    lib().oracle_vamana_batches(n, m, _ptr(f), _ptr(c), nb)
    return [(int(a), int(b)) for a, b in zip(f[:nb], c[:nb])]


def hnsw_batches(n, batch_base=2.0):
    """HNSW::HNSW's insert batches [(begin, end)] (HNSW.hpp:776-792)."""
    nb = lib().oracle_hnsw_batches(n, batch_base, None, None, 0)
    b = np.zeros(max(nb, 1), np.uint32)
    e = np.zeros(max(nb, 1), np.uint32)
    lib().oracle_hnsw_batches(n, batch_base, _ptr(b), _ptr(e), nb)
    return [(int(x), int(y)) for x, y in zip(b[:nb], e[:nb])]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _points(X: np.ndarray, metric: str) -> _Points:
    X = np.ascontiguousarray(X)
    dt = {np.dtype(np.float32): 0, np.dtype(np.uint8): 1, np.dtype(np.int8): 2}[X.dtype]
    return _Points(_ptr(X), X.shape[0], X.shape[1], X.strides[0], dt, MT[metric])


def _graph(adj: np.ndarray) -> _Graph:
    assert adj.dtype == np.uint32 and adj.flags.c_contiguous
    return _Graph(_ptr(adj), adj.shape[0], adj.shape[1] - 1)


def distance(p: np.ndarray, q: np.ndarray, metric: str) -> float:
    p = np.ascontiguousarray(p)
    q = np.ascontiguousarray(q, dtype=p.dtype)
    dt = {np.dtype(np.float32): 0, np.dtype(np.uint8): 1, np.dtype(np.int8): 2}[p.dtype]
    return float(lib().oracle_distance(_ptr(p), _ptr(q), p.shape[0], dt, MT[metric]))


class HashSet:
    """hashset<uint32_t> (hashset.h:12-55)."""

    def __init__(self, n: int):
        self._h = lib().oracle_hashset_new(n)

    def __call__(self, a: int) -> bool:
        return bool(lib().oracle_hashset_test(self._h, a))

    @property
    def num_entries(self):
        return lib().oracle_hashset_num_entries(self._h)

    @property
    def capacity(self):
        return lib().oracle_hashset_capacity(self._h)

    @property
    def grew(self):
        return bool(lib().oracle_hashset_grew(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_hashset_free(self._h)
            self._h = None


def sort_pairs(ids: np.ndarray, dists: np.ndarray):
    ids = np.ascontiguousarray(ids, dtype=np.uint32).copy()
    dists = np.ascontiguousarray(dists, dtype=np.float32).copy()
    lib().oracle_sort_pairs(_ptr(ids), _ptr(dists), len(ids))
    return ids, dists


def beam_search(adj, X, metric, query, starts, k, beam_size, limit=None, degree_limit=None,
                batch_factor=0.125, rerank_factor=100.0, self_id=-1, visited_cap=1 << 16):
    """One filtered_beam_search call (beamSearch.h:33-228).  Returns a dict
    with frontier ids/dists, sorted visited ids/dists and the counters."""
    n = X.shape[0]
    QP = _QP(k, beam_size, n if limit is None else limit,
             adj.shape[1] - 1 if degree_limit is None else degree_limit, rerank_factor,
             batch_factor)
    starts = np.ascontiguousarray(starts, dtype=np.uint32)
    q = np.ascontiguousarray(query, dtype=X.dtype)
    fi = np.zeros(beam_size + 1, np.uint32)
    fd = np.zeros(beam_size + 1, np.float32)
    vi = np.zeros(visited_cap, np.uint32)
    vd = np.zeros(visited_cap, np.float32)
    fs, vs = C.c_int(0), C.c_long(0)
    dc, fdc, grew = C.c_uint64(0), C.c_uint64(0), C.c_int(0)
    G, P = _graph(adj), _points(X, metric)
    rc = lib().oracle_beam_search(C.byref(G), C.byref(P), _ptr(q), self_id, _ptr(starts),
                                  len(starts), C.byref(QP), _ptr(fi), _ptr(fd), C.byref(fs),
                                  _ptr(vi), _ptr(vd), C.byref(vs), visited_cap, C.byref(dc),
                                  C.byref(fdc), C.byref(grew))
    if rc != 0:
        raise ValueError("oracle beam search rejected its arguments")
    nv = min(vs.value, visited_cap)
    return {"ids": fi[:fs.value].copy(), "dists": fd[:fs.value].copy(),
            "visited_ids": vi[:nv].copy(), "visited_dists": vd[:nv].copy(),
            "num_visited": vs.value, "dist_cmps": dc.value, "full_dist_cmps": fdc.value,
            "hash_grew": bool(grew.value)}


def search_batch(adj, X, metric, Q, k, beam_size, limit=None, degree_limit=None,
                 batch_factor=0.125, rerank_factor=100.0, start=0, nthreads=None,
                 Xq=None, Qq=None):
    """qsearchAll (beamSearch.h:599-646).  With (Xq, Qq) given, the search runs
    on the Xq store and the rerank branch of beam_search_rerank applies."""
    n = X.shape[0]
    nq = Q.shape[0]
    nthreads = nthreads or os.cpu_count() or 1
    QP = _QP(k, beam_size, n if limit is None else limit,
             adj.shape[1] - 1 if degree_limit is None else degree_limit, rerank_factor,
             batch_factor)
    Q = np.ascontiguousarray(Q, dtype=X.dtype)
    ids = np.zeros((nq, k), np.uint32)
    dists = np.zeros((nq, k), np.float32)
    vis = np.zeros(nq, np.uint32)
    cmps = np.zeros(nq, np.uint64)
    grew = np.zeros(nq, np.int32)
    G, P = _graph(adj), _points(X, metric)
    if Xq is not None:
        PQ = _points(Xq, metric)
        Qq = np.ascontiguousarray(Qq, dtype=Xq.dtype)
        pq_ptr, qq_ptr = C.byref(PQ), _ptr(Qq)
    else:
        pq_ptr, qq_ptr = None, None
    rc = lib().oracle_search_batch(C.byref(G), C.byref(P), pq_ptr, _ptr(Q), qq_ptr, nq, start,
                                   C.byref(QP), nthreads, _ptr(ids), _ptr(dists), _ptr(vis),
                                   _ptr(cmps), _ptr(grew))
    if rc != 0:
        raise ValueError("oracle search failed (k > beam or short beam)")
    return {"ids": ids, "dists": dists, "visited": vis, "dist_cmps": cmps,
            "hash_grew": grew.astype(bool)}


def robust_prune(adj, X, metric, p, cand_ids, cand_dists, alpha, R, add=True):
    cand_ids = np.ascontiguousarray(cand_ids, dtype=np.uint32)
    cand_dists = np.ascontiguousarray(cand_dists, dtype=np.float32)
    out = np.zeros(max(R, 1), np.uint32)
    comps = C.c_long(0)
    G, P = _graph(adj), _points(X, metric)
    n = lib().oracle_robust_prune(C.byref(G), C.byref(P), p, _ptr(cand_ids), _ptr(cand_dists),
                                  len(cand_ids), alpha, R, int(add), _ptr(out), C.byref(comps))
    return out[:n].copy(), comps.value


def build_vamana(X, metric, R, L, alpha, perms, num_passes=1, batch_factor=0.125,
                 sort_neighbors=True, nthreads=None):
    """build_index (vamana/index.h:150-186); perms: (num_passes, n) uint32."""
    n = X.shape[0]
    perms = np.ascontiguousarray(perms, dtype=np.uint32).reshape(num_passes, n)
    adj = np.zeros((n, R + 1), np.uint32)
    P = _points(X, metric)
    nthreads = nthreads or os.cpu_count() or 1
    rc = lib().oracle_build_vamana(C.byref(P), _ptr(adj), R, L, alpha, num_passes, batch_factor,
                                   _ptr(perms), int(sort_neighbors), nthreads)
    if rc != 0:
        raise ValueError("oracle build failed")
    # slots past each degree keep stale ids (update_neighbors writes only the
    # new list, graph.h:77-88); the graph file never contains them -- zero them
    tail = adj[:, 1:]
    tail[np.arange(R)[None, :] >= adj[:, :1].astype(np.int64)] = 0
    return adj


def groundtruth_f64(X, metric, Q, k, nthreads=None):
    Q = np.ascontiguousarray(Q, dtype=X.dtype)
    ids = np.zeros((Q.shape[0], k), np.uint32)
    d = np.zeros((Q.shape[0], k), np.float64)
    P = _points(X, metric)
    rc = lib().oracle_groundtruth_f64(C.byref(P), _ptr(Q), Q.shape[0], k,
                                      nthreads or os.cpu_count() or 1, _ptr(ids), _ptr(d))
    if rc != 0:
        raise ValueError("bad k")
    return ids, d


def _hnsw(levels_adj):
    """oracle_hnsw over a list of per-level (n, deg + 1) uint32 arrays; the
    second value keeps the pointer tables alive."""
    for a in levels_adj:
        assert a.dtype == np.uint32 and a.flags.c_contiguous
    ptrs = (C.c_void_p * len(levels_adj))(*[a.ctypes.data for a in levels_adj])
    degs = (C.c_uint * len(levels_adj))(*[a.shape[1] - 1 for a in levels_adj])
    H = _HNSW(levels_adj[0].shape[0], len(levels_adj) - 1, C.cast(ptrs, C.c_void_p),
              C.cast(degs, C.c_void_p))
    return H, (ptrs, degs, levels_adj)


def hnsw_search(levels_adj, X, metric, Q, k, ef, entrance, limit=0, nthreads=None):
    """HNSW::search (HNSW.hpp:1568-1613).  Returns (ids, dists) of shape
    (nq, ef) holding `count[q]` node ids each, count, cmps, visited."""
    Q = np.ascontiguousarray(Q, dtype=X.dtype)
    nq = Q.shape[0]
    ids = np.zeros((nq, ef), np.uint32)
    d = np.zeros((nq, ef), np.float32)
    cnt = np.zeros(nq, np.int32)
    cmps = np.zeros(nq, np.uint64)
    vis = np.zeros(nq, np.uint32)
    ent = np.ascontiguousarray(entrance, dtype=np.uint32)
    H, keep = _hnsw(levels_adj)
    P = _points(X, metric)
    rc = lib().oracle_hnsw_search(C.byref(H), C.byref(P), _ptr(ent), len(ent), _ptr(Q), nq, k, ef,
                                  limit, nthreads or os.cpu_count() or 1, _ptr(ids), _ptr(d),
                                  _ptr(cnt), _ptr(cmps), _ptr(vis))
    if rc != 0:
        raise ValueError("oracle hnsw search failed")
    return ids, d, cnt, cmps, vis


def build_hnsw(X, metric, m, efc, alpha, levels, batch_base=2.0, nthreads=None):
    """HNSW::HNSW + insert (HNSW.hpp:748-996) with the given node levels.
    Returns (per-level adjacency list, entrance node)."""
    n = X.shape[0]
    levels = np.ascontiguousarray(levels, dtype=np.uint32)
    top = int(levels.max()) if n else 0
    adj = [np.zeros((n, (2 * m if l == 0 else m) + 1), np.uint32) for l in range(top + 1)]
    H, keep = _hnsw(adj)
    P = _points(X, metric)
    ent = np.zeros(1, np.uint32)
    rc = lib().oracle_build_hnsw(C.byref(H), C.byref(P), _ptr(levels), efc, alpha, batch_base,
                                 _ptr(ent), nthreads or os.cpu_count() or 1)
    if rc != 0:
        raise ValueError("oracle hnsw build failed")
    for a in adj:  # slots past each degree are never read; keep them zero
        tail = a[:, 1:]
        tail[np.arange(a.shape[1] - 1)[None, :] >= a[:, :1].astype(np.int64)] = 0
    return adj, int(ent[0])


def q8_params(X, metric):
    """The reference's scalar quantization parameters for float points."""
    prm = Q8Params()
    P = _points(np.ascontiguousarray(X, dtype=np.float32), metric)
    if lib().oracle_q8_params_compute(C.byref(P), C.byref(prm)) != 0:
        raise ValueError("q8 parameters need float points")
    return prm


def q8_translate(prm, metric, X):
    """Codes of the rows of X (uint8 for L2, int8 for MIPS)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    out = np.zeros(X.shape, np.uint8 if MT[metric] == 0 else np.int8)
    for i in range(X.shape[0]):
        lib().oracle_q8_translate(C.byref(prm), MT[metric], X[i].ctypes.data, X.shape[1],
                                  out[i].ctypes.data)
    return out
