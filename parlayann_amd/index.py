"""Device index and the reference's Python search surface.

``DeviceIndex`` owns one libpann index (points + graph in HBM) and exposes
the batched search of qsearchAll / beam_search_rerank (beamSearch.h:528-646)
over host (numpy) or device (torch) buffers.

``load_index`` / ``GraphIndex`` mirror python/wrapper.py ``load_index`` and
python/graph_index.cpp ``GraphIndex`` (batch_search :192-216, single_search
:218-235, batch_search_from_string :237-257, check_recall :259-305), with the
intended QueryParams mapping (k = knn, beamSize = beam_width, limit =
visit_limit if > 0 else n, degree_limit = max degree) instead of the stale
``QueryParams(knn, beam_width, 1.35, ...)`` call (graph_index.cpp:198).
``quant=True`` searches the reference's own 8-bit scalar codes (uint8 affine
for Euclidian, trimmed int8 for MIPS, :141-186) and reranks in FP32; the FP8
and BF16 stores are available through DeviceIndex.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from . import formats

_DT = {np.dtype(np.float32): L.DT_F32, np.dtype(np.uint8): L.DT_U8, np.dtype(np.int8): L.DT_I8}
_METRIC = {"Euclidian": L.METRIC_L2, "l2": L.METRIC_L2, "L2": L.METRIC_L2,
           "mips": L.METRIC_MIPS, "ip": L.METRIC_MIPS, "MIPS": L.METRIC_MIPS}
_STORE = {"native": L.STORE_NATIVE, "bf16": L.STORE_BF16, "fp8": L.STORE_FP8, "q8": L.STORE_Q8,
          "fp8r": L.STORE_FP8_RESIDUAL}
_SUM = {"inorder": L.SUM_INORDER, "tree": L.SUM_TREE}


def _ptr(a) -> Optional[int]:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


class DeviceIndex:
    """Points (n x dim, float32/uint8/int8) and a fixed-stride adjacency
    (n x (max_degree + 1) uint32, slot 0 = degree; graph.h:46-64) in HBM."""

    def __init__(self, points: np.ndarray, adj: np.ndarray, metric: str = "l2", device: int = 0):
        points = np.ascontiguousarray(points)
        adj = np.ascontiguousarray(adj, dtype=np.uint32)
        if points.dtype not in _DT:
            raise TypeError(f"unsupported point dtype {points.dtype}")
        if adj.ndim != 2 or adj.shape[0] != points.shape[0]:
            raise ValueError("graph size and point size do not match")  # graph_index.cpp:113-116
        self.n, self.dim = points.shape
        self.max_degree = adj.shape[1] - 1
        self.dtype = points.dtype
        self.metric = metric
        self.device = device
        desc = L.IndexDesc(self.n, self.dim, self.max_degree, _DT[points.dtype], _METRIC[metric],
                           points.ctypes.data, adj.ctypes.data, device)
        h = C.c_void_p()
        L.check(L.lib().pann_open(C.byref(desc), C.byref(h)))
        self._h = h
        self._stores = {"native"}

    def close(self):
        if getattr(self, "_h", None):
            L.lib().pann_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stores --------------------------------------------------------------
    def build_store(self, kind: str, ncent: int = 0):
        """Reduced-precision copy of float points: 'bf16', 'fp8' (e4m3,
        per-dimension absmax scales), 'q8' (the reference's 8-bit scalar
        quantization: uint8 affine for L2, trimmed int8 for MIPS) or 'fp8r'
        (MIPS: e4m3 of each point minus the nearest of `ncent` sampled points,
        0 = automatic)."""
        if kind == "fp8r":
            if kind not in self._stores or ncent:
                L.check(L.lib().pann_build_fp8_residual(self._h, int(ncent)))
                self._stores.add(kind)
        elif kind not in self._stores:
            L.check(L.lib().pann_build_store(self._h, _STORE[kind]))
            self._stores.add(kind)

    def fp8_residual(self) -> dict:
        """The 'fp8r' store's centroids (ncent x dim), each point's centroid
        and the per-dimension scales."""
        nc = C.c_uint32()
        L.check(L.lib().pann_get_fp8_residual(self._h, C.byref(nc), None, None, None))
        cent = np.zeros((nc.value, self.dim), np.float32)
        assign = np.zeros(self.n, np.uint32)
        scales = np.zeros(self.dim, np.float32)
        L.check(L.lib().pann_get_fp8_residual(self._h, C.byref(nc), cent.ctypes.data,
                                              assign.ctypes.data, scales.ctypes.data))
        return {"centroids": cent, "assign": assign, "scales": scales}

    def q8_params(self) -> dict:
        """Parameters of the 'q8' store (the reference's 8-bit scalar quantization)."""
        sl, mx, sc = C.c_float(), C.c_float(), C.c_float()
        off = C.c_int32()
        L.check(L.lib().pann_get_q8_params(self._h, C.byref(sl), C.byref(off), C.byref(mx),
                                           C.byref(sc)))
        return {"slope": sl.value, "offset": off.value, "max_val": mx.value, "scale": sc.value}

    def fp8_scales(self) -> np.ndarray:
        s = np.zeros(self.dim, np.float32)
        L.check(L.lib().pann_get_fp8_scales(self._h, s.ctypes.data))
        return s

    def download_graph(self) -> np.ndarray:
        adj = np.zeros((self.n, self.max_degree + 1), np.uint32)
        L.check(L.lib().pann_download_graph(self._h, adj.ctypes.data))
        return adj

    def build_eval_counts(self, reset: bool = False) -> dict:
        """Distance evaluations of the build steps on this index so far:
        {"search": build-mode searches, "prune": robustPrune row distances}."""
        se, pe = C.c_uint64(0), C.c_uint64(0)
        L.check(L.lib().pann_build_eval_counts(self._h, C.byref(se), C.byref(pe), int(reset)))
        return {"search": int(se.value), "prune": int(pe.value)}

    # -- search --------------------------------------------------------------
    def _qparams(self, k, beam, limit, degree_limit, rerank_factor, batch_factor):
        return L.QParams(int(k), int(beam), int(self.n if limit is None else limit),
                         int(self.max_degree if degree_limit is None else degree_limit),
                         float(rerank_factor), float(batch_factor))

    @staticmethod
    def _opts(store, summation, rerank, hash_slots_log2, async_, waves_per_cu, exact_counts,
              visits_per_round=0, eval_cache_log2=0, frontier_filter=0, waves_per_query=0,
              query_order=0):
        return L.SearchOpts(_STORE[store], _SUM[summation], int(bool(rerank)),
                            int(hash_slots_log2), int(bool(async_)), int(waves_per_cu),
                            int(bool(exact_counts)), int(visits_per_round),
                            int(eval_cache_log2), int(frontier_filter), int(waves_per_query),
                            int(query_order))

    def search(self, queries, k: int, beam: int, *, limit=None, degree_limit=None,
               batch_factor: float = 0.125, rerank_factor: float = 100.0,
               store: str = "native", summation: str = "inorder", rerank: bool = False,
               starts: Sequence[int] = (0,), hash_slots_log2: int = 0, waves_per_cu: int = 0,
               exact_counts: bool = False, visits_per_round: int = 0,
               eval_cache_log2: int = 0, frontier_filter: int = 0, waves_per_query: int = 0):
        """Host-buffer batched search (qsearchAll, beamSearch.h:599-646).
        Returns dict(ids uint32[nq,k], dists float32[nq,k], visited, dist_cmps);
        dist_cmps is the reference's count with exact_counts, else the number
        of distance evaluations."""
        qdt = np.float32 if (self.dtype == np.float32) else self.dtype
        Q = np.ascontiguousarray(queries, dtype=qdt)
        if Q.ndim != 2 or Q.shape[1] != self.dim:
            raise ValueError(f"queries must be (nq, {self.dim})")
        nq = Q.shape[0]
        ids = np.zeros((nq, k), np.uint32)
        dists = np.zeros((nq, k), np.float32)
        vis = np.zeros(nq, np.uint32)
        cmps = np.zeros(nq, np.uint32)
        st = np.ascontiguousarray(starts, dtype=np.uint32)
        qp = self._qparams(k, beam, limit, degree_limit, rerank_factor, batch_factor)
        op = self._opts(store, summation, rerank, hash_slots_log2, False, waves_per_cu,
                        exact_counts, visits_per_round, eval_cache_log2, frontier_filter,
                        waves_per_query)
        L.check(L.lib().pann_search_host(self._h, Q.ctypes.data, nq, C.byref(qp), st.ctypes.data,
                                         len(st), C.byref(op), ids.ctypes.data,
                                         dists.ctypes.data, vis.ctypes.data, cmps.ctypes.data))
        return {"ids": ids, "dists": dists, "visited": vis, "dist_cmps": cmps}

    def beam_search(self, queries, beam: int, *, limit=None, degree_limit=None,
                    batch_factor: float = 0.125, starts: Sequence[int] = (0,),
                    summation: str = "inorder", exact_counts: bool = False, vis_cap: int = 0):
        """beam_search / filtered_beam_search without filtering (beamSearch.h:33-228,
        :419-445) for host queries: each query's whole final frontier sorted by
        (dist, id), and with vis_cap > 0 its visited list sorted the same way.
        Returns dict(frontier_ids, frontier_dists, frontier_count, visited,
        dist_cmps[, visited_ids, visited_dists])."""
        qdt = np.float32 if (self.dtype == np.float32) else self.dtype
        Q = np.ascontiguousarray(queries, dtype=qdt)
        if Q.ndim != 2 or Q.shape[1] != self.dim:
            raise ValueError(f"queries must be (nq, {self.dim})")
        nq = Q.shape[0]
        fi = np.zeros((nq, beam), np.uint32)
        fd = np.zeros((nq, beam), np.float32)
        fc = np.zeros(nq, np.uint32)
        vis = np.zeros(nq, np.uint32)
        cmps = np.zeros(nq, np.uint32)
        st = np.ascontiguousarray(starts, dtype=np.uint32)
        qp = self._qparams(beam, beam, limit, degree_limit, 100.0, batch_factor)
        op = self._opts("native", summation, False, 0, False, 0, exact_counts)
        out = {"frontier_ids": fi, "frontier_dists": fd, "frontier_count": fc, "visited": vis,
               "dist_cmps": cmps}
        if vis_cap > 0:
            vi = np.zeros((nq, vis_cap), np.uint32)
            vd = np.zeros((nq, vis_cap), np.float32)
            L.check(L.lib().pann_beam_search_visited_host(
                self._h, Q.ctypes.data, nq, C.byref(qp), st.ctypes.data, len(st), C.byref(op),
                fi.ctypes.data, fd.ctypes.data, fc.ctypes.data, int(vis_cap), vi.ctypes.data,
                vd.ctypes.data, vis.ctypes.data, cmps.ctypes.data))
            out["visited_ids"], out["visited_dists"] = vi, vd
        else:
            L.check(L.lib().pann_beam_search_host(
                self._h, Q.ctypes.data, nq, C.byref(qp), st.ctypes.data, len(st), C.byref(op),
                fi.ctypes.data, fd.ctypes.data, fc.ctypes.data, vis.ctypes.data,
                cmps.ctypes.data))
        return out

    def search_device(self, queries, k: int, beam: int, out_ids, out_dists, *, out_visited=None,
                      out_cmps=None, k_per_query=None, beam_per_query=None, limit=None,
                      degree_limit=None, batch_factor: float = 0.125,
                      rerank_factor: float = 100.0, store: str = "native",
                      summation: str = "inorder", rerank: bool = False,
                      starts: Sequence[int] = (0,), hash_slots_log2: int = 0,
                      waves_per_cu: int = 0, stream=None, async_: bool = False,
                      exact_counts: bool = False, visits_per_round: int = 0,
                      eval_cache_log2: int = 0, frontier_filter: int = 0,
                      waves_per_query: int = 0, query_order: int = 0):
        """Device-buffer search: queries/outputs are torch tensors on the
        index's device.  `beam` is the maximum beam when beam_per_query is
        given.  Launches on `stream` (torch stream or raw handle).  With
        async_ the tensors must stay alive, and unchanged, until
        search_status on the same stream."""
        nq = int(queries.shape[0])
        st = np.ascontiguousarray(starts, dtype=np.uint32)
        qp = self._qparams(k, beam, limit, degree_limit, rerank_factor, batch_factor)
        op = self._opts(store, summation, rerank, hash_slots_log2, async_, waves_per_cu,
                        exact_counts, visits_per_round, eval_cache_log2, frontier_filter,
                        waves_per_query, query_order)
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(L.lib().pann_search(self._h, _ptr(queries), nq, C.byref(qp), _ptr(k_per_query),
                                    _ptr(beam_per_query), st.ctypes.data, len(st), C.byref(op),
                                    _ptr(out_ids), _ptr(out_dists), int(out_ids.shape[1]),
                                    _ptr(out_visited), _ptr(out_cmps), stream))

    def search_status(self, stream=None):
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(L.lib().pann_search_status(self._h, stream))

    def last_kernel_ms(self) -> float:
        ms = C.c_float(0)
        L.check(L.lib().pann_last_kernel_ms(self._h, C.byref(ms)))
        return float(ms.value)


def device_alloc_count() -> int:
    """Device allocations libpann has made since it was loaded."""
    c = C.c_uint64(0)
    L.check(L.lib().pann_device_alloc_count(C.byref(c)))
    return int(c.value)


def device_info(device: int = 0):
    cus, lds, hbm = C.c_int32(), C.c_int64(), C.c_int64()
    L.check(L.lib().pann_device_info(device, C.byref(cus), C.byref(lds), C.byref(hbm)))
    return {"num_cus": cus.value, "lds_per_cu": lds.value, "hbm_bytes": hbm.value}


# ---------------------------------------------------------------------------
# the reference's Python surface (python/wrapper.py, python/graph_index.cpp)
# ---------------------------------------------------------------------------

_NP_DT = {"float": np.float32, "uint8": np.uint8, "int8": np.int8}


class GraphIndex:
    """GraphIndex<T, Point> (graph_index.cpp:48-118) over a `.bin` vector file
    and a graph file, searched on the GPU."""

    def __init__(self, metric: str, dtype: str, data_path: str, index_path: str,
                 hnsw: bool = False, device: int = 0):
        if hnsw:
            from .hnsw import HNSWIndex  # layered model file (HNSW.hpp:631-744)
            self._hnsw = HNSWIndex.load(index_path, data_path, metric, dtype, device=device)
            self._dev = self._hnsw.base
            self.points = self._hnsw.points
            self.metric, self.dtype = metric, _NP_DT[dtype]
            return
        self._hnsw = None
        if metric not in ("Euclidian", "mips"):
            raise ValueError("Invalid metric")
        if dtype not in _NP_DT:
            raise ValueError("Invalid data type")
        self.metric, self.dtype = metric, _NP_DT[dtype]
        pts = formats.read_vectors(data_path, self.dtype)
        if self.dtype == np.float32 and metric == "mips":
            # graph_index.cpp:92-95: float MIPS points are normalised on load
            pts = _normalize(pts)
        self.points = pts
        adj = formats.read_graph(index_path)
        self._dev = DeviceIndex(pts, adj, "l2" if metric == "Euclidian" else "mips", device)

    @property
    def device_index(self) -> DeviceIndex:
        return self._dev

    def _prep_queries(self, queries, quant):
        Q = np.ascontiguousarray(queries, dtype=self.dtype)
        if Q.ndim == 1:
            Q = Q[None, :]
        if quant and self.metric == "mips" and self.dtype == np.float32:
            Q = _normalize(Q)  # graph_index.cpp:171
        return Q

    def batch_search(self, queries, knn: int, beam_width: int, quant: bool = False,
                     visit_limit: int = -1):
        """(ids uint32[nq, knn], dists float32[nq, knn]) (graph_index.cpp:192-216).
        queries: numpy (or anything numpy converts, a CPU torch tensor too); a
        torch tensor on the index's GPU is searched in place and the results
        come back as device tensors (ids int32 holding the uint32 bits)."""
        if _on_device(queries):
            return self._batch_search_device(queries, knn, beam_width, quant, visit_limit)
        Q = self._prep_queries(queries, quant)
        limit = visit_limit if visit_limit > 0 else None
        if self._hnsw is not None:
            # HNSW::search returns a tie-extended list (HNSW.hpp:1590-1596); the
            # binding's fixed (nq, knn) output takes its first knn entries and
            # marks missing ones with id 0xFFFFFFFF, distance +inf
            use_q = quant and self.dtype == np.float32
            if use_q:
                self._hnsw.build_store("q8")
            r = self._hnsw.search(Q, knn, max(beam_width, knn), limit=limit or 0,
                                  store="q8" if use_q else "native", rerank=use_q)
            ids = np.full((len(Q), knn), 0xFFFFFFFF, np.uint32)
            dists = np.full((len(Q), knn), np.inf, np.float32)
            w = min(knn, r["ids"].shape[1])
            have = np.arange(w)[None, :] < r["count"][:, None]
            ids[:, :w][have] = r["ids"][:, :w][have]
            dists[:, :w][have] = r["dists"][:, :w][have]
            return ids, dists
        use_q = quant and self.dtype == np.float32
        if use_q:  # the reference's quant path: 8-bit scalar codes, float rerank
            self._dev.build_store("q8")
        r = self._dev.search(Q, knn, beam_width, limit=limit,
                             store="q8" if use_q else "native", rerank=use_q)
        return r["ids"], r["dists"]

    search = batch_search

    def _batch_search_device(self, queries, knn, beam_width, quant, visit_limit):
        import torch
        want = {np.dtype(np.float32): torch.float32, np.dtype(np.uint8): torch.uint8,
                np.dtype(np.int8): torch.int8}[np.dtype(self.dtype)]
        if queries.dtype != want:
            raise TypeError(f"queries must be {want} for this index")
        Q = queries.reshape(1, -1) if queries.dim() == 1 else queries
        Q = Q.contiguous()
        if quant and self.metric == "mips" and self.dtype == np.float32:
            # graph_index.cpp:171, as _normalize: double norm, x * float(1/norm)
            norm = torch.sqrt((Q.double() * Q.double()).sum(1))
            norm[norm == 0] = 1.0
            Q = (Q * (1.0 / norm).float()[:, None]).contiguous()
        use_q = quant and self.dtype == np.float32
        nq = Q.shape[0]
        dev = Q.device
        limit = visit_limit if visit_limit > 0 else None
        if self._hnsw is not None:
            if use_q:
                self._hnsw.build_store("q8")
            ef = max(beam_width, knn)
            oi = torch.empty((nq, ef), dtype=torch.int32, device=dev)
            od = torch.empty((nq, ef), dtype=torch.float32, device=dev)
            oc = torch.empty(nq, dtype=torch.int32, device=dev)
            self._hnsw.search_device(Q, knn, ef, oi, od, oc, limit=limit or 0,
                                     store="q8" if use_q else "native", rerank=use_q)
            w = min(knn, ef)
            ids = torch.full((nq, knn), -1, dtype=torch.int32, device=dev)
            dists = torch.full((nq, knn), float("inf"), dtype=torch.float32, device=dev)
            have = torch.arange(w, device=dev)[None, :] < oc[:, None]
            ids[:, :w] = torch.where(have, oi[:, :w], ids[:, :w])
            dists[:, :w] = torch.where(have, od[:, :w], dists[:, :w])
            return ids, dists
        if use_q:
            self._dev.build_store("q8")
        oi = torch.empty((nq, knn), dtype=torch.int32, device=dev)
        od = torch.empty((nq, knn), dtype=torch.float32, device=dev)
        self._dev.search_device(Q, knn, beam_width, oi, od, limit=limit,
                                store="q8" if use_q else "native", rerank=use_q)
        return oi, od

    def single_search(self, q, knn: int, beam_width: int, quant: bool = False,
                      visit_limit: int = -1):
        if _on_device(q):
            ids, _ = self._batch_search_device(q.reshape(1, -1), knn, beam_width, quant,
                                               visit_limit)
            return ids[0]
        ids, _ = self.batch_search(np.asarray(q)[None, :], knn, beam_width, quant, visit_limit)
        return ids[0]

    def batch_search_from_string(self, queries: str, knn: int, beam_width: int,
                                 quant: bool = False, visit_limit: int = -1):
        Q = formats.read_vectors(queries, self.dtype)
        return self.batch_search(Q, knn, beam_width, quant, visit_limit)

    def check_recall(self, queries_file: str, graph_file: str, neighbors, k: int) -> float:
        """Recall with ties at the k-th groundtruth distance counted
        (graph_index.cpp:259-305); prints and returns it."""
        from .recall import recall_with_ties
        Q = formats.read_vectors(queries_file, self.dtype)
        gt_ids, _ = formats.read_ibin(graph_file)
        nb = np.asarray(neighbors)
        if nb.min() < 0 or nb.max() >= len(self.points):
            raise ValueError("neighbor reported by query out of range")
        r = recall_with_ties(self.points, Q, gt_ids, nb, k,
                             "l2" if self.metric == "Euclidian" else "mips")
        print(f"Recall: {r:.6f}")
        return r


def _on_device(x) -> bool:
    """A torch tensor in GPU memory (numpy and CPU tensors go the host way)."""
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


def _normalize(X: np.ndarray) -> np.ndarray:
    """Mips_Point::normalize (mips_point.h:120-129): double norm, x * (1/norm)."""
    X = np.asarray(X, dtype=np.float32)
    norm = np.sqrt(np.einsum("ij,ij->i", X.astype(np.float64), X.astype(np.float64)))
    norm[norm == 0] = 1.0
    inv = (1.0 / norm).astype(np.float32)
    return (X * inv[:, None]).astype(np.float32)


def load_index(metric: str, dtype: str, data_dir: str, index_dir: str, hnsw: bool = False):
    """python/wrapper.py load_index."""
    if metric not in ("Euclidian", "mips"):
        raise Exception("Invalid metric")
    if dtype not in _NP_DT:
        raise Exception("Invalid data type")
    return GraphIndex(metric, dtype, data_dir, index_dir, hnsw)


def _variant(metric, dtype):
    def make(data_path, index_path, hnsw=False):
        return GraphIndex(metric, dtype, data_path, index_path, hnsw)
    return make


FloatEuclidianIndex = _variant("Euclidian", "float")
FloatMipsIndex = _variant("mips", "float")
UInt8EuclidianIndex = _variant("Euclidian", "uint8")
UInt8MipsIndex = _variant("mips", "uint8")
Int8EuclidianIndex = _variant("Euclidian", "int8")
Int8MipsIndex = _variant("mips", "int8")
