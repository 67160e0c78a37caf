"""HNSW on the device: the model file, search, and batch construction.

Restates algorithms/HNSW/HNSW.hpp:
  * the model file (load :631-744, save :1615-1686): "HNSW", version 3,
    code_U, sizeof(node), dim, m_l, m, efc, alpha, n, per-node (level, id),
    per node and level (size_t count, u32 node ids), then the entrance list;
  * search (:1568-1613): ef = 1 descent from the entrance through the upper
    levels (search_layer_to, :1545-1566), layer-0 beam search with ef, and
    the cut to k that keeps every entry tied with the k-th distance;
  * construction (HNSW::HNSW :748-807, insert :809-996): prefix-doubling
    batches, per batch an ef = 1 descent to each node's level, then per level
    a search with efc from the previous level's results, prune_heuristic
    (:430-484), forward lists and the reverse-edge step.

search_layer binds QueryParams(ef, ef, 1.35, limit, threshold_m(l)) (:1096),
so the literal 1.35 becomes `limit` = 1; the intended mapping
QueryParams(k = ef, beamSize = ef, limit, degree_limit = threshold_m(l)) is
used here (and in the reference's Python binding the HNSW search path is
commented out, graph_index.cpp:122-136).  Node levels come from
get_level_random's law, floor(-ln U * m_l) with U ~ U(FLT_MIN, 1) (:505-516),
drawn from a seeded generator instead of per-thread mt19937 streams.  A batch
never ends empty (the reference's 0.02 n batch limit is 0 for n < 50).
Everything per node and per level runs on the GPU; the host keeps the batch
schedule.
"""
from __future__ import annotations

import ctypes as C
import math
import struct
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib as L
from . import formats
from .index import DeviceIndex, _METRIC, _normalize


def random_levels(n: int, m_l: float, seed: int = 0) -> np.ndarray:
    """get_level_random (HNSW.hpp:505-516): uint32(-log(U) * m_l),
    U ~ U(FLT_MIN, 1)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.finfo(np.float32).tiny, 1.0, n)
    return np.floor(-np.log(u) * m_l).astype(np.uint32)


@dataclass
class HNSWModel:
    """The contents of an HNSW model file."""
    dim: int
    m_l: float
    m: int
    efc: int
    alpha: float
    node_levels: np.ndarray                 # uint32[n]
    node_ids: np.ndarray                    # uint32[n], point id of node i
    graphs: List[np.ndarray]                # per level: n x (deg + 1), Graph layout
    entrance: np.ndarray                    # uint32[], node ids
    code_u: int = 0
    size_node: int = 0

    @property
    def n(self) -> int:
        return int(self.node_levels.shape[0])

    @property
    def top_level(self) -> int:
        return int(self.node_levels[self.entrance[0]])


_HDR = struct.Struct("<4sIQQIfIIfI")  # HNSW.hpp:1660-1671


def read_model(path: str) -> HNSWModel:
    """HNSW::HNSW(filename_model, getter) (HNSW.hpp:631-744).  Per-level
    lists are returned as Graph-layout arrays of degree 2m (level 0) and m."""
    buf = np.fromfile(path, dtype=np.uint8)
    mv = memoryview(buf)
    magic, version, code_u, size_node, dim, m_l, m, efc, alpha, n = _HDR.unpack_from(mv, 0)
    if magic != b"HNSW":
        raise ValueError("Wrong type of model")
    if version != 3:
        raise ValueError("Unsupported version")
    pos = _HDR.size
    nodes = np.frombuffer(buf, dtype=np.uint32, count=2 * n, offset=pos).reshape(n, 2)
    node_levels = nodes[:, 0].copy()
    node_ids = nodes[:, 1].copy()
    pos += 8 * n
    top = int(node_levels.max()) if n else 0
    graphs = [np.zeros((n, (2 * m if l == 0 else m) + 1), np.uint32) for l in range(top + 1)]
    u64 = struct.Struct("<Q")
    for i in range(n):
        for l in range(int(node_levels[i]) + 1):
            (cnt,) = u64.unpack_from(mv, pos)
            pos += 8
            g = graphs[l]
            if cnt > g.shape[1] - 1:
                raise ValueError(f"node {i} level {l}: {cnt} neighbours exceed the level's "
                                 f"degree {g.shape[1] - 1}")
            g[i, 0] = cnt
            g[i, 1:1 + cnt] = np.frombuffer(buf, dtype=np.uint32, count=cnt, offset=pos)
            pos += 4 * cnt
    (ne,) = u64.unpack_from(mv, pos)
    pos += 8
    entrance = np.frombuffer(buf, dtype=np.uint32, count=ne, offset=pos).copy()
    return HNSWModel(dim, m_l, m, efc, alpha, node_levels, node_ids, graphs, entrance,
                     code_u, size_node)


def write_model(path: str, mdl: HNSWModel):
    """HNSW::save (HNSW.hpp:1615-1686).  code_U and sizeof(node) are written
    as stored in the model (the loader does not check them, :681-684)."""
    n = mdl.n
    parts = [_HDR.pack(b"HNSW", 3, mdl.code_u, mdl.size_node, mdl.dim, mdl.m_l, mdl.m, mdl.efc,
                       mdl.alpha, n)]
    parts.append(np.stack([mdl.node_levels, mdl.node_ids], 1).astype("<u4").tobytes())
    u64 = struct.Struct("<Q")
    for i in range(n):
        for l in range(int(mdl.node_levels[i]) + 1):
            g = mdl.graphs[l]
            cnt = int(g[i, 0])
            parts.append(u64.pack(cnt))
            parts.append(g[i, 1:1 + cnt].astype("<u4").tobytes())
    parts.append(u64.pack(len(mdl.entrance)))
    parts.append(np.asarray(mdl.entrance, "<u4").tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def _p(t):
    return None if t is None else t.data_ptr()


class HNSWIndex:
    """A layered index in HBM: level 0 (degree 2m) in a DeviceIndex plus one
    graph per upper level (degree m), over the points in node order."""

    def __init__(self, points: np.ndarray, model: HNSWModel, metric: str = "l2", device: int = 0):
        if points.shape[1] != model.dim:
            raise ValueError("point dimension does not match the model")
        ident = np.array_equal(model.node_ids, np.arange(model.n, dtype=np.uint32))
        self.points = points
        node_pts = points if ident else np.ascontiguousarray(points[model.node_ids])
        self.model = model
        self.metric = metric
        self.base = DeviceIndex(node_pts, model.graphs[0], metric, device)
        for l in range(1, len(model.graphs)):
            g = np.ascontiguousarray(model.graphs[l], dtype=np.uint32)
            L.check(L.lib().pann_hnsw_set_level(self.base._h, l, g.shape[1] - 1, g.ctypes.data))
        self._node_ids = None if ident else model.node_ids
        self._dev_node_ids = None

    @classmethod
    def load(cls, index_path: str, data_path: str, metric: str, dtype: str, device: int = 0):
        """GraphIndex(data_path, index_path, is_hnsw=true) (graph_index.cpp:82-110):
        float MIPS points are normalised as that constructor does (:93-95)."""
        X = formats.read_vectors(data_path, {"float": np.float32, "uint8": np.uint8,
                                             "int8": np.int8}[dtype])
        if X.dtype == np.float32 and _METRIC[metric] == L.METRIC_MIPS:
            X = _normalize(X)
        return cls(X, read_model(index_path), metric, device)

    def build_store(self, kind: str, ncent: int = 0):
        self.base.build_store(kind, ncent)

    def search_device(self, queries, k: int, ef: int, out_ids, out_dists, out_count, *,
                      limit: int = 0, store: str = "native", summation: str = "inorder",
                      rerank: bool = False, out_visited=None, out_cmps=None, stream=None,
                      async_: bool = False, k_per_query=None, ef_per_query=None,
                      exact_counts: bool = False, hash_slots_log2: int = 0,
                      waves_per_cu: int = 0, visits_per_round: int = 0,
                      eval_cache_log2: int = 0, frontier_filter: int = 0,
                      waves_per_query: int = 0, query_order: int = 0):
        """HNSW::search for a batch of device queries (torch tensors).  Rows of
        out_ids / out_dists (stride >= ef) hold out_count[q] results, point ids.
        k_per_query / ef_per_query (int32 tensors, ef_per_query <= ef) mix
        requests in one batch.  The launch options are DeviceIndex.search's
        (exact_counts also selects the reference's exact seen set, which at
        small ef and high degree gathers fewer rows).  With async_ the tensors
        must stay alive, and unchanged, until search_status on the same
        stream."""
        import torch
        if self._node_ids is not None and self._dev_node_ids is None:
            self._dev_node_ids = torch.from_numpy(self._node_ids.view(np.int32)).to(queries.device)
        nq = int(queries.shape[0])
        ent = np.ascontiguousarray(self.model.entrance, dtype=np.uint32)
        op = DeviceIndex._opts(store, summation, rerank, hash_slots_log2, async_, waves_per_cu,
                               exact_counts, visits_per_round, eval_cache_log2,
                               frontier_filter, waves_per_query, query_order)
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(L.lib().pann_hnsw_search(
            self.base._h, _p(queries), nq, k, ef, _p(k_per_query), _p(ef_per_query), limit,
            ent.ctypes.data, len(ent),
            self.model.top_level, _p(self._dev_node_ids), C.byref(op), _p(out_ids),
            _p(out_dists), int(out_ids.shape[1]), _p(out_count), _p(out_visited), _p(out_cmps),
            stream))

    def search(self, queries: np.ndarray, k: int, ef: int, *, limit: int = 0,
               store: str = "native", summation: str = "inorder", rerank: bool = False):
        """Host-buffer HNSW::search.  Returns dict(ids, dists: (nq, ef) arrays
        holding count[q] results each, count, visited, dist_cmps)."""
        import torch
        dev = torch.device("cuda", self.base.device)
        qdt = np.float32 if self.base.dtype == np.float32 else self.base.dtype
        Q = torch.from_numpy(np.ascontiguousarray(queries, dtype=qdt)).to(dev)
        nq = Q.shape[0]
        oi = torch.zeros((nq, ef), dtype=torch.int32, device=dev)
        od = torch.zeros((nq, ef), dtype=torch.float32, device=dev)
        oc = torch.zeros(nq, dtype=torch.int32, device=dev)
        ov = torch.zeros(nq, dtype=torch.int32, device=dev)
        om = torch.zeros(nq, dtype=torch.int32, device=dev)
        self.search_device(Q, k, ef, oi, od, oc, limit=limit, store=store, summation=summation,
                           rerank=rerank, out_visited=ov, out_cmps=om)
        return {"ids": oi.cpu().numpy().view(np.uint32), "dists": od.cpu().numpy(),
                "count": oc.cpu().numpy(), "visited": ov.cpu().numpy().view(np.uint32),
                "dist_cmps": om.cpu().numpy().view(np.uint32)}

    def close(self):
        self.base.close()


def batch_bounds(n: int, batch_base: float = 2.0):
    """The insert batches of HNSW::HNSW (HNSW.hpp:777-792)."""
    size_limit = max(1, int(n * 0.02))
    out, be = [], 1
    while be < n:
        bb = be
        be = min(n, int(math.ceil(float(np.float32(np.float32(bb) * np.float32(batch_base))))) + 1,
                 bb + size_limit)
        out.append((bb, be))
    return out


def build_hnsw(points: np.ndarray, metric: str, m: int, efc: int, alpha: float = 1.0,
               m_l: Optional[float] = None, levels: Optional[np.ndarray] = None, seed: int = 0,
               batch_base: float = 2.0, summation: str = "inorder", device: int = 0,
               progress=None) -> HNSWModel:
    """HNSW::HNSW (HNSW.hpp:748-807) on the GPU; node i is point i."""
    import torch

    n = points.shape[0]
    if m < 1 or 2 * m > 64:
        raise ValueError("m must be in [1, 32] (level-0 degree 2m <= 64)")
    m_l = (1.0 / math.log(m)) if (m_l is None and m > 1) else (m_l or 1.0)
    levels = random_levels(n, m_l, seed) if levels is None else \
        np.ascontiguousarray(levels, dtype=np.uint32)
    top = int(levels.max()) if n else 0
    idx = DeviceIndex(points, np.zeros((n, 2 * m + 1), np.uint32), metric, device)
    lib, h = L.lib(), idx._h
    for l in range(1, top + 1):
        L.check(lib.pann_hnsw_set_level(h, l, m, None))
    dev = torch.device("cuda", device)
    lev_t = torch.from_numpy(levels.astype(np.int32)).to(dev)
    op = DeviceIndex._opts("native", summation, False, 0, False, 0, False)
    entrance = 0
    for bi, (bb, be) in enumerate(batch_bounds(n, batch_base)):
        sz = be - bb
        ids = torch.arange(bb, be, dtype=torch.int32, device=dev)
        lu = lev_t[bb:be]
        level_ep = int(levels[entrance])
        eps = torch.full((sz, efc), entrance, dtype=torch.int32, device=dev)
        neps = torch.ones(sz, dtype=torch.int32, device=dev)
        # ef = 1 descent through the levels above each node's (:895-904)
        for l in range(level_ep, 0, -1):
            act = (lu < l).nonzero().flatten()
            if act.numel() == 0:
                continue
            na = act.numel()
            pts = ids[act].contiguous()
            st = eps[act, :1].contiguous()
            oi = torch.empty((na, 1), dtype=torch.int32, device=dev)
            od = torch.empty((na, 1), dtype=torch.float32, device=dev)
            oc = torch.empty(na, dtype=torch.int32, device=dev)
            L.check(lib.pann_hnsw_search_layer(h, l, None, _p(pts), na, 1, m, 0, _p(st), None, 1,
                                               C.byref(op), _p(oi), _p(od), 1, _p(oc), None,
                                               None, None))
            eps[act, 0] = oi[:, 0]
        # each level from the entrance's down (:907-985)
        for lc in range(level_ep, -1, -1):
            act = (lu >= lc).nonzero().flatten()
            if act.numel() == 0:
                continue
            na = act.numel()
            M = 2 * m if lc == 0 else m
            pts = ids[act].contiguous()
            st = eps[act].contiguous()
            ns = neps[act].contiguous()
            rid = torch.empty((na, efc), dtype=torch.int32, device=dev)
            rd = torch.empty((na, efc), dtype=torch.float32, device=dev)
            rc = torch.empty(na, dtype=torch.int32, device=dev)
            L.check(lib.pann_hnsw_search_layer(h, lc, None, _p(pts), na, efc, M, 0, _p(st),
                                               _p(ns), efc, C.byref(op), _p(rid), _p(rd), efc,
                                               _p(rc), None, None, None))
            sel = torch.empty((na, M), dtype=torch.int32, device=dev)
            scnt = torch.empty(na, dtype=torch.int32, device=dev)
            L.check(lib.pann_hnsw_select_neighbors(h, lc, _p(pts), na, _p(rid), _p(rd), _p(rc),
                                                   efc, alpha, M, _p(sel), _p(scnt), None))
            L.check(lib.pann_hnsw_set_neighbors(h, lc, _p(pts), na, _p(sel), _p(scnt), M, None))
            eps[act] = rid
            neps[act] = rc
            # reverse edges (pv, pu) grouped by pv, values in insertion order
            mask = torch.arange(M, device=dev)[None, :] < scnt[:, None]
            keys = sel[mask]
            if keys.numel() == 0:
                continue
            vals = pts[:, None].expand(-1, M)[mask]
            keys_s, order = torch.sort(keys, stable=True)
            vals_s = vals[order].contiguous()
            ukeys, counts = torch.unique_consecutive(keys_s, return_counts=True)
            ukeys = ukeys.contiguous()
            ng = ukeys.numel()
            offs = torch.zeros(ng + 1, dtype=torch.int32, device=dev)
            offs[1:] = torch.cumsum(counts, 0).to(torch.int32)
            out = torch.empty((ng, M), dtype=torch.int32, device=dev)
            ocnt = torch.empty(ng, dtype=torch.int32, device=dev)
            L.check(lib.pann_hnsw_add_reverse(h, lc, _p(ukeys), ng, _p(offs), _p(vals_s),
                                              int(counts.max()), M, _p(out), _p(ocnt), None))
            L.check(lib.pann_hnsw_set_neighbors(h, lc, _p(ukeys), ng, _p(out), _p(ocnt), M,
                                                None))
        hi = bb + int(np.argmax(levels[bb:be]))  # max_element: the first highest (:987-995)
        if levels[hi] > levels[entrance]:
            entrance = hi
        if progress:
            progress(bi, bb, be)
    graphs = []
    for l in range(top + 1):
        deg = 2 * m if l == 0 else m
        g = np.zeros((n, deg + 1), np.uint32)
        L.check(lib.pann_hnsw_download_level(h, l, g.ctypes.data))
        graphs.append(g)
    idx.close()
    return HNSWModel(points.shape[1], float(m_l), m, efc, float(alpha), levels,
                     np.arange(n, dtype=np.uint32), graphs, np.array([entrance], np.uint32))
