"""On-disk formats of ParlayANN, read and written unchanged.

- vectors ``.fbin/.u8bin/.i8bin``: ``u32 n, u32 d`` then ``n*d`` row-major
  elements (point_range.h:74-117);
- graph: ``u32 n, u32 max_deg, u32 deg[n]`` then the concatenated neighbour
  lists (graph.h:151-208 read, :210-237 write); in memory a fixed-stride
  ``n x (max_deg + 1)`` array whose slot 0 is the degree (graph.h:46-64);
- groundtruth ``.ibin``: ``i32 n, i32 k, u32 ids[n*k], f32 dists[n*k]``
  (types.h:38-73, compute_groundtruth.cpp:64-100).
"""
from __future__ import annotations

import numpy as np

_EXT_DTYPE = {".fbin": np.float32, ".u8bin": np.uint8, ".i8bin": np.int8}


def dtype_for_path(path: str):
    for ext, dt in _EXT_DTYPE.items():
        if path.endswith(ext):
            return dt
    return None


def read_vectors(path: str, dtype=None, mmap: bool = False) -> np.ndarray:
    """PointRange(char* filename) (point_range.h:74-117) without the in-memory
    64-byte row padding (the device store applies its own padding)."""
    dtype = dtype or dtype_for_path(path)
    if dtype is None:
        raise ValueError(f"cannot infer element type of {path}")
    hdr = np.fromfile(path, dtype=np.uint32, count=2)
    n, d = int(hdr[0]), int(hdr[1])
    if mmap:
        return np.memmap(path, dtype=dtype, mode="r", offset=8, shape=(n, d))
    X = np.fromfile(path, dtype=dtype, count=n * d, offset=8)
    if X.size != n * d:
        raise ValueError(f"{path}: truncated vector file")
    return X.reshape(n, d)


def write_vectors(path: str, X: np.ndarray) -> None:
    X = np.ascontiguousarray(X)
    with open(path, "wb") as f:
        np.array([X.shape[0], X.shape[1]], dtype=np.uint32).tofile(f)
        X.tofile(f)


def read_graph(path: str) -> np.ndarray:
    """Graph(char* gFile) (graph.h:151-208) -> fixed-stride adjacency."""
    with open(path, "rb") as f:
        n, max_deg = (int(v) for v in np.fromfile(f, dtype=np.uint32, count=2))
        deg = np.fromfile(f, dtype=np.uint32, count=n)
        edges = np.fromfile(f, dtype=np.uint32, count=int(deg.sum(dtype=np.int64)))
    if deg.size != n or edges.size != int(deg.sum(dtype=np.int64)):
        raise ValueError(f"{path}: truncated graph file")
    if n and int(deg.max()) > max_deg:
        raise ValueError(f"{path}: degree exceeds max degree")
    adj = np.zeros((n, max_deg + 1), dtype=np.uint32)
    adj[:, 0] = deg
    # scatter each list into its row
    rows = np.repeat(np.arange(n, dtype=np.int64), deg.astype(np.int64))
    starts = np.concatenate([[0], np.cumsum(deg, dtype=np.int64)[:-1]])
    cols = np.arange(edges.size, dtype=np.int64) - np.repeat(starts, deg.astype(np.int64)) + 1
    adj[rows, cols] = edges
    return adj


def write_graph(path: str, adj: np.ndarray) -> None:
    """Graph::save (graph.h:210-237)."""
    adj = np.ascontiguousarray(adj, dtype=np.uint32)
    n, w = adj.shape
    deg = adj[:, 0]
    mask = np.arange(w - 1)[None, :] < deg[:, None].astype(np.int64)
    with open(path, "wb") as f:
        np.array([n, w - 1], dtype=np.uint32).tofile(f)
        deg.tofile(f)
        adj[:, 1:][mask].astype(np.uint32).tofile(f)


def read_ibin(path: str):
    """groundTruth(char*) (types.h:48-73): (ids int/uint32 [n,k], dists f32 [n,k])."""
    hdr = np.fromfile(path, dtype=np.int32, count=2)
    n, k = int(hdr[0]), int(hdr[1])
    ids = np.fromfile(path, dtype=np.uint32, count=n * k, offset=8).reshape(n, k)
    dists = np.fromfile(path, dtype=np.float32, count=n * k, offset=8 + 4 * n * k).reshape(n, k)
    return ids, dists


def write_ibin(path: str, ids: np.ndarray, dists: np.ndarray) -> None:
    """write_ibin (compute_groundtruth.cpp:64-100)."""
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    dists = np.ascontiguousarray(dists, dtype=np.float32)
    with open(path, "wb") as f:
        np.array(ids.shape, dtype=np.int32).tofile(f)
        ids.tofile(f)
        dists.tofile(f)
