"""FP64 brute-force groundtruth (compute_groundtruth.cpp:22-60, in double,
ordered by (distance, id)).

Candidates come from a blocked float64 GEMM (|q|^2 + |x|^2 - 2 q.x) on the
GPU when torch sees one, else on the host with numpy; the top k + margin of
each query are then re-scored with direct float64 differences and sorted, so
the GEMM's cancellation error cannot reorder the result.
"""
from __future__ import annotations

import numpy as np

from . import formats


def _exact_f64(points, ids, q, metric):
    P = points[ids].astype(np.float64)
    qd = np.asarray(q, dtype=np.float64)
    if metric == "l2":
        return ((P - qd[None, :]) ** 2).sum(axis=1)
    return -(P @ qd)


def groundtruth(points: np.ndarray, queries: np.ndarray, k: int, metric: str = "l2",
                margin: int = 32, block: int = 0, use_gpu: bool | None = None):
    """Returns (ids uint32[nq,k], dists float64[nq,k])."""
    n = points.shape[0]
    if block <= 0:  # keep one block's distance matrix near 2 GB
        block = int(max(1, min(256, (2 << 30) // (8 * max(n, 1)))))
    kk = min(n, k + margin)
    nq = queries.shape[0]
    cand = np.empty((nq, kk), dtype=np.int64)
    torch = None
    if use_gpu is not False:
        try:
            import torch as _t
            if _t.cuda.is_available():
                torch = _t
        except Exception:
            torch = None
    if torch is not None:
        dev = torch.device("cuda")
        X = torch.from_numpy(np.ascontiguousarray(points)).to(dev, torch.float64)
        xx = (X * X).sum(1)
        for s in range(0, nq, block):
            Qb = torch.from_numpy(np.ascontiguousarray(queries[s:s + block])).to(dev, torch.float64)
            if metric == "l2":
                D = xx[None, :] - 2.0 * (Qb @ X.T)
            else:
                D = -(Qb @ X.T)
            cand[s:s + block] = torch.topk(D, kk, dim=1, largest=False).indices.cpu().numpy()
        del X
    else:
        X = np.asarray(points, dtype=np.float64)
        xx = (X * X).sum(1)
        for s in range(0, nq, block):
            Qb = np.asarray(queries[s:s + block], dtype=np.float64)
            D = xx[None, :] - 2.0 * (Qb @ X.T) if metric == "l2" else -(Qb @ X.T)
            part = np.argpartition(D, kk - 1, axis=1)[:, :kk]
            cand[s:s + block] = part
    # exact float64 re-score of every candidate, then (dist, id) order
    ids = np.empty((nq, k), np.uint32)
    dists = np.empty((nq, k), np.float64)
    for s in range(0, nq, 2048):
        e = min(nq, s + 2048)
        c = cand[s:e]
        P = points[c].astype(np.float64)                      # b x kk x d
        qd = np.asarray(queries[s:e], dtype=np.float64)[:, None, :]
        d = ((P - qd) ** 2).sum(-1) if metric == "l2" else -(P * qd).sum(-1)
        o = np.array([np.lexsort((c[i], d[i]))[:k] for i in range(e - s)])
        ids[s:e] = np.take_along_axis(c, o, 1)
        dists[s:e] = np.take_along_axis(d, o, 1)
    return ids, dists


def write_groundtruth(path: str, ids, dists):
    formats.write_ibin(path, ids, np.asarray(dists, dtype=np.float32))
