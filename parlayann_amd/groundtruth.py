"""FP64 brute-force groundtruth (compute_groundtruth.cpp:22-60, in double,
ordered by (distance, id)).

groundtruth_device takes its candidates from the library's brute-force
kernel on the matrix cores; groundtruth from a blocked GEMM as below.

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
    return _rescore(points, queries, cand, k, metric)


def _rescore(points, queries, cand, k, metric):
    """Exact float64 re-score of every candidate, then (dist, id) order."""
    nq = cand.shape[0]
    ids = np.empty((nq, k), np.uint32)
    dists = np.empty((nq, k), np.float64)
    for s in range(0, nq, 2048):
        e = min(nq, s + 2048)
        c = cand[s:e]
        P = points[c].astype(np.float64)                      # b x kk x d
        qd = np.asarray(queries[s:e], dtype=np.float64)[:, None, :]
        d = ((P - qd) ** 2).sum(-1) if metric == "l2" else -(P * qd).sum(-1)
        # (dist, id) order of every row: stable sort by id, then by distance
        o1 = np.argsort(c, axis=1, kind="stable")
        c1 = np.take_along_axis(c, o1, 1)
        d1 = np.take_along_axis(d, o1, 1)
        o2 = np.argsort(d1, axis=1, kind="stable")[:, :k]
        ids[s:e] = np.take_along_axis(c1, o2, 1)
        dists[s:e] = np.take_along_axis(d1, o2, 1)
    return ids, dists


def groundtruth_device(index, points: np.ndarray, queries: np.ndarray, k: int,
                       metric: str = "l2", margin: int = 32):
    """Groundtruth with the candidates from the brute-force MFMA kernel
    (pann_bruteforce_candidates: the k + margin smallest FP32 distances per
    query over the points resident in `index`, a DeviceIndex of `points`),
    then an exact float64 re-score of those candidates on the device and the
    (distance, id) order.  Returns (ids uint32[nq,k], dists float64[nq,k])."""
    import torch
    from . import _lib as L
    n, nq = points.shape[0], queries.shape[0]
    kk = min(224, k + margin, n)
    if kk < k:
        raise ValueError("k + margin exceeds the kernel's 224 candidates")
    dev = torch.device("cuda", index.device)
    Q = torch.from_numpy(np.ascontiguousarray(queries, dtype=np.float32)).to(dev)
    oi = torch.empty((nq, kk), dtype=torch.int32, device=dev)
    od = torch.empty((nq, kk), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    L.check(L.lib().pann_bruteforce_candidates(index._h, Q.data_ptr(), nq, kk, oi.data_ptr(),
                                               od.data_ptr(), None))
    del od
    # exact float64 re-score of the candidates, on the device, chunked
    Xd = torch.from_numpy(np.ascontiguousarray(points)).to(dev)
    cand = oi.to(torch.int64)
    del oi
    ids = np.empty((nq, k), np.uint32)
    dists = np.empty((nq, k), np.float64)
    step = max(1, (1 << 30) // (kk * points.shape[1] * 8))
    for s in range(0, nq, step):
        e = min(nq, s + step)
        c = cand[s:e]
        P = Xd[c].to(torch.float64)                               # b x kk x d
        q = Q[s:e].to(torch.float64)[:, None, :]
        d = ((P - q) ** 2).sum(-1) if metric == "l2" else -(P * q).sum(-1)
        c1, o1 = torch.sort(c, dim=1, stable=True)                # (dist, id) order:
        d1 = torch.gather(d, 1, o1)                               # by id, then by dist
        d2, o2 = torch.sort(d1, dim=1, stable=True)
        ids[s:e] = torch.gather(c1, 1, o2[:, :k]).cpu().numpy().astype(np.uint32)
        dists[s:e] = d2[:, :k].cpu().numpy()
    return ids, dists


def write_groundtruth(path: str, ids, dists):
    formats.write_ibin(path, ids, np.asarray(dists, dtype=np.float32))
