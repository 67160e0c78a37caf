"""Recall as the reference measures it.

``recall_with_ties`` follows checkRecall (check_nn_recall.h:86-111) and
GraphIndex::check_recall (graph_index.cpp:259-305): groundtruth ids past k
whose distance -- recomputed with the reference's float arithmetic -- equals
that of the k-th groundtruth id also count as correct.  ``recall_plain`` is
|reported[:k] & gt[:k]| / k.  Both average over queries.
"""
from __future__ import annotations

import numpy as np


def ref_distance(points: np.ndarray, ids: np.ndarray, q: np.ndarray, metric: str) -> np.ndarray:
    """Point::distance for rows `ids` against query q, with the reference's
    arithmetic: exact int32 for uint8/int8 (euclidian_point.h:45-81,
    mips_point.h:44-58), an in-order float32 sum of separately rounded terms
    for float (euclidian_point.h:83-90, mips_point.h:60-66)."""
    P = points[ids]
    if points.dtype == np.float32:
        qf = np.asarray(q, dtype=np.float32)
        if metric == "l2":
            t = (qf[None, :] - P).astype(np.float32)
            terms = (t * t).astype(np.float32)
            return np.cumsum(terms, axis=1, dtype=np.float32)[:, -1]
        terms = (qf[None, :] * P).astype(np.float32)
        return -np.cumsum(terms, axis=1, dtype=np.float32)[:, -1]
    qi = np.asarray(q).astype(np.int64)
    Pi = P.astype(np.int64)
    if metric == "l2":
        return ((qi[None, :] - Pi) ** 2).sum(axis=1).astype(np.int32).astype(np.float32)
    return -((qi[None, :] * Pi).sum(axis=1).astype(np.int32).astype(np.float32))


def recall_with_ties(points, queries, gt_ids, reported, k: int, metric: str) -> float:
    gt_ids = np.asarray(gt_ids)
    reported = np.asarray(reported)
    nq = reported.shape[0]
    correct = 0
    for i in range(nq):
        res = list(gt_ids[i, :k])
        if gt_ids.shape[1] > k:
            d = ref_distance(points, gt_ids[i], queries[i], metric)
            last = d[k - 1]
            res += [int(g) for g, dd in zip(gt_ids[i, k:], d[k:]) if dd == last]
        rep = set(int(x) for x in reported[i, :k])
        correct += sum(1 for g in res if int(g) in rep)
    return correct / float(k * nq)


def recall_plain(gt_ids, reported, k: int) -> float:
    gt_ids = np.asarray(gt_ids)[:, :k]
    reported = np.asarray(reported)[:, :k]
    hits = 0
    for g, r in zip(gt_ids, reported):
        hits += len(set(g.tolist()) & set(r.tolist()))
    return hits / float(k * gt_ids.shape[0])
