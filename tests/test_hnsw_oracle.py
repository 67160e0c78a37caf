"""Pins the oracle's HNSW restatement (oracle_hnsw_search / oracle_build_hnsw,
HNSW.hpp:748-996, :1545-1613) and the model file layout (:631-744,
:1615-1686) against what the mathematics fixes: exact kNN when the layer-0
search sees every node, the tie-extended cut, the greedy descent on complete
upper levels, counters summed over levels, build invariants, recall and
thread-count independence.  The exact graph the reference would build is
not pinned (its levels come from per-thread mt19937 streams, :505-516, and
parlay's group_by_key order is unspecified): parity unpinned for graph
bit-equality with the reference."""
import struct

import numpy as np
import pytest

from parlayann_amd import hnsw
from tests.hnsw_data import build_small


def complete_level(n, members, deg):
    """Graph-layout level where every member links to every other member."""
    g = np.zeros((n, deg + 1), np.uint32)
    for u in members:
        others = [v for v in members if v != u][:deg]
        g[u, 0] = len(others)
        g[u, 1:1 + len(others)] = others
    return g


def brute(X, q, metric):
    if metric == "l2":
        d = ((X.astype(np.float64) - q.astype(np.float64)) ** 2).sum(1)
    else:
        d = -(X.astype(np.float64) @ q.astype(np.float64))
    return np.lexsort((np.arange(len(X)), d)), d


def test_exact_knn_when_layer0_sees_all(oracle, rng):
    n, d = 40, 5
    X = rng.standard_normal((n, d)).astype(np.float32)
    Q = rng.standard_normal((7, d)).astype(np.float32)
    g0 = complete_level(n, list(range(n)), 64)[:, :65]
    ids, ds, cnt, cmps, vis = oracle.hnsw_search([g0], X, "l2", Q, 5, 64, [3])
    for i, q in enumerate(Q):
        order, _ = brute(X, q, "l2")
        assert cnt[i] >= 5
        assert list(ids[i, :5]) == list(order[:5])
        # one start + every other node evaluated once
        assert cmps[i] == n


def test_tie_extended_cut(oracle):
    # query at the origin; 6 points at squared distance exactly 1, 3 farther
    d = 3
    pts = [np.eye(d)[j] * s for j in range(d) for s in (1, -1)]
    pts += [np.full(d, 2.0), np.full(d, -2.0), np.array([3.0, 0, 0])]
    X = np.array(pts, np.float32)
    n = len(X)
    g0 = complete_level(n, list(range(n)), 16)
    Q = np.zeros((1, d), np.float32)
    for k, expect in [(1, 6), (2, 6), (6, 6), (7, 7), (0, 0)]:
        ids, ds, cnt, _, _ = oracle.hnsw_search([g0], X, "l2", Q, k, 16, [8])
        assert cnt[0] == expect, k
        assert np.all(np.diff(ds[0, :cnt[0]]) >= 0)
    # k >= |frontier|: everything the search found, no cut
    ids, ds, cnt, _, _ = oracle.hnsw_search([g0], X, "l2", Q, 50, 4, [8])
    assert cnt[0] == 4


def test_descent_reaches_level_nearest(oracle, rng):
    n, d = 60, 4
    X = rng.standard_normal((n, d)).astype(np.float32)
    lvl1 = list(range(0, n, 5))       # nodes on level 1
    lvl2 = list(range(0, n, 15))      # nodes on level 2
    g2 = complete_level(n, lvl2, 4)
    g1 = complete_level(n, lvl1, 16)
    # level 0: a path 0-1-2-...: the layer-0 search must start where the descent ends
    g0 = np.zeros((n, 3), np.uint32)
    for u in range(n):
        nb = [v for v in (u - 1, u + 1) if 0 <= v < n]
        g0[u, 0] = len(nb)
        g0[u, 1:1 + len(nb)] = nb
    q = rng.standard_normal((1, d)).astype(np.float32)
    _, dist = brute(X, q[0], "l2")
    near1 = lvl1[int(np.argmin(dist[lvl1]))]
    ids, ds, cnt, cmps, vis = oracle.hnsw_search([g0, g1, g2], X, "l2", q, 1, 1, [lvl2[0]])
    # ef = 1 on the path from near1 walks downhill to a local minimum next to it
    walk = near1
    while True:
        nb = [v for v in (walk - 1, walk + 1) if 0 <= v < n]
        best = min(nb, key=lambda v: (dist[v], v))
        if dist[best] < dist[walk]:
            walk = best
        else:
            break
    assert ids[0, 0] == walk
    # level 2: 1 start + 3 others; level 1: start + 11 others; level 0 at least 1
    assert cmps[0] >= (1 + len(lvl2) - 1) + (1 + len(lvl1) - 1) + 1


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_build_invariants_and_recall(oracle, metric):
    m = 8
    X, lev, adj, ent = build_small(oracle, metric, m=m)
    n = len(X)
    assert lev[ent] == lev.max()
    for l, g in enumerate(adj):
        deg = 2 * m if l == 0 else m
        assert g.shape == (n, deg + 1)
        assert np.all(g[:, 0] <= deg)
        for u in range(n):
            nb = g[u, 1:1 + g[u, 0]]
            if lev[u] < l:
                assert g[u, 0] == 0
                continue
            assert u not in nb
            assert len(set(nb.tolist())) == len(nb)
            assert np.all(lev[nb] >= l)
    assert adj[0][:, 0].min() >= 1
    rng = np.random.default_rng(9)
    Q = X[rng.integers(0, n, 200)] + 0.1 * rng.standard_normal((200, X.shape[1])).astype(np.float32)
    Q = Q.astype(np.float32)
    ids, ds, cnt, cmps, vis = oracle.hnsw_search(adj, X, metric, Q, 10, 64, [ent])
    gt, _ = oracle.groundtruth_f64(X, metric, Q, 10)
    rec = np.mean([len(set(ids[i, :10].tolist()) & set(gt[i].tolist())) / 10
                   for i in range(len(Q))])
    assert rec > 0.95


def test_build_thread_independent(oracle):
    _, _, a1, e1 = build_small(oracle, nthreads=1, n=800)
    _, _, a8, e8 = build_small(oracle, nthreads=8, n=800)
    assert e1 == e8
    for x, y in zip(a1, a8):
        assert np.array_equal(x, y)


def test_batch_bounds():
    for n in (2, 10, 49, 50, 1000, 12345):
        b = hnsw.batch_bounds(n)
        assert b[0][0] == 1 and b[-1][1] == n
        assert all(b[i][1] == b[i + 1][0] for i in range(len(b) - 1))
        lim = max(1, int(n * 0.02))
        assert all(0 < e - s <= lim for s, e in b)
    # ceil(2 b) + 1 while below the 0.02 n limit
    assert hnsw.batch_bounds(10 ** 6)[:4] == [(1, 3), (3, 7), (7, 15), (15, 31)]


def test_model_file_layout(tmp_path):
    n = 5
    g0 = np.zeros((n, 5), np.uint32)
    g0[0, :3] = [2, 1, 2]
    g0[1, :2] = [1, 0]
    g1 = np.zeros((n, 3), np.uint32)
    g1[2, :2] = [1, 0]
    mdl = hnsw.HNSWModel(dim=7, m_l=0.5, m=2, efc=9, alpha=1.25,
                         node_levels=np.array([1, 0, 1, 0, 0], np.uint32),
                         node_ids=np.array([4, 3, 2, 1, 0], np.uint32),
                         graphs=[g0, g1], entrance=np.array([2], np.uint32), code_u=77,
                         size_node=40)
    p = str(tmp_path / "m.hnsw")
    hnsw.write_model(p, mdl)
    b = open(p, "rb").read()
    # the loader's read order (HNSW.hpp:667-744)
    assert b[:4] == b"HNSW"
    assert struct.unpack_from("<I", b, 4)[0] == 3
    assert struct.unpack_from("<QQ", b, 8) == (77, 40)
    dim, m_l, m, efc, alpha, nn = struct.unpack_from("<IfIIfI", b, 24)
    assert (dim, m, efc, nn) == (7, 2, 9, n) and m_l == 0.5 and alpha == 1.25
    pos = 48
    assert list(struct.unpack_from("<10I", b, pos)) == [1, 4, 0, 3, 1, 2, 0, 1, 0, 0]
    pos += 40
    # node 0: level 0 (2 ids), level 1 (0 ids); node 1: level 0 (1 id)
    assert struct.unpack_from("<Q", b, pos)[0] == 2
    assert struct.unpack_from("<2I", b, pos + 8) == (1, 2)
    pos += 16
    assert struct.unpack_from("<Q", b, pos)[0] == 0
    pos += 8
    assert struct.unpack_from("<QI", b, pos) == (1, 0)
    assert b[-12:] == struct.pack("<QI", 1, 2)
    r = hnsw.read_model(p)
    assert np.array_equal(r.graphs[0], g0) and np.array_equal(r.graphs[1], g1)
    assert np.array_equal(r.node_ids, mdl.node_ids) and r.top_level == 1
