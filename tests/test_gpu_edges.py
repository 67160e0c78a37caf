"""GPU edge cases against the oracle, through the C ABI, on uint8 data (so the
bar is bit-exact: ids, distances, visited and dist_cmps counts).

Cases the reference's search meets in practice (beamSearch.h:33-228,
:599-646): an empty batch, k = beam = 1, the largest beam on a small graph
(beam >= n gives the exact nearest neighbours of the reachable set),
ragged adjacency rows (degree 0, partial rows whose slots past the degree
hold stale ids that must be ignored, graph.h:59-64), repeated neighbour
entries (the seen set drops the second, beamSearch.h:151-159), duplicated
points (ties broken by id, :62-64) and queries that are base points
(distance 0; same_as is pointer identity, so the point itself is a result).
"""
import sys

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

from parlayann_amd import DeviceIndex, datagen  # noqa: E402
from parlayann_amd._lib import PannError  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _assert_exact(g, o):
    assert np.array_equal(g["ids"], o["ids"]), "ids differ"
    assert np.array_equal(g["dists"], o["dists"]), "dists differ"
    assert np.array_equal(g["visited"], o["visited"]), "visited counts differ"
    assert np.array_equal(g["dist_cmps"].astype(np.uint64), o["dist_cmps"]), "cmps differ"


@pytest.fixture(scope="module")
def small(oracle):
    X, Q = datagen.workload_u8_l2(n=3000, nq=120, d=32, centers=40)
    perm = datagen.insertion_orders(len(X), 1, 21)
    adj = oracle.build_vamana(X, "l2", 24, 48, 1.2, perm)
    return X, Q, adj


def test_empty_batch(small):
    X, Q, adj = small
    _say("[edges] empty batch")
    idx = DeviceIndex(X, adj, "l2")
    r = idx.search(Q[:0], 10, 32)
    assert r["ids"].shape == (0, 10) and r["dists"].shape == (0, 10)
    with pytest.raises(PannError):  # qsearchAll checks k <= beam before any query
        idx.search(Q[:0], 11, 10)
    idx.close()


def test_k1_beam1(oracle, small):
    X, Q, adj = small
    _say("[edges] k = beam = 1")
    o = oracle.search_batch(adj, X, "l2", Q, 1, 1)
    g = DeviceIndex(X, adj, "l2").search(Q, 1, 1, exact_counts=True)
    _assert_exact(g, o)


def test_beam_at_maximum_is_exact_knn(oracle, small):
    """beam 4096 > n with limit = n: every reachable point enters the
    frontier, so the result is the exact kNN of the reachable set."""
    X, Q, adj = small
    _say("[edges] beam 4096 on n = 3000")
    k = 100
    o = oracle.search_batch(adj, X, "l2", Q[:24], k, 4096)
    g = DeviceIndex(X, adj, "l2").search(Q[:24], k, 4096, exact_counts=True)
    _assert_exact(g, o)
    reach = np.zeros(len(X), bool)  # reachable from the start point 0
    stack = [0]
    reach[0] = True
    while stack:
        u = stack.pop()
        for v in adj[u, 1:1 + adj[u, 0]]:
            if not reach[v]:
                reach[v] = True
                stack.append(v)
    ids = np.nonzero(reach)[0]
    for i in range(24):
        d = ((X[ids].astype(np.int64) - Q[i].astype(np.int64)) ** 2).sum(1)
        order = np.lexsort((ids, d))[:k]  # (distance, id)
        assert np.array_equal(g["ids"][i], ids[order].astype(np.uint32))
        assert np.array_equal(g["dists"][i], d[order].astype(np.float32))


def test_ragged_rows_and_stale_slots(oracle, small):
    X, Q, adj = small
    _say("[edges] ragged rows, degree 0, stale slots")
    rng = np.random.default_rng(5)
    a = adj.copy()
    md = a.shape[1] - 1
    n = len(X)
    for u in range(1, n):
        r = rng.random()
        if r < 0.05:
            a[u, 0] = 0
        elif r < 0.5:
            a[u, 0] = rng.integers(0, a[u, 0] + 1)
    # slots past each row's degree hold valid-looking stale ids
    for u in range(n):
        deg = a[u, 0]
        a[u, 1 + deg:] = rng.integers(0, n, md - deg, dtype=np.uint32)
    o = oracle.search_batch(a, X, "l2", Q, 10, 64)
    g = DeviceIndex(X, a, "l2").search(Q, 10, 64, exact_counts=True)
    _assert_exact(g, o)
    # a start point of degree 0: the frontier is the start alone
    a0 = a.copy()
    a0[0, 0] = 0
    idx = DeviceIndex(X, a0, "l2")
    o1 = oracle.search_batch(a0, X, "l2", Q, 1, 16)
    _assert_exact(idx.search(Q, 1, 16, exact_counts=True), o1)
    with pytest.raises(PannError):  # fewer than k results (beamSearch.h:557-562)
        idx.search(Q, 10, 16)
    idx.close()


def test_repeated_neighbour_entries(oracle, small):
    X, Q, adj = small
    _say("[edges] repeated neighbour entries")
    a = adj.copy()
    for u in range(len(X)):
        deg = int(a[u, 0])
        if deg >= 4:  # repeat the first two neighbours over the last two slots
            a[u, 1 + deg - 2:1 + deg] = a[u, 1:3]
    o = oracle.search_batch(a, X, "l2", Q, 10, 48)
    idx = DeviceIndex(X, a, "l2")
    _assert_exact(idx.search(Q, 10, 48, exact_counts=True), o)
    f = idx.search(Q, 10, 48)  # default seen set: same results, more evaluations
    for key in ("ids", "dists", "visited"):
        assert np.array_equal(f[key], o[key]), key
    assert np.all(f["dist_cmps"].astype(np.uint64) >= o["dist_cmps"])
    idx.close()


def test_duplicate_points_and_base_queries(oracle, small):
    X, Q, adj = small
    _say("[edges] duplicate points, queries that are base points")
    X2 = X.copy()
    X2[1500:1800] = X2[0:300]  # 300 exact duplicates: ties broken by id
    Qb = np.concatenate([X2[[0, 5, 1500, 1799, 2999]], Q[:40]])
    o = oracle.search_batch(adj, X2, "l2", Qb, 10, 64)
    g = DeviceIndex(X2, adj, "l2").search(Qb, 10, 64, exact_counts=True)
    _assert_exact(g, o)
    assert np.all(g["dists"][:5, 0] == 0.0)


def test_exact_seen_set_grows():
    """The exact seen set outgrows its global table and grows, as the
    reference's hashset does past half load (hashset.h:25-38).  On a path
    graph (node i linked to i - 1 and i + 1; point i has coordinates
    clip(i - 255 j, 0, 255), a u8 staircase whose distance to the last point
    falls at every step) a query at the far end walks the whole path from
    start 0: every one of the n ids is
    seen exactly once and every node is visited, so dist_cmps = visited = n,
    far past the table's first size (2^13 slots at beam 10, degree 2).  A
    synchronous search grows the table and returns the default seen set's
    ids and distances; an async one reports PANN_ERR_CAPACITY."""
    import torch
    _say("exact seen set growth (path graph)")
    n = 20000
    X = np.clip(np.arange(n)[:, None] - 255 * np.arange(80)[None, :], 0, 255).astype(np.uint8)
    adj = np.zeros((n, 3), np.uint32)
    for i in range(n):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
        adj[i, 0] = len(nb)
        adj[i, 1:1 + len(nb)] = nb
    Q = X[[n - 1, n - 1, n - 2]].copy()
    Q[1, 79] = 3  # off the line
    # async on a fresh index: the table cannot grow, the status says so
    idx = DeviceIndex(X, adj, "l2")
    dev = torch.device("cuda", 0)
    tq = torch.from_numpy(Q).to(dev)
    oi = torch.zeros((len(Q), 10), dtype=torch.int32, device=dev)
    od = torch.zeros((len(Q), 10), dtype=torch.float32, device=dev)
    idx.search_device(tq, 10, 10, oi, od, exact_counts=True, async_=True)
    with pytest.raises(PannError):
        idx.search_status()
    g = idx.search(Q, 10, 10, exact_counts=True)
    d = idx.search(Q, 10, 10)
    for key in ("ids", "dists", "visited"):
        assert np.array_equal(g[key], d[key]), key
    assert np.all(g["visited"] == n) and np.all(g["dist_cmps"] == n)
    assert np.all(g["ids"][:2, 0] == n - 1)
    # the grown table stays: an async search now fits
    idx.search_device(tq, 10, 10, oi, od, exact_counts=True, async_=True)
    idx.search_status()
    assert np.array_equal(oi.cpu().numpy().view(np.uint32), g["ids"])
    idx.close()
