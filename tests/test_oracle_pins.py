"""Pins the oracle (oracle/pann_oracle.c) against what the mathematics fixes:
closed forms, exact integer arithmetic, an independent in-order float32
emulation, Python set semantics, brute force on tiny inputs and special
graphs whose beam search must return the exact k nearest neighbours.
Each check is chosen so that a plausible restatement mistake (a fused or
reordered float sum, a dropped deferral rule, a wrong sign or index) fails."""
import numpy as np
import pytest


def inorder_f32(terms):
    acc = np.float32(0.0)
    for t in terms:
        acc = np.float32(acc + np.float32(t))
    return acc


# ---------------------------------------------------------------- distances
@pytest.mark.parametrize("d", [1, 3, 8, 17, 128])
def test_u8_i8_distances_exact(oracle, rng, d):
    for _ in range(50):
        p = rng.integers(0, 256, d, dtype=np.uint8)
        q = rng.integers(0, 256, d, dtype=np.uint8)
        l2 = int(((p.astype(np.int64) - q.astype(np.int64)) ** 2).sum())
        ip = int((p.astype(np.int64) * q.astype(np.int64)).sum())
        assert oracle.distance(p, q, "l2") == float(np.float32(l2))
        assert oracle.distance(p, q, "ip") == float(-np.float32(ip))
        ps, qs = p.view(np.int8), q.view(np.int8)
        l2s = int(((ps.astype(np.int64) - qs.astype(np.int64)) ** 2).sum())
        ips = int((ps.astype(np.int64) * qs.astype(np.int64)).sum())
        assert oracle.distance(ps, qs, "l2") == float(np.float32(l2s))
        assert oracle.distance(ps, qs, "ip") == float(-np.float32(ips))


def test_u8_extremes(oracle):
    z = np.zeros(128, np.uint8)
    f = np.full(128, 255, np.uint8)
    assert oracle.distance(z, f, "l2") == 128 * 255 * 255  # 8,323,200 < 2^24: exact
    assert oracle.distance(f, f, "ip") == -128 * 255 * 255


@pytest.mark.parametrize("d", [96, 128, 200])
def test_f32_distance_is_inorder_sum(oracle, rng, d):
    """Bit-exact against an independent in-order float32 emulation; a fused
    multiply-add or a pairwise/tree sum would differ in the last bits."""
    mism_tree = 0
    for _ in range(40):
        p = rng.standard_normal(d).astype(np.float32)
        q = rng.standard_normal(d).astype(np.float32)
        t = (q - p).astype(np.float32)
        sq = (t * t).astype(np.float32)
        assert oracle.distance(p, q, "l2") == inorder_f32(sq)
        prod = (q * p).astype(np.float32)
        assert oracle.distance(p, q, "ip") == -inorder_f32(prod)
        # closeness to float64 (bounded by d * eps * sum)
        ref = float(((p.astype(np.float64) - q.astype(np.float64)) ** 2).sum())
        assert abs(oracle.distance(p, q, "l2") - ref) <= d * 1.2e-7 * ref
        mism_tree += oracle.distance(p, q, "l2") != np.float32(np.sum(sq))  # numpy: pairwise
    assert mism_tree > 0  # the emulation really distinguishes summation orders


def test_mips_negative_zero(oracle):
    p = np.zeros(8, np.float32)
    q = np.ones(8, np.float32)
    d = oracle.distance(p, q, "ip")
    assert d == 0.0 and np.signbit(np.float32(d))  # -(+0.0) as in mips_point.h:65


# ------------------------------------------------------------------ hashset
def test_hashset_matches_set_until_growth(oracle, rng):
    h = oracle.HashSet(2 * (10 + 64) * 32)  # beamSearch.h:66 sizing
    assert h.capacity == 8192
    ref = set()
    keys = rng.integers(0, 5000, 3000)
    for k in keys:
        assert h(int(k)) == (int(k) in ref)
        ref.add(int(k))
    assert not h.grew and h.num_entries == len(ref)


def test_hashset_colliding_keys(oracle):
    # keys that share the low hash bits probe linearly and stay exact
    h = oracle.HashSet(64)
    mask = 63
    keys = [k for k in range(200000) if ((k * 0xbf58476d1ce4e5b9) & ((1 << 64) - 1)) & mask == 5][:20]
    for k in keys:
        assert h(k) is False
    for k in keys:
        assert h(k) is True


def test_hashset_grows_past_half(oracle):
    h = oracle.HashSet(16)
    for k in range(9):
        h(k)
    assert not h.grew
    h(100)  # num_entries 9 > 16/2: grows (hashset.h:25)
    assert h.grew and h.capacity == 32


# --------------------------------------------------------------------- sort
def test_sort_pairs(oracle, rng):
    for n in [0, 1, 5, 24, 25, 100, 1000]:
        ids = rng.permutation(10 * n + 1)[:n].astype(np.uint32)
        d = rng.integers(0, 7, n).astype(np.float32)  # many ties
        si, sd = oracle.sort_pairs(ids, d)
        order = sorted(zip(d.tolist(), ids.tolist()))
        assert list(zip(sd.tolist(), si.tolist())) == order


# -------------------------------------------------------------- beam search
def _complete_graph(n):
    adj = np.zeros((n, n), np.uint32)
    for i in range(n):
        nb = [j for j in range(n) if j != i]
        adj[i, 0] = len(nb)
        adj[i, 1:1 + len(nb)] = nb
    return adj


def _brute(X, q, metric, k):
    Xd, qd = X.astype(np.float64), q.astype(np.float64)
    d = ((Xd - qd) ** 2).sum(1) if metric == "l2" else -(Xd @ qd)
    o = np.lexsort((np.arange(len(X)), d))
    return o[:k]


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_complete_graph_is_exact_knn(oracle, rng, metric):
    X = rng.standard_normal((60, 16)).astype(np.float32)
    adj = _complete_graph(60)
    for _ in range(10):
        q = rng.standard_normal(16).astype(np.float32)
        r = oracle.beam_search(adj, X, metric, q, [0], 5, 8)
        assert list(r["ids"][:5]) == list(_brute(X, q, metric, 5))
        assert r["num_visited"] >= 1 and r["full_dist_cmps"] == 60  # start + 59 neighbours


def test_beam_ge_n_on_connected_graph_is_exact(oracle, rng):
    n = 200
    X = rng.standard_normal((n, 8)).astype(np.float32)
    # ring plus random chords: connected, degree <= 6
    adj = np.zeros((n, 7), np.uint32)
    for i in range(n):
        nb = {(i + 1) % n, (i - 1) % n}
        while len(nb) < 6:
            j = int(rng.integers(0, n))
            if j != i:
                nb.add(j)
        nb = sorted(nb)
        adj[i, 0] = len(nb)
        adj[i, 1:1 + len(nb)] = nb
    for _ in range(10):
        q = rng.standard_normal(8).astype(np.float32)
        r = oracle.beam_search(adj, X, "l2", q, [0], 10, n)
        assert list(r["ids"][:10]) == list(_brute(X, q, "l2", 10))
        assert r["num_visited"] == n  # every reachable node is visited


def test_path_graph_greedy_walk(oracle):
    """Points 0..n-1 on a line, path graph, query beyond the far end, beam 1:
    the search walks the whole path one node per step."""
    n = 30
    X = np.arange(n, dtype=np.float32)[:, None]
    adj = np.zeros((n, 3), np.uint32)
    for i in range(n):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
        adj[i, 0] = len(nb)
        adj[i, 1:1 + len(nb)] = nb
    r = oracle.beam_search(adj, X, "l2", np.array([n + 5], np.float32), [0], 1, 1)
    assert list(r["ids"]) == [n - 1]
    assert r["num_visited"] == n
    assert list(r["visited_ids"]) == list(range(n - 1, -1, -1))  # sorted by distance


def test_deferred_merge_rule(oracle):
    """Hand-traced deferral (beamSearch.h:185-192) on 1-D points, query at 0.
    Node 0 (x=10) -> A=1 (x=5), B=2 (x=6); A -> C=3 (x=1); B -> D=4 (x=7).
    Beam 2, batch_factor 1 (defer while |C| < 2):
      visit 0: cands A,B -> merge -> frontier [A, B];
      visit A: cutoff 36, cand C (1) -> |C|=1 < 2 and B still unvisited: defer;
      visit B (stale frontier): D (49 >= 36) compared and dropped -> merge,
        frontier [C, A];  visit C: no neighbours.
    -> 4 visits, 5 distance comparisons.  Merging eagerly (batch_factor 0)
    never visits B: 3 visits, 4 comparisons."""
    X = np.array([[10.0], [5.0], [6.0], [1.0], [7.0]], np.float32)
    adj = np.zeros((5, 3), np.uint32)
    adj[0, 0], adj[0, 1:3] = 2, [1, 2]
    adj[1, 0], adj[1, 1] = 1, 3
    adj[2, 0], adj[2, 1] = 1, 4
    q = np.array([0.0], np.float32)
    r = oracle.beam_search(adj, X, "l2", q, [0], 1, 2, batch_factor=1.0)
    assert list(r["ids"]) == [3, 1]
    assert list(r["visited_ids"]) == [3, 1, 2, 0]
    assert r["num_visited"] == 4 and r["full_dist_cmps"] == 5
    r = oracle.beam_search(adj, X, "l2", q, [0], 1, 2, batch_factor=0.0)
    assert list(r["ids"]) == [3, 1]
    assert list(r["visited_ids"]) == [3, 1, 0]
    assert r["num_visited"] == 3 and r["full_dist_cmps"] == 4
    # deferral is off when limit < 2 * beam (:186)
    r = oracle.beam_search(adj, X, "l2", q, [0], 1, 2, batch_factor=1.0, limit=3)
    assert r["num_visited"] == 3 and r["full_dist_cmps"] == 4


def test_limit_bounds_visits(oracle, rng):
    X = rng.standard_normal((100, 4)).astype(np.float32)
    adj = _complete_graph(100)
    r = oracle.beam_search(adj, X, "l2", X[3] + 0.01, [0], 5, 50, limit=1)
    assert r["num_visited"] == 1


def test_self_is_skipped(oracle, rng):
    X = rng.standard_normal((30, 4)).astype(np.float32)
    adj = _complete_graph(30)
    r = oracle.beam_search(adj, X, "l2", X[7], [0], 5, 30, self_id=7)
    assert 7 not in set(r["ids"].tolist())


def test_invariants_random_graph(oracle, rng):
    n, d = 500, 16
    X = rng.standard_normal((n, d)).astype(np.float32)
    adj = np.zeros((n, 9), np.uint32)
    for i in range(n):
        nb = rng.choice(n, 8, replace=False)
        nb = nb[nb != i]
        adj[i, 0] = len(nb)
        adj[i, 1:1 + len(nb)] = nb
    for B in [4, 16, 64]:
        q = rng.standard_normal(d).astype(np.float32)
        r = oracle.beam_search(adj, X, "l2", q, [0], 1, B)
        ids, ds = r["ids"], r["dists"]
        assert len(ids) <= B and len(set(ids.tolist())) == len(ids)
        assert all((ds[i], ids[i]) < (ds[i + 1], ids[i + 1]) for i in range(len(ids) - 1))
        for i, dd in zip(ids, ds):
            assert dd == oracle.distance(X[i], q, "l2")
        assert r["dist_cmps"] == r["full_dist_cmps"]


def test_short_beam_and_errors(oracle, rng):
    X = rng.standard_normal((10, 4)).astype(np.float32)
    adj = np.zeros((10, 3), np.uint32)  # no edges
    with pytest.raises(ValueError):
        oracle.search_batch(adj, X, "l2", X[:2], 5, 8)  # beam returns 1 < k
    with pytest.raises(ValueError):
        oracle.search_batch(adj, X, "l2", X[:2], 9, 8)  # k > beam


# -------------------------------------------------------------- robustPrune
def test_prune_huge_alpha_keeps_R_nearest(oracle, rng):
    n = 50
    X = rng.standard_normal((n, 6)).astype(np.float32)
    adj = np.zeros((n, 11), np.uint32)
    p = 0
    cand = np.arange(1, n, dtype=np.uint32)
    cd = np.array([oracle.distance(X[c], X[p], "l2") for c in cand], np.float32)
    out, _ = oracle.robust_prune(adj, X, "l2", p, cand, cd, 1e30, 10)
    expect = [int(c) for _, c in sorted(zip(cd.tolist(), cand.tolist()))[:10]]
    assert list(out) == expect


def test_prune_alpha1_collinear(oracle):
    # p = 0 at x=0; candidates at 1,2,3 (right) and -1.5,-2.5 (left)
    X = np.array([[0.0], [1.0], [2.0], [3.0], [-1.5], [-2.5]], np.float32)
    adj = np.zeros((6, 6), np.uint32)
    cand = np.arange(1, 6, dtype=np.uint32)
    cd = np.array([oracle.distance(X[c], X[0], "l2") for c in cand], np.float32)
    out, _ = oracle.robust_prune(adj, X, "l2", 0, cand, cd, 1.0, 5)
    assert sorted(out.tolist()) == [1, 4]  # nearest on each side


def test_prune_adds_out_neighbours_and_dedups(oracle):
    X = np.array([[0.0], [1.0], [5.0], [-2.0]], np.float32)
    adj = np.zeros((4, 4), np.uint32)
    adj[0, 0] = 2
    adj[0, 1:3] = [2, 3]  # current out-neighbours of p = 0
    cand = np.array([1, 2], np.uint32)  # 2 duplicated with the out-list
    cd = np.array([1.0, 25.0], np.float32)
    out, _ = oracle.robust_prune(adj, X, "l2", 0, cand, cd, 1.0, 3)
    assert sorted(out.tolist()) == [1, 3]  # 2 pruned by 1 (d(1,2)=16 <= 25)


# -------------------------------------------------------------------- build
def test_build_invariants_and_recall(oracle):
    from parlayann_amd import datagen
    X, Q = datagen.workload_f32_l2(n=3000, nq=200)
    perm = datagen.insertion_orders(len(X), 1, 5)
    adj = oracle.build_vamana(X, "l2", 16, 32, 1.2, perm, nthreads=4)
    deg = adj[:, 0]
    assert deg.max() <= 16 and deg.min() >= 1
    for i in range(0, 3000, 97):
        row = adj[i, 1:1 + deg[i]]
        assert i not in row and len(set(row.tolist())) == len(row)
    gt, _ = oracle.groundtruth_f64(X, "l2", Q, 10)
    r = oracle.search_batch(adj, X, "l2", Q, 10, 32)
    rec = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(r["ids"], gt)])
    assert rec > 0.95
    # deterministic in the thread count
    adj1 = oracle.build_vamana(X, "l2", 16, 32, 1.2, perm, nthreads=1)
    assert np.array_equal(adj, adj1)


# --------------------------------------------------------------- groundtruth
def test_groundtruth_f64(oracle, rng):
    X = rng.standard_normal((500, 12)).astype(np.float32)
    Q = rng.standard_normal((20, 12)).astype(np.float32)
    ids, d = oracle.groundtruth_f64(X, "l2", Q, 7)
    for i in range(20):
        assert list(ids[i]) == list(_brute(X, Q[i], "l2", 7))
    ids, d = oracle.groundtruth_f64(X, "ip", Q, 7)
    for i in range(20):
        assert list(ids[i]) == list(_brute(X, Q[i], "ip", 7))
