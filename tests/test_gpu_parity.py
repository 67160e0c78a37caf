"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
seeded inputs and the same graph file.

Bars (the tolerance is part of each test):
  * uint8/int8 L2 and IP: bit-exact ids, distances, visited counts and
    distance-comparison counts;
  * float32 with in-order summation: bit-exact, like the integer paths;
  * float32 with the tree (DPP) summation: identical ids except where the
    two distances differ by at most 1e-6 relative (ties), >= 99% of queries
    fully identical;
  * BF16 / FP8 gathers + FP32 rerank: recall@10 against FP64 groundtruth
    within 0.02 of the plain FP32 search, reranked distances equal to the
    in-order FP32 distance (relative error vs FP64 <= 1e-6).
"""
import os
import sys

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

from parlayann_amd import DeviceIndex, datagen  # noqa: E402
from parlayann_amd.recall import recall_plain  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def u8_case(oracle):
    X, Q = datagen.workload_u8_l2(n=20000, nq=400)
    perm = datagen.insertion_orders(len(X), 1, 11)
    adj = oracle.build_vamana(X, "l2", 64, 128, 1.2, perm)
    return X, Q, adj


@pytest.fixture(scope="module")
def f32_case(oracle):
    X, Q = datagen.workload_f32_l2(n=20000, nq=400)
    perm = datagen.insertion_orders(len(X), 1, 12)
    adj = oracle.build_vamana(X, "l2", 32, 64, 1.2, perm)
    return X, Q, adj


@pytest.fixture(scope="module")
def ip_case(oracle):
    X, Q = datagen.workload_aniso_ip(n=20000, nq=400)
    perm = datagen.insertion_orders(len(X), 1, 13)
    adj = oracle.build_vamana(X, "ip", 32, 64, 1.0, perm)
    return X, Q, adj


def _assert_exact(g, o, k):
    assert np.array_equal(g["ids"], o["ids"]), "ids differ"
    assert np.array_equal(g["dists"], o["dists"]), "dists differ"
    assert np.array_equal(g["visited"], o["visited"]), "visited counts differ"
    assert np.array_equal(g["dist_cmps"].astype(np.uint64), o["dist_cmps"]), "cmps differ"


@pytest.mark.parametrize("B", [10, 20, 50, 100, 200])
def test_u8_l2_bit_exact(oracle, u8_case, B):
    X, Q, adj = u8_case
    _say(f"u8 l2 B={B}")
    o = oracle.search_batch(adj, X, "l2", Q, 10, B)
    assert not o["hash_grew"].any()
    idx = DeviceIndex(X, adj, "l2")
    g = idx.search(Q, 10, B, exact_counts=True)
    _assert_exact(g, o, 10)
    # the accepted-only seen set: same results, at least as many evaluations
    f = idx.search(Q, 10, B)
    assert np.array_equal(f["ids"], o["ids"]) and np.array_equal(f["dists"], o["dists"])
    assert np.array_equal(f["visited"], o["visited"])
    assert np.all(f["dist_cmps"].astype(np.uint64) >= o["dist_cmps"])


def test_u8_ip_and_i8_bit_exact(oracle, u8_case):
    X, Q, adj = u8_case
    _say("u8 ip / i8 l2")
    o = oracle.search_batch(adj, X, "ip", Q, 10, 32)
    g = DeviceIndex(X, adj, "ip").search(Q, 10, 32, exact_counts=True)
    _assert_exact(g, o, 10)
    Xs = (X.astype(np.int16) - 128).astype(np.int8)
    Qs = (Q.astype(np.int16) - 128).astype(np.int8)
    for metric in ("l2", "ip"):
        o = oracle.search_batch(adj, Xs, metric, Qs, 10, 32)
        g = DeviceIndex(Xs, adj, metric).search(Qs, 10, 32, exact_counts=True)
        _assert_exact(g, o, 10)


@pytest.mark.parametrize("metric,B", [("l2", 10), ("l2", 64), ("l2", 128), ("ip", 16),
                                      ("ip", 64)])
def test_f32_inorder_bit_exact(oracle, f32_case, ip_case, metric, B):
    X, Q, adj = f32_case if metric == "l2" else ip_case
    _say(f"f32 inorder {metric} B={B}")
    o = oracle.search_batch(adj, X, metric, Q, 10, B)
    assert not o["hash_grew"].any()
    idx = DeviceIndex(X, adj, metric)
    g = idx.search(Q, 10, B, summation="inorder", exact_counts=True)
    _assert_exact(g, o, 10)
    f = idx.search(Q, 10, B, summation="inorder")
    assert np.array_equal(f["ids"], o["ids"]) and np.array_equal(f["dists"], o["dists"])
    assert np.array_equal(f["visited"], o["visited"])


@pytest.mark.parametrize("dim", [5, 13, 32, 40, 99, 100, 101, 130, 200, 203, 256, 300, 301, 600,
                                 601])
@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_f32_inorder_bit_exact_dims(oracle, metric, dim):
    """In-order FP32 at every row shape: rows of 1 to 8 128-byte lines (8-lane
    gather, sum carried lane to lane in element order) and rows of 10 and 19
    lines (8 lines at a time, query in LDS); dimensions that are not multiples
    of 4 end with the fused terms of the reference loop's compiled tail, in
    every lane position of the group (dim / 4 % 8 = 1, 3, 0, 0, 2, 5, 3, 2)."""
    rng = np.random.default_rng(1000 + dim)
    X = rng.standard_normal((3000, dim)).astype(np.float32)
    Q = rng.standard_normal((100, dim)).astype(np.float32)
    _say(f"f32 inorder {metric} dim={dim}")
    perm = datagen.insertion_orders(len(X), 1, 21)
    adj = oracle.build_vamana(X, metric, 16, 32, 1.2 if metric == "l2" else 1.0, perm)
    o = oracle.search_batch(adj, X, metric, Q, 10, 32)
    assert not o["hash_grew"].any()
    idx = DeviceIndex(X, adj, metric)
    g = idx.search(Q, 10, 32, summation="inorder", exact_counts=True)
    _assert_exact(g, o, 10)
    f = idx.search(Q, 10, 32, summation="inorder")
    assert np.array_equal(f["ids"], o["ids"]) and np.array_equal(f["dists"], o["dists"])
    idx.close()


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_f32_tree_ids_up_to_ties(oracle, f32_case, ip_case, metric):
    X, Q, adj = f32_case if metric == "l2" else ip_case
    _say(f"f32 tree {metric}")
    o = oracle.search_batch(adj, X, metric, Q, 10, 64)
    g = DeviceIndex(X, adj, metric).search(Q, 10, 64, summation="tree")
    same = np.all(g["ids"] == o["ids"], axis=1)
    assert same.mean() >= 0.99, same.mean()
    for i in np.nonzero(~same)[0]:
        # every differing slot holds a distance tied (1e-6 rel) with the oracle's
        for a, b in zip(g["ids"][i], o["ids"][i]):
            if a != b:
                da = oracle.distance(X[a], Q[i], metric)
                db = oracle.distance(X[b], Q[i], metric)
                assert abs(da - db) <= 1e-6 * max(abs(da), abs(db)), (i, a, b, da, db)
    # tree distances agree with the in-order ones to float rounding
    rel = np.abs(g["dists"] - o["dists"]) / np.maximum(np.abs(o["dists"]), 1e-30)
    assert rel[same].max() <= 1e-5


@pytest.mark.parametrize("d", [140, 180, 200])  # f32 rows of 5, 6 and 7 128-byte lines
def test_f32_tree_rows_of_5_to_7_lines(d):
    """The gather's 8-chunk shape on rows of 5-7 lines: the chunks past a
    row's last line are not loaded (they are the next row's, or past the last
    row).  On a complete graph with beam = n every row, the last one included,
    is evaluated; ids come back in exact order and distances match FP64."""
    _say(f"f32 tree rows d={d}")
    rng = np.random.default_rng(d)
    n = 61
    X = rng.standard_normal((n, d)).astype(np.float32)
    Q = rng.standard_normal((6, d)).astype(np.float32)
    for metric in ("l2", "ip"):
        r = DeviceIndex(X, _complete(n), metric).search(Q, n, n, summation="tree")
        for i in range(len(Q)):
            q = Q[i].astype(np.float64)
            ex = (((X.astype(np.float64) - q) ** 2).sum(1) if metric == "l2"
                  else -(X.astype(np.float64) @ q))
            assert np.array_equal(r["ids"][i], np.argsort(ex, kind="stable")), (metric, i)
            got = r["dists"][i].astype(np.float64)
            want = ex[r["ids"][i]]
            assert np.all(np.abs(got - want) <= 1e-5 * np.maximum(np.abs(want), 1.0)), (metric, i)


def test_hash_overflow_path_is_exact(oracle, u8_case, f32_case):
    """A 256-slot LDS table overflows into the per-wave global table after
    128 ids; results must not change."""
    X, Q, adj = u8_case
    _say("hash overflow")
    o = oracle.search_batch(adj, X, "l2", Q, 10, 100)
    idx = DeviceIndex(X, adj, "l2")
    g = idx.search(Q, 10, 100, hash_slots_log2=8, exact_counts=True)
    _assert_exact(g, o, 10)
    f = idx.search(Q, 10, 100, hash_slots_log2=8)
    assert np.array_equal(f["ids"], o["ids"]) and np.array_equal(f["visited"], o["visited"])
    X, Q, adj = f32_case
    o = oracle.search_batch(adj, X, "l2", Q, 10, 64)
    idx = DeviceIndex(X, adj, "l2")
    g = idx.search(Q, 10, 64, hash_slots_log2=8, exact_counts=True)
    _assert_exact(g, o, 10)
    f = idx.search(Q, 10, 64, hash_slots_log2=8)
    assert np.array_equal(f["ids"], o["ids"]) and np.array_equal(f["visited"], o["visited"])


def test_limit_and_degree_limit(oracle, f32_case):
    X, Q, adj = f32_case
    _say("limit / degree limit")
    for lim, dl in [(15, 32), (20, 16), (35, 8)]:
        B = max(lim, 10)
        o = oracle.search_batch(adj, X, "l2", Q, 10, B, limit=lim, degree_limit=dl)
        g = DeviceIndex(X, adj, "l2").search(Q, 10, B, limit=lim, degree_limit=dl,
                                             exact_counts=True)
        _assert_exact(g, o, 10)


def test_per_query_k_and_beam(oracle, u8_case):
    import torch
    X, Q, adj = u8_case
    _say("per-query k / beam")
    rng = np.random.default_rng(7)
    nq = len(Q)
    ks = rng.choice([1, 10, 50], nq).astype(np.int32)
    bs = np.array([rng.integers(max(16, k), 129) for k in ks], np.int32)
    idx = DeviceIndex(X, adj, "l2")
    dev = torch.device("cuda")
    tq = torch.from_numpy(Q).to(dev)
    out = []
    # query_order 0 runs the batch longest-first (a device sort by beam), 1 as
    # given: the same results either way
    for qo in (0, 1):
        oi = torch.zeros((nq, 50), dtype=torch.int32, device=dev)
        od = torch.zeros((nq, 50), dtype=torch.float32, device=dev)
        ov = torch.zeros(nq, dtype=torch.int32, device=dev)
        oc = torch.zeros(nq, dtype=torch.int32, device=dev)
        idx.search_device(tq, 50, int(bs.max()), oi, od, out_visited=ov, out_cmps=oc,
                          k_per_query=torch.from_numpy(ks).to(dev),
                          beam_per_query=torch.from_numpy(bs).to(dev), query_order=qo)
        torch.cuda.synchronize()
        out.append([t.cpu().numpy() for t in (oi, od, ov, oc)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    gi, gd = out[0][0].view(np.uint32), out[0][1]
    for B in np.unique(bs):
        sel = np.nonzero(bs == B)[0]
        kmax = int(ks[sel].max())
        o = oracle.search_batch(adj, X, "l2", Q[sel], kmax, int(B))
        for j, qi in enumerate(sel):
            k = ks[qi]
            assert np.array_equal(gi[qi, :k], o["ids"][j, :k])
            assert np.array_equal(gd[qi, :k], o["dists"][j, :k])


@pytest.mark.parametrize("store", ["bf16", "fp8"])
def test_reduced_precision_rerank(oracle, ip_case, store):
    from parlayann_amd.groundtruth import groundtruth
    X, Q, adj = ip_case
    _say(f"{store} + rerank")
    gt, gd = groundtruth(X, Q, 10, "ip")
    idx = DeviceIndex(X, adj, "ip")
    idx.build_store(store)
    plain = idx.search(Q, 10, 64, summation="inorder")
    red = idx.search(Q, 10, 64, store=store, rerank=True)
    r_plain = recall_plain(gt, plain["ids"], 10)
    r_red = recall_plain(gt, red["ids"], 10)
    assert r_red >= r_plain - 0.02, (r_plain, r_red)
    # reranked distances are the in-order FP32 distances of the returned ids
    for i in range(0, len(Q), 37):
        for a, d in zip(red["ids"][i], red["dists"][i]):
            assert d == oracle.distance(X[a], Q[i], "ip")
    # and close to FP64
    ex = -(X[red["ids"]].astype(np.float64) * Q[:, None, :].astype(np.float64)).sum(-1)
    rel = np.abs(red["dists"] - ex) / np.maximum(np.abs(ex), 1e-30)
    assert rel.max() <= 1e-5
    # sorted by (dist, id)
    assert np.all(np.diff(red["dists"], axis=1) >= 0)


@pytest.mark.parametrize("dim", [20, 23, 96, 99, 128, 200, 201, 256, 300, 301])
@pytest.mark.parametrize("metric", ["l2", "ip"])
@pytest.mark.parametrize("store", ["bf16", "fp8"])
def test_rerank_distances_bit_exact_dims(oracle, store, metric, dim):
    """The fused FP32 rerank: the 8-lane in-order chain (BF16, FP8 inner
    product) at rows of 1 to 4 lines, of 5 to 8 lines and past 8 lines (8
    lines at a time), and one lane per row (FP8 L2); dimensions that are not
    multiples of 4 included.  Every returned distance is the oracle's FP32
    distance."""
    rng = np.random.default_rng(2000 + dim)
    X = rng.standard_normal((3000, dim)).astype(np.float32)
    Q = rng.standard_normal((100, dim)).astype(np.float32)
    _say(f"{store} + rerank {metric} dim={dim}")
    perm = datagen.insertion_orders(len(X), 1, 22)
    adj = oracle.build_vamana(X, metric, 16, 32, 1.2 if metric == "l2" else 1.0, perm)
    idx = DeviceIndex(X, adj, metric)
    idx.build_store(store)
    r = idx.search(Q, 10, 40, store=store, rerank=True)
    for i in range(len(Q)):
        for a, d in zip(r["ids"][i], r["dists"][i]):
            assert d == oracle.distance(X[a], Q[i], metric), (i, a)
    assert np.all(np.diff(r["dists"], axis=1) >= 0)
    idx.close()


def _e4m3_table():
    v = np.zeros(256, np.float64)
    for c in range(256):
        s = -1.0 if c & 0x80 else 1.0
        e, m = (c >> 3) & 0xF, c & 7
        if e == 0xF and m == 7:
            v[c] = np.nan
        elif e == 0:
            v[c] = s * m * 2.0 ** -9
        else:
            v[c] = s * (1 + m / 8.0) * 2.0 ** (e - 7)
    return v


def _fp16_query(q, s):
    """The inner-product FP8 kernels' query: q~_j = q_j * s_j in FP32, times
    2^e with max |q~| * 2^e in [2^14, 2^15), rounded to FP16 (RNE), as
    float64 values of q~ (the 2^e taken back out, exactly)."""
    qs = q.astype(np.float32) * s.astype(np.float32)
    mx = float(np.abs(qs).max())
    e = int(np.clip(15 - np.frexp(mx)[1], -120, 120)) if mx > 0 else 0
    h = (qs * np.float32(2.0 ** e)).astype(np.float16)
    return h.astype(np.float64) * 2.0 ** -e


def _complete(n):
    adj = np.zeros((n, n), np.uint32)
    for i in range(n):
        nb = [j for j in range(n) if j != i]
        adj[i, 0] = len(nb)
        adj[i, 1:] = nb
    return adj


def _e4m3_encode_bruteforce(v, tab):
    finite = np.where(np.isnan(tab), np.inf, tab)
    out = np.zeros(v.shape, np.int64)
    for idx, x in np.ndenumerate(v):
        dif = np.abs(finite - float(x))
        best = np.flatnonzero(dif == dif.min())
        if len(best) > 1:  # ties: even mantissa
            best = [b for b in best if (b & 1) == 0] or list(best)
        out[idx] = int(best[0])
    return out


def test_fp8_store_codec():
    """FP8 store: scales s_j = max|x_j| / 448; codes = nearest e4m3fn value of
    x/s (ties to even, saturating), decoded exactly.  Checked through the
    search itself: on a complete graph with beam = n every row's FP8 IP
    distance comes back, and must equal -sum_j q_j s_j dec(code_ij)
    computed from a brute-force encoder over the 256-code table, the query
    term q_j s_j in the kernel's FP16 form (_fp16_query)."""
    _say("fp8 codec")
    rng = np.random.default_rng(3)
    n, d = 60, 16
    X = (rng.standard_normal((n, d)) * 10.0 ** rng.uniform(-4, 1, (n, d))).astype(np.float32)
    Q = rng.standard_normal((8, d)).astype(np.float32)
    idx = DeviceIndex(X, _complete(n), "ip")
    idx.build_store("fp8")
    s = idx.fp8_scales()
    assert np.array_equal(s, (np.abs(X).max(0) / np.float32(448.0)).astype(np.float32))
    tab = _e4m3_table()
    codes = _e4m3_encode_bruteforce((X / s[None, :]).astype(np.float32), tab)
    assert (codes & 0x7F).min() < 8  # subnormal codes are exercised
    r = idx.search(Q, n, n, store="fp8")
    for i in range(len(Q)):
        h = _fp16_query(Q[i], s)
        exp = -(tab[codes[r["ids"][i]]] @ h)
        assert np.allclose(r["dists"][i], exp, rtol=1e-5, atol=1e-6 * np.abs(exp).max())
        # the FP16 rounding shows at this tolerance: the FP32 query term misses it
        full = -(tab[codes[r["ids"][i]]] @ (Q[i].astype(np.float64) * s.astype(np.float64)))
        assert not np.allclose(r["dists"][i], full, rtol=1e-6, atol=1e-7 * np.abs(full).max())


@pytest.mark.parametrize("d", [20, 200, 253])  # rows of 1, 2 and 3 128-byte chunks
def test_fp8_residual_store_codec(d):
    """FP8 residual store: the centroids are points floor(n / ncent) * i; a
    point's centroid is its nearest in L2; scales s_j = max|x_j - c_j| / 448;
    codes = nearest e4m3fn value of (x - c) / s.  Checked through the search:
    on a complete graph with beam = n every row's distance comes back and must
    equal -(q.c + sum_j q_j s_j dec(code_ij)), from a brute-force encoder, the
    term q_j s_j in the kernel's FP16 form (_fp16_query)."""
    _say(f"fp8 residual codec d={d}")
    rng = np.random.default_rng(5)
    n, nc = 64, 8
    centers = rng.standard_normal((6, d)) * 3.0
    X = (centers[rng.integers(0, 6, n)] + 0.1 * rng.standard_normal((n, d))
         * 10.0 ** rng.uniform(-3, 0, (n, d))).astype(np.float32)
    Q = rng.standard_normal((8, d)).astype(np.float32)
    idx = DeviceIndex(X, _complete(n), "ip")
    idx.build_store("fp8r", ncent=nc)
    p = idx.fp8_residual()
    step = n // nc
    cent = X[np.arange(nc) * step]
    assert np.array_equal(p["centroids"], cent)
    D = ((X[:, None, :].astype(np.float64) - cent[None].astype(np.float64)) ** 2).sum(-1)
    srt = np.sort(D, axis=1)
    clear = srt[:, 1] - srt[:, 0] > 1e-4 * srt[:, 1]  # no FP32 near-tie between two centroids
    assert clear.mean() > 0.9
    a = D.argmin(1)
    assert np.array_equal(p["assign"][clear], a[clear])
    a = p["assign"].astype(np.int64)
    R = (X - cent[a]).astype(np.float32)
    s = p["scales"]
    assert np.array_equal(s, (np.abs(R).max(0) / np.float32(448.0)).astype(np.float32))
    tab = _e4m3_table()
    codes = _e4m3_encode_bruteforce((R / s[None, :]).astype(np.float32), tab)
    r = idx.search(Q, n, n, store="fp8r")
    for i in range(len(Q)):
        ids = r["ids"][i].astype(np.int64)
        q = Q[i].astype(np.float64)
        exp = -(cent[a[ids]].astype(np.float64) @ q + tab[codes[ids]] @ _fp16_query(Q[i], s))
        assert np.allclose(r["dists"][i], exp, rtol=1e-5, atol=1e-5 * np.abs(exp).max())
    # the store is inner product only
    from parlayann_amd._lib import PannError
    with pytest.raises(PannError):
        DeviceIndex(X, _complete(n), "l2").build_store("fp8r")


def test_fp8_residual_small_index():
    """Automatic centroid count on a 10-point index (the minimum, 4), and the
    rerank over it: with beam = n every point comes back in exact order."""
    _say("fp8 residual small index")
    rng = np.random.default_rng(6)
    X = rng.standard_normal((10, 8)).astype(np.float32)
    Q = rng.standard_normal((3, 8)).astype(np.float32)
    idx = DeviceIndex(X, _complete(10), "ip")
    idx.build_store("fp8r")
    assert idx.fp8_residual()["centroids"].shape == (4, 8)
    r = idx.search(Q, 10, 10, store="fp8r", rerank=True)
    for i in range(len(Q)):
        exact = -(X.astype(np.float64) @ Q[i].astype(np.float64))
        assert np.array_equal(r["ids"][i], np.argsort(exact, kind="stable"))


def test_bf16_store_codec():
    import torch
    _say("bf16 codec")
    rng = np.random.default_rng(4)
    n, d = 60, 24
    X = (rng.standard_normal((n, d)) * 3).astype(np.float32)
    Q = rng.standard_normal((8, d)).astype(np.float32)
    idx = DeviceIndex(X, _complete(n), "l2")
    idx.build_store("bf16")
    Xb = torch.from_numpy(X).to(torch.bfloat16).to(torch.float64).numpy()  # RNE on the CPU
    r = idx.search(Q, n, n, store="bf16")
    for i in range(len(Q)):
        exp = ((Xb[r["ids"][i]] - Q[i].astype(np.float64)) ** 2).sum(1)
        assert np.allclose(r["dists"][i], exp, rtol=1e-5)


def test_errors(oracle, u8_case):
    from parlayann_amd._lib import PannError
    X, Q, adj = u8_case
    idx = DeviceIndex(X, adj, "l2")
    with pytest.raises(PannError):
        idx.search(Q[:4], 11, 10)  # k > beam (beamSearch.h:612-616)
    with pytest.raises(PannError):
        idx.search(Q[:4], 5, 10, starts=[])  # no start point
    with pytest.raises(PannError):
        idx.search(Q[:4], 5, 10, starts=[1, 1])  # duplicate starts
    with pytest.raises(PannError):
        idx.search(Q[:4], 5, 10, starts=[len(X)])  # out of range
    empty = np.zeros_like(adj)
    with pytest.raises(PannError):
        DeviceIndex(X, empty, "l2").search(Q[:4], 5, 10)  # short beam (:557-562)


# ------------------------------------------------------------- Vamana build
@pytest.fixture(scope="module")
def small_u8():
    return datagen.workload_u8_l2(n=4000, nq=100)


def test_build_mode_visited_and_prune(oracle, u8_case):
    import torch
    from parlayann_amd import _lib as L
    X, Q, adj = u8_case
    _say("build-mode search + prune")
    idx = DeviceIndex(X, adj, "l2")
    dev = torch.device("cuda")
    ids = np.arange(0, 20000, 257, dtype=np.int32)
    nb = len(ids)
    cap = 512
    tid = torch.from_numpy(ids).to(dev)
    vid = torch.zeros((nb, cap), dtype=torch.int32, device=dev)
    vd = torch.zeros((nb, cap), dtype=torch.float32, device=dev)
    vc = torch.zeros(nb, dtype=torch.int32, device=dev)
    L.check(L.lib().pann_search_visited(idx._h, tid.data_ptr(), nb, 128, 0.125, 0, cap,
                                        vid.data_ptr(), vd.data_ptr(), vc.data_ptr(), None))
    vid_h, vd_h, vc_h = vid.cpu().numpy().view(np.uint32), vd.cpu().numpy(), vc.cpu().numpy()
    R = 64
    out = torch.zeros((nb, R), dtype=torch.int32, device=dev)
    cnt = torch.zeros(nb, dtype=torch.int32, device=dev)
    L.check(L.lib().pann_robust_prune_batch(idx._h, tid.data_ptr(), nb, vid.data_ptr(),
                                            vd.data_ptr(), vc.data_ptr(), cap, 1.2, R, 1,
                                            out.data_ptr(), cnt.data_ptr(), None))
    out_h, cnt_h = out.cpu().numpy().view(np.uint32), cnt.cpu().numpy()
    for j, p in enumerate(ids):
        o = oracle.beam_search(adj, X, "l2", X[p], [0], 0, 128, self_id=int(p))
        assert vc_h[j] == o["num_visited"]
        assert np.array_equal(vid_h[j, :vc_h[j]], o["visited_ids"])
        assert np.array_equal(vd_h[j, :vc_h[j]], o["visited_dists"])
        pr, _ = oracle.robust_prune(adj, X, "l2", int(p), o["visited_ids"], o["visited_dists"],
                                    1.2, R)
        assert np.array_equal(out_h[j, :cnt_h[j]], pr)


def test_build_eval_counts(u8_case):
    """pann_build_eval_counts: the build-mode search evaluates at least every
    node it visits; robustPrune's sort-and-cut mode (add = 3) evaluates exactly
    the candidates and the point's current neighbours; reset clears both."""
    import torch
    from parlayann_amd import _lib as L
    X, Q, adj = u8_case
    idx = DeviceIndex(X, adj, "l2")
    dev = torch.device("cuda")
    ids = np.arange(0, 20000, 997, dtype=np.int32)
    nb, cap = len(ids), 512
    tid = torch.from_numpy(ids).to(dev)
    vid = torch.zeros((nb, cap), dtype=torch.int32, device=dev)
    vd = torch.zeros((nb, cap), dtype=torch.float32, device=dev)
    vc = torch.zeros(nb, dtype=torch.int32, device=dev)
    idx.build_eval_counts(reset=True)
    L.check(L.lib().pann_search_visited(idx._h, tid.data_ptr(), nb, 64, 0.125, 0, cap,
                                        vid.data_ptr(), vd.data_ptr(), vc.data_ptr(), None))
    ev = idx.build_eval_counts(reset=True)
    nvis = int(vc.sum())
    assert ev["search"] >= nvis > 0 and ev["prune"] == 0, ev
    R = int(adj.shape[1] - 1)
    out = torch.zeros((nb, R), dtype=torch.int32, device=dev)
    cnt = torch.zeros(nb, dtype=torch.int32, device=dev)
    L.check(L.lib().pann_robust_prune_batch(idx._h, tid.data_ptr(), nb, vid.data_ptr(), None,
                                            vc.data_ptr(), cap, 1.2, R, 3, out.data_ptr(),
                                            cnt.data_ptr(), None))
    ev = idx.build_eval_counts()
    assert ev == {"search": 0, "prune": nvis + int(adj[ids, 0].astype(np.int64).sum())}, ev
    assert idx.build_eval_counts(reset=True) == ev
    assert idx.build_eval_counts() == {"search": 0, "prune": 0}
    idx.close()


@pytest.mark.parametrize("kind", ["u8", "f32", "f32-ip-d99", "i8-d200"])
def test_gpu_vamana_build_matches_oracle(oracle, kind):
    """The whole build, bit-equal: search, robustPrune (its distances on the
    8-lane row gathers: the in-order FP32 sum with the fused tail at d = 99,
    rows of two lines for int8 at d = 200) and the reverse-edge rule."""
    from parlayann_amd.vamana import build_vamana
    _say(f"gpu vamana build {kind}")
    if kind == "u8":
        X, _ = datagen.workload_u8_l2(n=4000, nq=10)
        R, Lb, alpha, metric = 32, 64, 1.2, "l2"
    elif kind == "f32":
        X, _ = datagen.workload_f32_l2(n=4000, nq=10)
        R, Lb, alpha, metric = 32, 64, 1.2, "l2"
    elif kind == "f32-ip-d99":
        X = datagen.anisotropic_unit(3000, datagen.anisotropic_sigmas(99, 5), 6)
        R, Lb, alpha, metric = 32, 64, 1.0, "ip"
    else:
        X = (datagen.workload_u8_l2(n=3000, nq=10, d=200)[0].astype(np.int16) - 128).astype(np.int8)
        R, Lb, alpha, metric = 32, 64, 1.2, "l2"
    perm = datagen.insertion_orders(len(X), 2, 21)
    o = oracle.build_vamana(X, metric, R, Lb, alpha, perm, num_passes=2)
    g = build_vamana(X, metric, R, Lb, alpha, perm, num_passes=2)
    assert np.array_equal(g, o)
