"""The batched uneven search (per-query k and beam, one persistent launch)
captured into a hipGraph and replayed: every replay, with eager work and
synchronisation between replays, returns the eager result bit for bit."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


@pytest.mark.parametrize("store,metric", [("native", "l2"), ("q8", "l2"), ("fp8r", "ip")])
def test_gpu_graph_replay_equals_eager(oracle, store, metric):
    """fp8r: the per-query centroid table is the library's own MFMA kernel, so
    that store's search is captured too."""
    import torch
    from parlayann_amd import DeviceIndex, datagen
    print(f"[graph] store={store} metric={metric}", flush=True)
    if metric == "l2":
        X, Q = datagen.workload_f32_l2(n=5000, nq=700, d=32)
    else:
        X, Q = datagen.workload_aniso_ip(n=5000, nq=700, d=40)
    adj = oracle.build_vamana(X, metric, 24, 48, 1.2, datagen.insertion_orders(len(X), 1, 9))
    idx = DeviceIndex(X, adj, metric)
    if store != "native":
        idx.build_store(store)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(3)
    ks = rng.choice([1, 10, 50], len(Q)).astype(np.int32)
    bs = np.array([rng.integers(max(16, k), 97) for k in ks], np.int32)
    tq = torch.from_numpy(Q).to(dev)
    tk, tb = torch.from_numpy(ks).to(dev), torch.from_numpy(bs).to(dev)
    oi = torch.zeros((len(Q), 50), dtype=torch.int32, device=dev)
    od = torch.zeros((len(Q), 50), dtype=torch.float32, device=dev)
    kw = dict(k_per_query=tk, beam_per_query=tb, store=store, rerank=store != "native",
              summation="inorder")
    # columns past a query's k are not written: fill them the way every
    # replay below does
    oi.fill_(-1)
    od.fill_(0.0)
    idx.search_device(tq, 50, 96, oi, od, **kw)
    ref_i, ref_d = oi.clone(), od.clone()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):  # first use of the stream happens outside the capture
        idx.search_device(tq, 50, 96, oi, od, stream=s, async_=True, **kw)
    s.synchronize()
    idx.search_status(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        idx.search_device(tq, 50, 96, oi, od, stream=s, async_=True, **kw)
    for rep in range(4):
        oi.fill_(-1)  # eager work between replays
        od.fill_(0.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        idx.search_status(s)
        assert torch.equal(oi, ref_i), rep
        assert torch.equal(od, ref_d), rep
    idx.close()


def test_gpu_results_independent_of_launch_shape(oracle):
    """The autotuner may pick any (waves per CU, visits per round, evaluated-id
    cache size, frontier-id filter): every choice returns the oracle's ids,
    distances and visit counts."""
    from parlayann_amd import DeviceIndex, autotune, datagen
    X, Q = datagen.workload_u8_l2(n=8000, nq=300, d=32)
    adj = oracle.build_vamana(X, "l2", 24, 48, 1.2, datagen.insertion_orders(len(X), 1, 4))
    o = oracle.search_batch(adj, X, "l2", Q, 10, 48)
    idx = DeviceIndex(X, adj, "l2")
    for wpc in autotune.WAVES_PER_CU:
        for vpr in autotune.VISITS_PER_ROUND:
            for ec in autotune.EVAL_CACHE_LOG2 + (6, 13):
                for ff in autotune.FRONTIER_FILTER:
                    r = idx.search(Q, 10, 48, waves_per_cu=wpc, visits_per_round=vpr,
                                   eval_cache_log2=ec, frontier_filter=ff)
                    shape = (wpc, vpr, ec, ff)
                    assert np.array_equal(r["ids"], o["ids"]), shape
                    assert np.array_equal(r["dists"], o["dists"]), shape
                    assert np.array_equal(r["visited"], o["visited"]), shape
                    # cache and filter only remove evaluations, never below the
                    # reference's count
                    assert np.all(r["dist_cmps"].astype(np.uint64) >= o["dist_cmps"]), shape
    idx.close()


@pytest.mark.parametrize("beam", [48, 512, 600, 1500])
def test_gpu_frontier_filter_exact_f32(oracle, beam):
    """FP32 rows of 200 bytes (the filter's automatic choice is on) with the
    in-order summation: bit-exact against the oracle with the filter on, off
    and automatic, at beams inside one 512-id chunk of the exact check and
    across several."""
    from parlayann_amd import DeviceIndex, datagen
    X, Q = datagen.workload_f32_l2(n=6000, nq=120, d=50)
    adj = oracle.build_vamana(X, "l2", 24, 48, 1.2, datagen.insertion_orders(len(X), 1, 9))
    o = oracle.search_batch(adj, X, "l2", Q, 10, beam)
    idx = DeviceIndex(X, adj, "l2")
    for ff in (0, 1, 2):
        for vpr in (1, 2, 3):
            r = idx.search(Q, 10, beam, summation="inorder", frontier_filter=ff,
                           visits_per_round=vpr)
            for key in ("ids", "dists", "visited"):
                assert np.array_equal(r[key], o[key]), (key, ff, vpr)
            assert np.all(r["dist_cmps"].astype(np.uint64) >= o["dist_cmps"]), (ff, vpr)
    idx.close()


@pytest.mark.parametrize("metric,dtype,k", [("l2", "f32", 10), ("ip", "f32", 100),
                                            ("l2", "u8", 10), ("l2", "f32", 100),
                                            ("ip", "f32-200", 10), ("l2", "i8", 100),
                                            ("ip", "u8-200", 10), ("l2", "f32-128", 10),
                                            ("l2", "u8-128", 100)])
def test_gpu_bruteforce_groundtruth_equals_fp64(oracle, metric, dtype, k):
    """K7 (brute-force candidates on MFMA) + FP64 re-score gives exactly the
    oracle's FP64 brute-force top k in (distance, id) order.  The dims cover
    every register layout of the kernel: f32 dim 50 / 96 / 128 / 200, u8 / i8
    dim 64 / 100 / 128 / 200 (one or two query tiles per wave, one or two
    waves per SIMD)."""
    from parlayann_amd import DeviceIndex, datagen
    from parlayann_amd.groundtruth import groundtruth_device
    from parlayann_amd.index import _normalize
    print(f"[bruteforce] metric={metric} dtype={dtype} k={k}", flush=True)
    if dtype == "u8":
        X, Q = datagen.workload_u8_l2(n=30000, nq=700, d=64)
    elif dtype == "u8-128":
        X, Q = datagen.workload_u8_l2(n=24000, nq=400, d=128)
    elif dtype == "f32-128":
        X, Q = datagen.workload_f32_l2(n=30007, nq=401, d=128)
    elif dtype == "u8-200":
        X, Q = datagen.workload_u8_l2(n=20011, nq=333, d=200)
    elif dtype == "i8":
        X, Q = datagen.workload_u8_l2(n=25000, nq=500, d=100)
        X = (X.astype(np.int16) - 128).astype(np.int8)
        Q = (Q.astype(np.int16) - 128).astype(np.int8)
    elif dtype == "f32-200":
        X, Q = datagen.workload_hnsw_ip(n=20000, nq=300, d=200)
    elif metric == "ip":
        X, Q = datagen.workload_aniso_ip(n=30000, nq=700, d=96)
        X, Q = _normalize(X), _normalize(Q)
    else:
        X, Q = datagen.workload_f32_l2(n=30007, nq=701, d=50)
    adj = np.zeros((len(X), 2), np.uint32)
    idx = DeviceIndex(X, adj, metric)
    ids, d = groundtruth_device(idx, X, Q, k, metric)
    oi, od = oracle.groundtruth_f64(X, metric, Q, k)
    assert np.array_equal(ids, oi)
    # both FP64; the host re-score sums pairwise, the oracle in order
    assert np.allclose(d, od, rtol=1e-12, atol=1e-12)
    idx.close()


def test_gpu_autotune_returns_a_valid_shape(oracle, tmp_path, monkeypatch):
    """tune() grids the launch shapes on a sample, confirms the best few on
    the whole batch, caches the winner; the chosen shape reproduces the
    oracle's results and a second call is served from the cache."""
    import torch
    from parlayann_amd import DeviceIndex, autotune, datagen
    monkeypatch.setenv("PANN_TUNE_CACHE", str(tmp_path / "tune.json"))
    X, Q = datagen.workload_u8_l2(n=6000, nq=600, d=32)
    adj = oracle.build_vamana(X, "l2", 24, 48, 1.2, datagen.insertion_orders(len(X), 1, 5))
    o = oracle.search_batch(adj, X, "l2", Q, 10, 48)
    idx = DeviceIndex(X, adj, "l2")
    tq = torch.from_numpy(Q).to(torch.device("cuda"))
    res = autotune.tune(idx, tq, 10, 48, sample=128)
    assert not res["cached"] and res["queries"] == len(Q) and res["sample"] == 128
    shape = {k: res[k] for k in ("waves_per_cu", "visits_per_round", "eval_cache_log2",
                                 "frontier_filter", "exact_counts", "hash_slots_log2",
                                 "waves_per_query")}
    assert shape["waves_per_cu"] in autotune.WAVES_PER_CU + (0,) + tuple(
        w for _, w in autotune.COOP_SHAPES)
    assert shape["waves_per_query"] in (1,) + tuple(q for q, _ in autotune.COOP_SHAPES)
    assert shape["frontier_filter"] in autotune.FRONTIER_FILTER
    assert not shape["exact_counts"] or (shape["hash_slots_log2"], shape["waves_per_cu"]) \
        in autotune.EXACT_SHAPES
    r = idx.search(Q, 10, 48, **shape)
    for key in ("ids", "dists", "visited"):
        assert np.array_equal(r[key], o[key]), key
    again = autotune.tune(idx, tq, 10, 48, sample=128)
    assert again["cached"] and {k: again[k] for k in shape} == shape
    idx.close()


@pytest.mark.parametrize("store,metric,dim,beam", [("native", "l2", 40, 48), ("native", "l2", 40, 700),
                                                   ("fp8", "ip", 96, 300), ("bf16", "l2", 150, 90),
                                                   ("q8", "ip", 40, 200)])
def test_gpu_waves_per_query_equal_one_wave(oracle, store, metric, dim, beam):
    """Several waves per query (a workgroup shares each visit's gathers, the
    merges and the rerank) return exactly what one wave returns: ids,
    distances, visit counts and evaluations, for the in-order FP32 sum (also
    bit-exact against the oracle) and the reduced-precision stores with rerank,
    at beams that fill several merge rounds, and with per-query k and beam."""
    import torch
    from parlayann_amd import DeviceIndex, datagen
    rng = np.random.default_rng(dim + beam)
    X = rng.standard_normal((6000, dim)).astype(np.float32)
    if metric == "ip":
        X /= np.linalg.norm(X, axis=1, keepdims=True)
    Q = X[rng.integers(0, len(X), 150)] + 0.05 * rng.standard_normal((150, dim)).astype(np.float32)
    Q = Q.astype(np.float32)
    print(f"[waves/query] store={store} metric={metric} dim={dim} beam={beam}", flush=True)
    adj = oracle.build_vamana(X, metric, 24, 48, 1.2 if metric == "l2" else 1.0,
                              datagen.insertion_orders(len(X), 1, 17))
    idx = DeviceIndex(X, adj, metric)
    rr = store != "native"
    if rr:
        idx.build_store(store)
    # the cooperative launch runs without the frontier-id filter: so does the
    # one-wave reference here, whose evaluation count then matches too
    kw = dict(store=store, rerank=rr, summation="inorder", frontier_filter=2)
    one = idx.search(Q, 10, beam, **kw)
    if store == "native":
        o = oracle.search_batch(adj, X, metric, Q, 10, beam)
        for key in ("ids", "dists", "visited"):
            assert np.array_equal(one[key], o[key]), key
    for wpq in (2, 4):
        for wpc in (0, 8):
            r = idx.search(Q, 10, beam, waves_per_query=wpq, waves_per_cu=wpc, **kw)
            for key in ("ids", "dists", "visited", "dist_cmps"):
                assert np.array_equal(r[key], one[key]), (key, wpq, wpc)
    # per-query k and beam in one cooperative launch
    dev = torch.device("cuda", 0)
    kmax = min(50, beam)
    ks = rng.choice([1, 10, kmax], len(Q)).astype(np.int32)
    bs = np.array([rng.integers(max(16, k), beam + 1) for k in ks], np.int32)
    tq = torch.from_numpy(Q).to(dev)
    tk, tb = torch.from_numpy(ks).to(dev), torch.from_numpy(bs).to(dev)
    outs = []
    for wpq in (1, 2, 4):
        oi = torch.full((len(Q), kmax), -1, dtype=torch.int32, device=dev)
        od = torch.zeros((len(Q), kmax), dtype=torch.float32, device=dev)
        idx.search_device(tq, kmax, beam, oi, od, k_per_query=tk, beam_per_query=tb,
                          waves_per_query=wpq, **kw)
        outs.append((oi.cpu().numpy(), od.cpu().numpy()))
    for oi, od in outs[1:]:
        assert np.array_equal(oi, outs[0][0]) and np.array_equal(od, outs[0][1])
    idx.close()


def test_gpu_waves_per_query_hnsw_and_limits(oracle):
    """The HNSW search (descent and layer 0 in one launch) with several waves
    per query equals one wave; the options it does not take are errors."""
    from parlayann_amd import hnsw
    from parlayann_amd._lib import PannError
    from tests.hnsw_data import build_small, model_of
    X, lev, adj, ent = build_small(oracle, "l2", m=8, efc=48)
    idx = hnsw.HNSWIndex(X, model_of(X, lev, adj, ent, 8, 48), "l2")
    rng = np.random.default_rng(3)
    Q = (X[rng.integers(0, len(X), 200)] + 0.2 * rng.standard_normal((200, X.shape[1]))).astype(
        np.float32)
    import torch
    dev = torch.device("cuda", 0)
    tq = torch.from_numpy(Q).to(dev)
    res = []
    for wpq in (1, 2, 4):
        oi = torch.zeros((len(Q), 64), dtype=torch.int32, device=dev)
        od = torch.zeros((len(Q), 64), dtype=torch.float32, device=dev)
        oc = torch.zeros(len(Q), dtype=torch.int32, device=dev)
        idx.search_device(tq, 10, 64, oi, od, oc, waves_per_query=wpq)
        res.append((oi.cpu().numpy(), od.cpu().numpy(), oc.cpu().numpy()))
    for r in res[1:]:
        for a, b in zip(r, res[0]):
            assert np.array_equal(a, b)
    with pytest.raises(PannError):
        idx.search_device(tq, 10, 64, oi, od, oc, waves_per_query=2, exact_counts=True)
    with pytest.raises(PannError):
        idx.search_device(tq, 10, 64, oi, od, oc, waves_per_query=3)
    idx.close()
