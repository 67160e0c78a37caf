"""The batched uneven search (per-query k and beam, one persistent launch)
captured into a hipGraph and replayed: every replay, with eager work and
synchronisation between replays, returns the eager result bit for bit."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


@pytest.mark.parametrize("store", ["native", "q8"])
def test_gpu_graph_replay_equals_eager(oracle, store):
    import torch
    from parlayann_amd import DeviceIndex, datagen
    print(f"[graph] store={store}", flush=True)
    X, Q = datagen.workload_f32_l2(n=5000, nq=700, d=32)
    adj = oracle.build_vamana(X, "l2", 24, 48, 1.2, datagen.insertion_orders(len(X), 1, 9))
    idx = DeviceIndex(X, adj, "l2")
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
        for vpr in (1, 2):
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
                                 "frontier_filter", "exact_counts", "hash_slots_log2")}
    assert shape["waves_per_cu"] in autotune.WAVES_PER_CU + (0,)
    assert shape["frontier_filter"] in autotune.FRONTIER_FILTER
    assert not shape["exact_counts"] or (shape["hash_slots_log2"], shape["waves_per_cu"]) \
        in autotune.EXACT_SHAPES
    r = idx.search(Q, 10, 48, **shape)
    for key in ("ids", "dists", "visited"):
        assert np.array_equal(r[key], o[key]), key
    again = autotune.tune(idx, tq, 10, 48, sample=128)
    assert again["cached"] and {k: again[k] for k in shape} == shape
    idx.close()
