"""Oracle parity at the shapes of the benchmark workloads (smaller n):

* batched uneven search: unit-normalised anisotropic inner-product data
  (d = 96) on an oracle Vamana graph with R = 64, L = 128, alpha = 1.0; each
  query has its own k in {1, 10, 100} and beam in [max(16, k), 256]; the
  launch is captured into a hipGraph and replayed (qsearchAll,
  beamSearch.h:599-646).  In-order FP32: bit-exact ids, distances and visited
  counts against the oracle, query by query, on every replay;
* reduced-precision gathers + FP32 rerank on the same graph at beams up to
  256 (beam_search_rerank, beamSearch.h:528-597): the reference's 8-bit codes
  bit-exact against the oracle; BF16 and FP8 distances equal to the oracle's
  FP32 distances and recall@10 within 0.02 of the FP32 search;
* HNSW serving: a 200-dimensional inner-product mixture (mean-shifted
  queries), m = 32, efc = 200: the GPU build equals the oracle's bit for bit,
  searches at k in {10, 100} equal the oracle's, and a mixed batch served
  through two overlapping slots returns each request's result alone.
"""
import sys

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def aniso(oracle):
    from parlayann_amd import datagen
    X, Q = datagen.workload_aniso_ip(n=50000, nq=600)
    _say("[aniso] oracle Vamana build n=50000 d=96 R=64 L=128 alpha=1.0")
    adj = oracle.build_vamana(X, "ip", 64, 128, 1.0, datagen.insertion_orders(len(X), 1, 41))
    return X, Q, adj


def test_uneven_batch_in_replayed_graph_equals_oracle(oracle, aniso):
    import torch
    from parlayann_amd import DeviceIndex
    X, Q, adj = aniso
    rng = np.random.default_rng(17)
    ks = rng.choice([1, 10, 100], len(Q)).astype(np.int32)
    bs = np.array([rng.integers(max(16, k), 257) for k in ks], np.int32)
    _say(f"[uneven] nq={len(Q)} k in {{1,10,100}} beam in [16,256]")
    idx = DeviceIndex(X, adj, "ip")
    dev = torch.device("cuda", 0)
    tq = torch.from_numpy(Q).to(dev)
    tk, tb = torch.from_numpy(ks).to(dev), torch.from_numpy(bs).to(dev)
    oi = torch.zeros((len(Q), 100), dtype=torch.int32, device=dev)
    od = torch.zeros((len(Q), 100), dtype=torch.float32, device=dev)
    ov = torch.zeros(len(Q), dtype=torch.int32, device=dev)
    kw = dict(k_per_query=tk, beam_per_query=tb, summation="inorder", out_visited=ov)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):  # first use of the stream outside the capture
        idx.search_device(tq, 100, 256, oi, od, stream=s, async_=True, **kw)
    s.synchronize()
    idx.search_status(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        idx.search_device(tq, 100, 256, oi, od, stream=s, async_=True, **kw)
    # the oracle, one (k, beam) group at a time
    want_i = np.zeros((len(Q), 100), np.uint32)
    want_d = np.zeros((len(Q), 100), np.float32)
    want_v = np.zeros(len(Q), np.uint32)
    for k, b in sorted(set(zip(ks.tolist(), bs.tolist()))):
        sel = np.nonzero((ks == k) & (bs == b))[0]
        o = oracle.search_batch(adj, X, "ip", Q[sel], k, b)
        assert not o["hash_grew"].any()
        want_i[sel, :k], want_d[sel, :k] = o["ids"], o["dists"]
        want_v[sel] = o["visited"]
    for rep in range(3):
        oi.fill_(-1)
        od.fill_(0.0)
        ov.fill_(0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        idx.search_status(s)
        gi = oi.cpu().numpy().view(np.uint32)
        gd = od.cpu().numpy()
        for q in range(len(Q)):
            k = ks[q]
            assert np.array_equal(gi[q, :k], want_i[q, :k]), (rep, q, k, bs[q])
            assert np.array_equal(gd[q, :k], want_d[q, :k]), (rep, q, k, bs[q])
        assert np.array_equal(ov.cpu().numpy().view(np.uint32), want_v), rep
    idx.close()


@pytest.mark.parametrize("beam", [64, 128, 256])
def test_reduced_precision_rerank_at_full_beams(oracle, aniso, beam):
    from parlayann_amd import DeviceIndex
    from parlayann_amd.recall import recall_plain
    X, Q, adj = aniso
    _say(f"[rerank] beam={beam}")
    gt, _ = oracle.groundtruth_f64(X, "ip", Q, 10)
    idx = DeviceIndex(X, adj, "ip")
    f32 = idx.search(Q, 10, beam, summation="inorder")
    o = oracle.search_batch(adj, X, "ip", Q, 10, beam)
    assert np.array_equal(f32["ids"], o["ids"]) and np.array_equal(f32["dists"], o["dists"])
    rec32 = recall_plain(f32["ids"], gt, 10)
    for store in ("bf16", "fp8"):
        idx.build_store(store)
        r = idx.search(Q, 10, beam, store=store, rerank=True)
        for i in range(len(Q)):
            for a, d in zip(r["ids"][i], r["dists"][i]):
                assert d == np.float32(oracle.distance(X[a], Q[i], "ip")), (store, i, a)
        assert np.all(np.diff(r["dists"], axis=1) >= 0)
        assert recall_plain(r["ids"], gt, 10) >= rec32 - 0.02, store
    # the reference's own 8-bit codes: bit-exact with the oracle's Q8 search
    idx.build_store("q8")
    prm = oracle.q8_params(X, "ip")
    Xq, Qq = oracle.q8_translate(prm, "ip", X), oracle.q8_translate(prm, "ip", Q)
    oq = oracle.search_batch(adj, X, "ip", Q, 10, beam, Xq=Xq, Qq=Qq)
    q8 = idx.search(Q, 10, beam, store="q8", rerank=True)
    assert np.array_equal(q8["ids"], oq["ids"]) and np.array_equal(q8["dists"], oq["dists"])
    assert np.array_equal(q8["visited"], oq["visited"])
    idx.close()


@pytest.mark.parametrize("store", ["fp8", "q8", "bf16"])
@pytest.mark.parametrize("beam,rf,k", [(64, 100.0, 10), (256, 100.0, 10), (352, 100.0, 10),
                                       (256, 3.0, 10), (128, 100.0, 1), (256, 2.0, 100),
                                       (352, 1.5, 100)])
def test_pruned_rerank_equals_rescored_frontier(oracle, aniso, store, beam, rf, k):
    """FP8, BF16 or 8-bit-code search + FP32 rerank (beam_search_rerank,
    beamSearch.h:569-587): the result must be the first k by (FP32 distance,
    id) of the search frontier's first min(k * rerank_factor, |frontier|)
    entries, whatever rows the kernel's per-row error bounds let it leave
    unread.  The frontier comes from the same search at k = beam; the FP32
    distances from the oracle."""
    from parlayann_amd import DeviceIndex
    X, Q, adj = aniso
    _say(f"[rerank pruning] store={store} beam={beam} rerank_factor={rf} k={k}")
    idx = DeviceIndex(X, adj, "ip")
    idx.build_store(store)
    front = idx.search(Q, beam, beam, store=store)
    r = idx.search(Q, k, beam, store=store, rerank=True, rerank_factor=rf)
    nchk = min(int(k * rf), beam)
    for i in range(len(Q)):
        ids = front["ids"][i][:nchk].astype(np.int64)
        d = np.array([np.float32(oracle.distance(X[a], Q[i], "ip")) for a in ids])
        order = np.lexsort((ids, d))[:k]
        assert np.array_equal(r["ids"][i].astype(np.int64), ids[order]), (beam, k, i)
        assert np.array_equal(r["dists"][i], d[order]), (beam, k, i)
    idx.close()


@pytest.fixture(scope="module")
def f32_l2(oracle):
    from parlayann_amd import datagen
    X, Q = datagen.workload_f32_l2(n=20000, nq=300, d=128)
    _say("[f32 l2] oracle Vamana build n=20000 d=128 R=32 L=64 alpha=1.2")
    adj = oracle.build_vamana(X, "l2", 32, 64, 1.2, datagen.insertion_orders(len(X), 1, 43))
    return X, Q, adj


@pytest.mark.parametrize("store", ["fp8", "bf16"])
@pytest.mark.parametrize("beam,rf,k", [(64, 100.0, 10), (256, 100.0, 10), (1000, 100.0, 10),
                                       (256, 100.0, 1), (512, 2.0, 100)])
def test_pruned_l2_rerank_equals_rescored_frontier(oracle, f32_l2, store, beam, rf, k):
    """FP8 or BF16 L2 search + FP32 rerank (beamSearch.h:569-587): as for inner
    product, the first k by (FP32 distance, id) of the frontier's first
    min(k * rerank_factor, |frontier|) entries, whatever rows the per-row
    bounds leave unread."""
    from parlayann_amd import DeviceIndex
    X, Q, adj = f32_l2
    _say(f"[l2 rerank pruning] store={store} beam={beam} rerank_factor={rf} k={k}")
    idx = DeviceIndex(X, adj, "l2")
    idx.build_store(store)
    front = idx.search(Q, beam, beam, store=store)
    r = idx.search(Q, k, beam, store=store, rerank=True, rerank_factor=rf)
    nchk = min(int(k * rf), beam)
    for i in range(len(Q)):
        ids = front["ids"][i][:nchk].astype(np.int64)
        d = np.array([np.float32(oracle.distance(X[a], Q[i], "l2")) for a in ids])
        order = np.lexsort((ids, d))[:k]
        assert np.array_equal(r["ids"][i].astype(np.int64), ids[order]), (beam, k, i)
        assert np.array_equal(r["dists"][i], d[order]), (beam, k, i)
    idx.close()


@pytest.fixture(scope="module")
def hnsw_ip(oracle):
    from parlayann_amd import datagen, hnsw
    from parlayann_amd.index import _normalize
    X, Q = datagen.workload_hnsw_ip(n=20000, nq=300)
    X, Q = _normalize(X), _normalize(Q)
    lev = hnsw.random_levels(len(X), 1.0 / np.log(32), 5)
    _say("[hnsw ip] oracle build n=20000 d=200 m=32 efc=200")
    adj, ent = oracle.build_hnsw(X, "ip", 32, 200, 1.0, lev)
    return X, Q, lev, adj, ent


def test_hnsw_serving_build_equals_oracle(hnsw_ip):
    from parlayann_amd import hnsw
    X, Q, lev, adj, ent = hnsw_ip
    _say("[hnsw ip] GPU build")
    mdl = hnsw.build_hnsw(X, "ip", 32, 200, 1.0, levels=lev)
    assert int(mdl.entrance[0]) == ent
    assert len(mdl.graphs) == len(adj)
    for l, (g, o) in enumerate(zip(mdl.graphs, adj)):
        bad = np.nonzero((g != o).any(1))[0]
        assert bad.size == 0, f"level {l}: {bad.size} rows differ, first {bad[:5]}"


@pytest.mark.parametrize("k,ef", [(10, 32), (10, 100), (100, 100), (100, 200)])
def test_hnsw_serving_search_equals_oracle(oracle, hnsw_ip, k, ef):
    from parlayann_amd import hnsw
    X, Q, lev, adj, ent = hnsw_ip
    _say(f"[hnsw ip] search k={k} ef={ef}")
    mdl = hnsw.HNSWModel(X.shape[1], float(1.0 / np.log(32)), 32, 200, 1.0, lev,
                         np.arange(len(X), dtype=np.uint32), adj, np.array([ent], np.uint32))
    idx = hnsw.HNSWIndex(X, mdl, "ip")
    r = idx.search(Q, k, ef)
    oi, od, oc, om, ov = oracle.hnsw_search(adj, X, "ip", Q, k, ef, [ent])
    assert np.array_equal(r["count"], oc)
    for i in range(len(Q)):
        c = oc[i]
        assert np.array_equal(r["ids"][i, :c], oi[i, :c]), i
        assert np.array_equal(r["dists"][i, :c], od[i, :c]), i
    assert np.array_equal(r["visited"], ov)
    idx.close()


def test_hnsw_serving_mixed_requests_equal_alone(oracle, hnsw_ip):
    from parlayann_amd import hnsw
    from parlayann_amd.serve import BatchServer, poisson_arrivals
    X, Q, lev, adj, ent = hnsw_ip
    mdl = hnsw.HNSWModel(X.shape[1], float(1.0 / np.log(32)), 32, 200, 1.0, lev,
                         np.arange(len(X), dtype=np.uint32), adj, np.array([ent], np.uint32))
    idx = hnsw.HNSWIndex(X, mdl, "ip")
    rng = np.random.default_rng(8)
    ks = rng.choice([10, 100], len(Q)).astype(np.int32)
    efs = np.maximum(ks, 64).astype(np.int32)
    _say("[hnsw ip] mixed k in {10, 100} through the server")
    srv = BatchServer(idx, 128, 1 << 20, slots=2, summation="inorder")
    ids, cnt = srv.search_all(Q, ks, efs)
    for k in (10, 100):
        sel = np.nonzero(ks == k)[0]
        oi, od, oc, _, _ = oracle.hnsw_search(adj, X, "ip", Q[sel], k, max(k, 64), [ent])
        assert np.array_equal(cnt[sel], oc), k
        for j, q in enumerate(sel):
            assert np.array_equal(ids[q, :oc[j]], oi[j, :oc[j]]), (k, q)
    arr = poisson_arrivals(len(Q), 50000.0, seed=2)
    res, sids, scnt = srv.serve(Q, ks, efs, arr, keep_results=True)
    assert np.array_equal(scnt, cnt)
    for q in range(len(Q)):
        assert np.array_equal(sids[q, :cnt[q]], ids[q, :cnt[q]]), q
    idx.close()
