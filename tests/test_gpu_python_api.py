"""The reference's Python surface (python/wrapper.py load_index,
graph_index.cpp:48-305) over libpann, on files written in the reference's
formats: batch, single and file searches equal to the oracle's qsearchAll
(in-order float sums, so bit-exact), numpy and device-tensor queries giving
the same results, recall with ties, and the quant path (the reference's
8-bit codes + float rerank)."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _files(tmp_path, X, Q, adj):
    from parlayann_amd import formats
    data, graph, qf = (str(tmp_path / n) for n in ("base.bin", "graph.bin", "query.bin"))
    formats.write_vectors(data, X)
    formats.write_graph(graph, adj)
    formats.write_vectors(qf, Q)
    return data, graph, qf


@pytest.mark.parametrize("metric,dtype", [("Euclidian", "float"), ("mips", "float"),
                                          ("Euclidian", "uint8")])
def test_load_index_searches_equal_oracle(oracle, tmp_path, metric, dtype):
    import torch
    from parlayann_amd import datagen, formats, load_index
    from parlayann_amd.index import _normalize
    from parlayann_amd.recall import recall_with_ties
    print(f"[python api] metric={metric} dtype={dtype}", flush=True)
    m = "l2" if metric == "Euclidian" else "ip"
    if dtype == "uint8":
        X, Q = datagen.workload_u8_l2(n=5000, nq=120, d=32, centers=40)
    else:
        X, Q = datagen.workload_f32_l2(n=5000, nq=120, d=32)
    Xs = _normalize(X) if (m == "ip" and dtype == "float") else X  # graph_index.cpp:92-95
    adj = oracle.build_vamana(Xs, m, 24, 48, 1.2, datagen.insertion_orders(len(X), 1, 8))
    data, graph, qf = _files(tmp_path, X, Q, adj)
    idx = load_index(metric, dtype, data, graph)
    o = oracle.search_batch(adj, Xs, m, Q, 10, 40)
    ids, dists = idx.batch_search(Q, 10, 40)
    assert np.array_equal(ids, o["ids"]) and np.array_equal(dists, o["dists"])
    # queries already in GPU memory: searched in place, same results
    tq = torch.from_numpy(Q).to(torch.device("cuda", 0))
    ti, td = idx.batch_search(tq, 10, 40)
    assert ti.is_cuda and np.array_equal(ti.cpu().numpy().view(np.uint32), ids)
    assert np.array_equal(td.cpu().numpy(), dists)
    assert np.array_equal(idx.single_search(Q[7], 10, 40), ids[7])
    assert np.array_equal(idx.single_search(tq[7], 10, 40).cpu().numpy().view(np.uint32),
                          ids[7])
    fi, fd = idx.batch_search_from_string(qf, 10, 40)
    assert np.array_equal(fi, ids) and np.array_equal(fd, dists)
    # recall with ties at the k-th groundtruth distance (graph_index.cpp:259-305)
    gi, gd = oracle.groundtruth_f64(Xs, m, Q, 10)
    gt = str(tmp_path / "gt.ibin")
    formats.write_ibin(gt, gi, gd.astype(np.float32))
    r = idx.check_recall(qf, gt, ids, 10)
    assert r == pytest.approx(recall_with_ties(Xs, Q, gi, ids, 10, m)) and r > 0.8
    if dtype == "float":  # quant: the reference's 8-bit codes, float rerank
        qi, _ = idx.batch_search(Q, 10, 40, quant=True)
        tqi, _ = idx.batch_search(tq, 10, 40, quant=True)
        assert np.array_equal(tqi.cpu().numpy().view(np.uint32), qi)
        rq = recall_with_ties(Xs, Q, gi, qi, 10, m)
        assert rq >= r - 0.05
