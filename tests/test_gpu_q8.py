"""The reference's 8-bit scalar quantization on the GPU (PANN_STORE_Q8):
parameters equal to the oracle's bit for bit, and the quantized search +
float rerank (beam_search_rerank's rerank branch, beamSearch.h:569-587) equal
to the oracle's search over the oracle's codes: ids, distances, visited and
distance-comparison counts."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_gpu_q8_search_equals_oracle(oracle, metric):
    from parlayann_amd import DeviceIndex, datagen
    from parlayann_amd.index import _normalize
    print(f"[q8] metric={metric}", flush=True)
    if metric == "l2":
        X, Q = datagen.workload_f32_l2(n=6000, nq=200, d=48)
    else:
        X, Q = datagen.workload_aniso_ip(n=6000, nq=200, d=48)
        X, Q = _normalize(X), _normalize(Q)
    adj = oracle.build_vamana(X, metric, 32, 64, 1.2 if metric == "l2" else 1.0,
                              datagen.insertion_orders(len(X), 1, 5))
    prm = oracle.q8_params(X, metric)
    Xq = oracle.q8_translate(prm, metric, X)
    Qq = oracle.q8_translate(prm, metric, Q)
    idx = DeviceIndex(X, adj, metric)
    idx.build_store("q8")
    g = idx.q8_params()
    if metric == "l2":
        assert g["slope"] == prm.slope and g["offset"] == prm.offset
    else:
        assert g["max_val"] == prm.max_val and g["scale"] == prm.scale
    for k, B in ((10, 32), (10, 64), (1, 16)):
        r = idx.search(Q, k, B, store="q8", rerank=True, exact_counts=True)
        o = oracle.search_batch(adj, X, metric, Q, k, B, Xq=Xq, Qq=Qq)
        assert np.array_equal(r["ids"], o["ids"]), (k, B)
        assert np.array_equal(r["dists"], o["dists"]), (k, B)
        assert np.array_equal(r["visited"], o["visited"]), (k, B)
        assert np.array_equal(r["dist_cmps"].astype(np.uint64), o["dist_cmps"]), (k, B)
    idx.close()
