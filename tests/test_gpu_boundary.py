"""The host-buffer boundary: beam_search's whole result (frontier, sorted
visited list, comparisons; beamSearch.h:33-228, :419-445) against the
oracle, steady-state calls that allocate no device memory, and the rerank
argument check (beamSearch.h:571-584 reads past its vector when
k * rerank_factor < k)."""
import sys

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

from parlayann_amd import DeviceIndex, datagen, device_alloc_count  # noqa: E402
from parlayann_amd._lib import PannError, PANN_ERR_ARG  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def case(oracle):
    X, Q = datagen.workload_f32_l2(n=5000, nq=64, d=40)
    perm = datagen.insertion_orders(len(X), 1, 31)
    adj = oracle.build_vamana(X, "l2", 24, 48, 1.2, perm)
    return X, Q, adj


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("B", [8, 40, 150])
def test_beam_search_frontier_and_visited_lists(oracle, case, exact, B):
    X, Q, adj = case
    _say(f"beam_search visited lists B={B} exact={exact}")
    idx = DeviceIndex(X, adj, "l2")
    r = idx.beam_search(Q, B, exact_counts=exact, vis_cap=1024)
    for i in range(len(Q)):
        o = oracle.beam_search(adj, X, "l2", Q[i], [0], B, B)
        assert not o["hash_grew"]
        fc = int(r["frontier_count"][i])
        assert fc == len(o["ids"])
        assert np.array_equal(r["frontier_ids"][i, :fc], o["ids"])
        assert np.array_equal(r["frontier_dists"][i, :fc], o["dists"])
        vc = int(r["visited"][i])
        assert vc == o["num_visited"]
        assert np.array_equal(r["visited_ids"][i, :vc], o["visited_ids"])
        assert np.array_equal(r["visited_dists"][i, :vc], o["visited_dists"])
        if exact:
            assert int(r["dist_cmps"][i]) == o["full_dist_cmps"]
        else:
            assert int(r["dist_cmps"][i]) >= o["full_dist_cmps"]
    idx.close()


def test_visited_list_capacity_is_an_error(case):
    X, Q, adj = case
    idx = DeviceIndex(X, adj, "l2")
    with pytest.raises(PannError):
        idx.beam_search(Q[:4], 64, vis_cap=8)  # a beam-64 search visits more than 8
    idx.close()


def test_repeated_host_searches_allocate_nothing(oracle, case):
    X, Q, adj = case
    _say("host searches: allocation count")
    idx = DeviceIndex(X, adj, "l2")
    idx.build_store("fp8")
    first = idx.search(Q, 10, 32)
    idx.search(Q, 10, 32, store="fp8", rerank=True)
    idx.beam_search(Q, 32, vis_cap=256)
    a0 = device_alloc_count()
    for _ in range(3):
        again = idx.search(Q, 10, 32)
        idx.search(Q[:7], 10, 32)  # smaller: same staging
        idx.search(Q, 10, 32, store="fp8", rerank=True)
        idx.beam_search(Q, 32, vis_cap=256)
        idx.beam_search(Q[:1], 16)
    assert device_alloc_count() == a0
    assert np.array_equal(again["ids"], first["ids"])
    o = oracle.search_batch(adj, X, "l2", Q, 10, 32)
    assert np.array_equal(first["ids"], o["ids"]) and np.array_equal(first["dists"], o["dists"])
    idx.close()


def test_rerank_factor_below_one_is_rejected(case):
    X, Q, adj = case
    idx = DeviceIndex(X, adj, "l2")
    idx.build_store("bf16")
    with pytest.raises(PannError) as e:
        idx.search(Q[:4], 10, 32, store="bf16", rerank=True, rerank_factor=0.5)
    assert e.value.status == PANN_ERR_ARG
    r = idx.search(Q[:4], 10, 32, store="bf16", rerank=True, rerank_factor=1.0)
    assert r["ids"].shape == (4, 10)
    idx.close()
