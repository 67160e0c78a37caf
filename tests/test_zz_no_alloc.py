"""pann_search_noalloc: the search with no device allocation (its launch
memory is a file-scope device array) returns exactly what pann_search
returns, on two calls in a row, the second with a smaller batch, for each
store it supports; the library's device-allocation count does not move; and
the limits are errors, not faults."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

from parlayann_amd import DeviceIndex, datagen, device_alloc_count  # noqa: E402
from parlayann_amd import _lib as L  # noqa: E402


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(idx, tq, k, beam, oi, od, ov, om, store, rerank, summation):
    qp = idx._qparams(k, beam, None, None, 100.0, 0.125)
    op = idx._opts(store, summation, rerank, 0, False, 0, False)
    st = np.zeros(1, np.uint32)
    return L.lib().pann_search_noalloc(idx._h, tq.data_ptr(), int(tq.shape[0]), C.byref(qp),
                                       None, None, st.ctypes.data, 1, C.byref(op),
                                       oi.data_ptr(), od.data_ptr(), int(oi.shape[1]),
                                       ov.data_ptr(), om.data_ptr(), None)


@pytest.fixture(scope="module")
def case(oracle):
    X, Q = datagen.workload_f32_l2(n=4000, nq=300, d=40)
    adj = oracle.build_vamana(X, "l2", 24, 48, 1.2, datagen.insertion_orders(len(X), 1, 51))
    return X, Q, adj


@pytest.mark.parametrize("store,rerank,summation", [("native", False, "inorder"),
                                                    ("native", False, "tree"),
                                                    ("q8", True, "inorder"),
                                                    ("fp8", True, "inorder"),
                                                    ("bf16", True, "inorder")])
def test_noalloc_equals_search(oracle, case, store, rerank, summation):
    import torch
    X, Q, adj = case
    idx = DeviceIndex(X, adj, "l2")
    if store != "native":
        idx.build_store(store)
    dev = torch.device("cuda", 0)
    k, beam = 10, 48
    for nq in (300, 77):  # two calls in a row, the second smaller
        _say(f"[noalloc] store={store} rerank={rerank} summation={summation} nq={nq}")
        tq = torch.from_numpy(np.ascontiguousarray(Q[:nq])).to(dev)
        outs = []
        for _ in range(2):
            outs.append((torch.full((nq, k), -1, dtype=torch.int32, device=dev),
                         torch.zeros((nq, k), dtype=torch.float32, device=dev),
                         torch.zeros(nq, dtype=torch.int32, device=dev),
                         torch.zeros(nq, dtype=torch.int32, device=dev)))
        (ri, rd, rv, rm), (ni, nd, nv, nm) = outs
        idx.search_device(tq, k, beam, ri, rd, out_visited=rv, out_cmps=rm, store=store,
                          rerank=rerank, summation=summation)
        torch.cuda.synchronize()
        a0 = device_alloc_count()
        L.check(_noalloc(idx, tq, k, beam, ni, nd, nv, nm, store, rerank, summation))
        assert device_alloc_count() == a0
        assert torch.equal(ni, ri) and torch.equal(nd, rd)
        assert torch.equal(nv, rv) and torch.equal(nm, rm)
        if store == "native" and summation == "inorder":
            o = oracle.search_batch(adj, X, "l2", Q[:nq], k, beam)
            assert np.array_equal(ni.cpu().numpy().view(np.uint32), o["ids"])
            assert np.array_equal(nd.cpu().numpy(), o["dists"])
    idx.close()


def test_noalloc_limits_are_errors(case):
    import torch
    X, Q, adj = case
    idx = DeviceIndex(X, adj, "l2")
    dev = torch.device("cuda", 0)
    tq = torch.from_numpy(np.ascontiguousarray(Q[:8])).to(dev)
    oi = torch.zeros((8, 10), dtype=torch.int32, device=dev)
    od = torch.zeros((8, 10), dtype=torch.float32, device=dev)
    ov = torch.zeros(8, dtype=torch.int32, device=dev)
    om = torch.zeros(8, dtype=torch.int32, device=dev)
    qp = idx._qparams(10, 32, None, None, 100.0, 0.125)
    st = np.zeros(1, np.uint32)
    for kw, want in ((dict(async_=True), L.PANN_ERR_ARG),
                     (dict(exact_counts=True), L.PANN_ERR_UNSUPPORTED)):
        op = idx._opts("native", "inorder", False, 0, kw.get("async_", False), 0,
                       kw.get("exact_counts", False))
        rc = L.lib().pann_search_noalloc(idx._h, tq.data_ptr(), 8, C.byref(qp), None, None,
                                         st.ctypes.data, 1, C.byref(op), oi.data_ptr(),
                                         od.data_ptr(), 10, ov.data_ptr(), om.data_ptr(), None)
        assert rc == want, kw
    # Q8 codes past the array: refused on the host before any launch
    idx.build_store("q8")
    big = 8 * 1024 * 1024 // X.shape[1] + 1
    op = idx._opts("q8", "inorder", True, 0, False, 0, False)
    rc = L.lib().pann_search_noalloc(idx._h, tq.data_ptr(), big, C.byref(qp), None, None,
                                     st.ctypes.data, 1, C.byref(op), oi.data_ptr(),
                                     od.data_ptr(), 10, ov.data_ptr(), om.data_ptr(), None)
    assert rc == L.PANN_ERR_CAPACITY
    idx.close()
