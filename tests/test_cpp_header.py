"""include/pann/beam_search.h: the reference's C++ search names (QueryParams,
stats, qsearchAll, beam_search_rerank, beam_search; beamSearch.h:419-646,
types.h:233-268, stats.h:121-155) over the C ABI.  The CPU test compiles
and links a caller; the GPU test runs it and compares every output with the
oracle bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, gpu_available

LIB = os.path.join(ROOT, "parlayann_amd", "libpann.so")
BIN = os.path.join(ROOT, "tests", "cpp", "header_check")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libpann.so not built")
def test_header_compiles_and_links(tmp_path):
    import __graft_entry__ as ge
    out = ge.build_header_check(str(tmp_path / "hc"))
    assert os.path.exists(out)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_gpu_cpp_header_matches_oracle(oracle, tmp_path):
    from parlayann_amd import datagen
    assert os.path.exists(BIN), "tests/cpp/header_check is built by __graft_entry__.build()"
    X, Q = datagen.workload_f32_l2(n=3000, nq=40, d=32)
    R, k, B, start = 24, 10, 40, 0
    adj = oracle.build_vamana(X, "l2", R, 48, 1.2, datagen.insertion_orders(len(X), 1, 3))
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        f.write(np.array([len(X), X.shape[1], R, len(Q), k, B, start], np.uint32).tobytes())
        f.write(np.ascontiguousarray(X).tobytes())
        f.write(np.ascontiguousarray(adj).tobytes())
        f.write(np.ascontiguousarray(Q).tobytes())
    print(f"[cpp header] n={len(X)} nq={len(Q)} k={k} beam={B}", flush=True)
    subprocess.run([BIN, fin, fout], check=True, timeout=120)
    b = open(fout, "rb").read()
    nq = len(Q)
    pos = 0

    def take(dt, cnt):
        nonlocal pos
        a = np.frombuffer(b, dtype=dt, count=cnt, offset=pos)
        pos += a.nbytes
        return a
    ids = take(np.uint32, nq * k).reshape(nq, k)
    vis = take(np.uint32, nq)
    cmps = take(np.uint32, nq)
    fc = int(take(np.uint32, 1)[0])
    fids = take(np.uint32, fc)
    fd = take(np.float32, fc)
    fcmps = int(take(np.uint64, 1)[0])
    rids = take(np.uint32, k)
    rd = take(np.float32, k)
    o = oracle.search_batch(adj, X, "l2", Q, k, B)
    assert np.array_equal(ids, o["ids"])
    assert np.array_equal(vis, o["visited"])
    assert np.array_equal(cmps.astype(np.uint64), o["dist_cmps"])
    ob = oracle.beam_search(adj, X, "l2", Q[0], [start], k, B)
    assert fc == len(ob["ids"])
    assert np.array_equal(fids, ob["ids"]) and np.array_equal(fd, ob["dists"])
    assert fcmps == int(ob["full_dist_cmps"])
    assert np.array_equal(rids, o["ids"][0]) and np.array_equal(rd, o["dists"][0])
