"""include/pann/beam_search.h: the reference's C++ search surface
(filtered_beam_search, beam_search, beam_search_rerank, qsearchAll,
QueryParams, stats, Graph, PointRange; beamSearch.h:33-646, types.h:233-268,
stats.h:121-155, graph.h:130-255, point_range.h:43-141) over the C ABI.  The
CPU test compiles and links a caller written against the reference's
signatures; the GPU test runs it on files in the reference's formats and
compares every output with the oracle bit for bit."""
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
    from parlayann_amd import datagen, formats
    assert os.path.exists(BIN), "tests/cpp/header_check is built by __graft_entry__.build()"
    X, Q = datagen.workload_f32_l2(n=3000, nq=40, d=32)
    R, k, B, start = 24, 10, 40, 0
    adj = oracle.build_vamana(X, "l2", R, 48, 1.2, datagen.insertion_orders(len(X), 1, 3))
    fx, fg, fq = str(tmp_path / "x.fbin"), str(tmp_path / "g.graph"), str(tmp_path / "q.fbin")
    formats.write_vectors(fx, X)
    formats.write_graph(fg, adj)
    formats.write_vectors(fq, Q)
    fout = str(tmp_path / "out.bin")
    print(f"[cpp header] n={len(X)} nq={len(Q)} k={k} beam={B}", flush=True)
    subprocess.run([BIN, fx, fg, fq, str(k), str(B), str(start), fout], check=True, timeout=120)
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
    fids, fd = take(np.uint32, fc), take(np.float32, fc)
    vc = int(take(np.uint32, 1)[0])
    vids, vd = take(np.uint32, vc), take(np.float32, vc)
    fcmps = int(take(np.uint64, 1)[0])
    rids, rd = take(np.uint32, k), take(np.float32, k)
    qids = take(np.uint32, nq * k).reshape(nq, k)
    qrids, qrd = take(np.uint32, k), take(np.float32, k)
    allocs = int(take(np.uint64, 1)[0])
    assert pos == len(b)

    # qsearchAll with the reference's seen set: ids, visited and dist_cmps
    o = oracle.search_batch(adj, X, "l2", Q, k, B)
    assert np.array_equal(ids, o["ids"])
    assert np.array_equal(vis, o["visited"])
    assert np.array_equal(cmps.astype(np.uint64), o["dist_cmps"])
    # beam_search: frontier, the sorted visited list and the comparisons
    ob = oracle.beam_search(adj, X, "l2", Q[0], [start], k, B)
    assert fc == len(ob["ids"])
    assert np.array_equal(fids, ob["ids"]) and np.array_equal(fd, ob["dists"])
    assert vc == ob["num_visited"]
    assert np.array_equal(vids, ob["visited_ids"]) and np.array_equal(vd, ob["visited_dists"])
    assert fcmps == int(ob["full_dist_cmps"])
    # beam_search_rerank on the float points
    assert np.array_equal(rids, o["ids"][0]) and np.array_equal(rd, o["dists"][0])
    # FP8 copy + FP32 rerank: recall against the float search, and reranked
    # distances equal to the oracle's FP32 distances
    agree = np.mean([len(set(qids[i]) & set(o["ids"][i])) / k for i in range(nq)])
    assert agree >= 0.9
    assert np.array_equal(qrids, qids[0])
    for a, d in zip(qrids, qrd):
        assert d == np.float32(oracle.distance(X[a], Q[0], "l2"))
    # a repeated host search of the same size allocates no device memory
    assert allocs == 0
