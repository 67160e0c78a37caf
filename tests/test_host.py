"""Host-side logic: file formats, the C-ABI library's exports, the batch
schedule, recall definitions, data generators, the groundtruth tool."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------------ formats
def test_vector_roundtrip(tmp_path, rng):
    from parlayann_amd import formats
    for dt, ext in [(np.float32, ".fbin"), (np.uint8, ".u8bin"), (np.int8, ".i8bin")]:
        X = (rng.standard_normal((37, 9)) * 50).astype(dt)
        p = str(tmp_path / ("v" + ext))
        formats.write_vectors(p, X)
        raw = open(p, "rb").read()
        assert np.frombuffer(raw[:8], np.uint32).tolist() == [37, 9]  # point_range.h:86-90
        assert np.array_equal(formats.read_vectors(p), X)


def test_graph_roundtrip_and_bytes(tmp_path):
    from parlayann_amd import formats
    adj = np.zeros((4, 4), np.uint32)
    adj[0, :3] = [2, 1, 3]
    adj[1, :2] = [1, 0]
    adj[3, :4] = [3, 0, 1, 2]
    p = str(tmp_path / "g")
    formats.write_graph(p, adj)
    raw = np.fromfile(p, np.uint32)
    # n, max_deg, degrees, then concatenated lists (graph.h:210-237)
    assert raw.tolist() == [4, 3, 2, 1, 0, 3, 1, 3, 0, 0, 1, 2]
    assert np.array_equal(formats.read_graph(p), adj)


def test_ibin_roundtrip(tmp_path, rng):
    from parlayann_amd import formats
    ids = rng.integers(0, 1000, (5, 7)).astype(np.uint32)
    d = rng.standard_normal((5, 7)).astype(np.float32)
    p = str(tmp_path / "gt.ibin")
    formats.write_ibin(p, ids, d)
    i2, d2 = formats.read_ibin(p)
    assert np.array_equal(i2, ids) and np.array_equal(d2, d)
    assert np.fromfile(p, np.int32, 2).tolist() == [5, 7]


# -------------------------------------------------------------------- C ABI
def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "pann.h")).read()
    return sorted(set(re.findall(r"\b(pann_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from parlayann_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    decl = _declared_symbols()
    assert len(decl) >= 14
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == decl


def test_library_errors_without_device_are_status_codes():
    from parlayann_amd import _lib
    L = _lib.lib()
    st = L.pann_open(None, None)
    assert st == _lib.PANN_ERR_ARG
    assert b"null" in L.pann_last_error()


# ------------------------------------------------------------ batch schedule
def test_batch_schedule_closed_form():
    from parlayann_amd.vamana import batch_schedule
    s = batch_schedule(100_000, 100_000)
    # prefix doubling 1, 2, 4, ... while 2^inc <= 0.02 n = 2000, then 2000 each
    assert s[:3] == [(0, 1), (1, 3), (3, 7)]
    assert s[10] == (1023, 2047)
    assert s[11] == (2047, 4047)
    assert s[-1][1] == 100_000
    assert all(a[1] == b[0] for a, b in zip(s, s[1:]))
    # 0.02 n < 1: max batch falls back to n (index.h:208-209), doubling still applies
    assert batch_schedule(10, 10) == [(0, 1), (1, 3), (3, 7), (7, 10)]


def test_batch_schedule_matches_oracle_build(oracle):
    """The GPU driver's schedule is the one the oracle builds with
    (oracle_vamana_batches, which its batch_insert runs), batch for batch."""
    from parlayann_amd.vamana import batch_schedule
    for n in [1, 2, 7, 100, 3001]:
        assert batch_schedule(n, n) == oracle.vamana_batches(n)


# ------------------------------------------------------------------- recall
def test_recall_definitions():
    from parlayann_amd.recall import recall_plain, recall_with_ties
    X = np.array([[0.0], [1.0], [1.0], [3.0]], np.float32)
    Q = np.array([[0.0]], np.float32)
    gt = np.array([[0, 1, 2, 3]], np.uint32)
    # k=2: gt[:2] = {0,1}; id 2 ties with the 2nd gt distance and also counts
    assert recall_plain(gt, np.array([[0, 2]]), 2) == 0.5
    assert recall_with_ties(X, Q, gt, np.array([[0, 2]]), 2, "l2") == 1.0


def test_ref_distance_matches_oracle(oracle, rng):
    from parlayann_amd.recall import ref_distance
    X = rng.standard_normal((20, 128)).astype(np.float32)
    q = rng.standard_normal(128).astype(np.float32)
    d = ref_distance(X, np.arange(20), q, "l2")
    for i in range(20):
        assert d[i] == oracle.distance(X[i], q, "l2")
    d = ref_distance(X, np.arange(20), q, "ip")
    for i in range(20):
        assert d[i] == oracle.distance(X[i], q, "ip")


# --------------------------------------------------------------- generators
def test_generators_deterministic_and_shaped():
    from parlayann_amd import datagen
    a, qa = datagen.workload_f32_l2(n=1000, nq=10)
    b, qb = datagen.workload_f32_l2(n=1000, nq=10)
    assert np.array_equal(a, b) and a.shape == (1000, 128) and a.dtype == np.float32
    u, _ = datagen.workload_u8_l2(n=500, nq=5)
    assert u.dtype == np.uint8 and u.shape == (500, 128)
    x, _ = datagen.workload_aniso_ip(n=300, nq=3)
    assert np.allclose(np.linalg.norm(x, axis=1), 1.0, atol=1e-5)


def test_groundtruth_numpy_matches_oracle(oracle, rng):
    from parlayann_amd.groundtruth import groundtruth
    X = rng.standard_normal((3000, 16)).astype(np.float32)
    Q = rng.standard_normal((40, 16)).astype(np.float32)
    for metric in ("l2", "ip"):
        g, gd = groundtruth(X, Q, 10, metric, use_gpu=False)
        o, od = oracle.groundtruth_f64(X, metric, Q, 10)
        assert np.array_equal(g, o)
        assert np.allclose(gd, od, rtol=1e-12)
