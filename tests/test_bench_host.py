"""Host-side pieces of the measurement harness (bench.py), the launch-shape
cache and the serving helpers, on the CPU: the one-GPU guard, the committed
traffic lookup, the row-byte and relative-error formulas against closed
forms, the tuner's cache file, and Poisson arrivals / latency summaries."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_measured_traffic_reads_the_committed_profiles():
    import bench
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        rows = json.load(f)
    assert rows, "profiles/traffic.json holds the rocprofv3 traffic of the headline kernels"
    for r in rows:
        w = dict(bench.WORKLOADS[r["workload"]])
        assert bench.measured_traffic(r["workload"], w, r["leg"], r["beam"]) == \
            r["traffic_bytes_per_launch"]
        # scaled FETCH + WRITE, as the rows say
        tb = r["fetch_kib"] * 1024 * r["fetch_scale"] + r["write_kib"] * 1024 * r["write_scale"]
        assert tb == pytest.approx(r["traffic_bytes_per_launch"])
    w = dict(bench.WORKLOADS["aniso-ip-10M"])
    assert bench.measured_traffic("aniso-ip-10M", w, "fp8-rerank", 12345) is None


def test_row_bytes_per_store():
    import bench
    assert bench.row_bytes("fp32-tree", 96, "f32") == 384
    assert bench.row_bytes("fp32-inorder", 96, "f32") == 384
    assert bench.row_bytes("bf16-rerank", 96, "f32") == 192
    assert bench.row_bytes("fp8-rerank", 96, "f32") == 96
    assert bench.row_bytes("fp8r-rerank", 200, "f32") == 204  # + the centroid id
    assert bench.row_bytes("q8-rerank", 200, "f32") == 200
    assert bench.row_bytes("u8", 128, "u8") == 128


def test_max_rel_dist_error_closed_form():
    import bench
    X = np.array([[1.0, 0.0], [0.0, 2.0], [3.0, 4.0]], np.float32)
    Q = np.array([[0.0, 0.0]], np.float32)
    ids = np.array([[0, 1, 2]])
    exact = np.array([[1.0, 4.0, 25.0]], np.float32)  # squared L2 to the origin
    assert bench.max_rel_dist_error(X, Q, ids, exact, "l2") == 0.0
    off = exact.copy()
    off[0, 2] = 25.5  # relative error 0.02 on the last item
    assert bench.max_rel_dist_error(X, Q, ids, off, "l2") == pytest.approx(0.02)
    Qi = np.array([[1.0, 1.0]], np.float32)
    ip = np.array([[-1.0, -2.0, -7.0]], np.float32)  # -<q, x>
    assert bench.max_rel_dist_error(X, Qi, ids, ip, "ip") == 0.0


def test_tune_cache_roundtrip(tmp_path, monkeypatch):
    from parlayann_amd import autotune
    monkeypatch.setenv("PANN_TUNE_CACHE", str(tmp_path / "sub" / "tune.json"))
    assert autotune.cache_path() == str(tmp_path / "sub" / "tune.json")
    assert autotune._load() == {}
    autotune._store({"k": {"waves_per_cu": 12}})
    assert autotune._load() == {"k": {"waves_per_cu": 12}}
    # a damaged file reads as empty rather than failing a run
    with open(autotune.cache_path(), "w") as f:
        f.write("{not json")
    assert autotune._load() == {}


def test_poisson_arrivals_and_latency_summary():
    from parlayann_amd.serve import ServeResult, poisson_arrivals
    a = poisson_arrivals(200_000, 1e6, seed=3)
    assert np.all(np.diff(a) >= 0)
    assert a[-1] / len(a) == pytest.approx(1e-6, rel=0.02)  # mean gap = 1 / rate
    assert np.array_equal(a, poisson_arrivals(200_000, 1e6, seed=3))
    lat = np.arange(1, 101) * 1e-3  # 1 .. 100 ms
    s = ServeResult(lat, batches=10, mean_batch=10.0, wall_s=2.0, max_in_flight=2).summary()
    assert s["p50_ms"] == pytest.approx(50.5) and s["p99_ms"] == pytest.approx(99.01)
    assert s["qps"] == pytest.approx(50.0) and s["max_batches_in_flight"] == 2
