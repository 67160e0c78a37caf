"""Committed golden vectors (tests/golden/*.npz, written by
tests/golden/make_golden.py): the oracle must keep reproducing them, and the
device path must match them bit-exactly (uint8 and in-order float32)."""
import os

import numpy as np
import pytest

from conftest import gpu_available

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [("u8_l2", "l2"), ("f32_l2", "l2"), ("f32_ip", "ip")]


def _load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz")))


@pytest.mark.parametrize("name,metric", CASES)
def test_oracle_reproduces_golden(oracle, name, metric):
    g = _load(name)
    adj = oracle.build_vamana(g["X"], metric, int(g["R"]), int(g["L"]), float(g["alpha"]),
                              g["perm"], nthreads=2)
    assert np.array_equal(adj, g["adj"])
    for B in g["beams"]:
        r = oracle.search_batch(g["adj"], g["X"], metric, g["Q"], 10, int(B), nthreads=2)
        assert np.array_equal(r["ids"], g[f"ids_{B}"])
        assert np.array_equal(r["dists"], g[f"dists_{B}"])
        assert np.array_equal(r["visited"], g[f"visited_{B}"])
        assert np.array_equal(r["dist_cmps"], g[f"cmps_{B}"])


@pytest.mark.gpu
@pytest.mark.parametrize("name,metric", CASES)
def test_device_matches_golden(name, metric):
    if not gpu_available():
        pytest.skip("no HIP device")
    from parlayann_amd import DeviceIndex
    g = _load(name)
    idx = DeviceIndex(g["X"], g["adj"], metric)
    for B in g["beams"]:
        r = idx.search(g["Q"], 10, int(B), summation="inorder", exact_counts=True)
        assert np.array_equal(r["ids"], g[f"ids_{B}"])
        assert np.array_equal(r["dists"], g[f"dists_{B}"])
        assert np.array_equal(r["visited"], g[f"visited_{B}"])
        assert np.array_equal(r["dist_cmps"].astype(np.uint64), g[f"cmps_{B}"])
    idx.close()
