"""Small seeded HNSW inputs shared by the oracle and GPU HNSW tests."""
import numpy as np

from parlayann_amd import hnsw


def clustered(n, d, dtype="float32", metric="l2", seed=3):
    rng = np.random.default_rng(seed)
    if dtype == "uint8":
        C = rng.integers(0, 256, (20, d)).astype(np.float64)
        X = np.clip(np.rint(C[rng.integers(0, 20, n)] + 12 * rng.standard_normal((n, d))), 0, 255)
        return X.astype(np.uint8)
    C = rng.standard_normal((20, d)).astype(np.float32) * 3
    X = (C[rng.integers(0, 20, n)] + rng.standard_normal((n, d))).astype(np.float32)
    if metric == "ip":
        X /= np.linalg.norm(X, axis=1, keepdims=True)
    return X


def build_small(oracle, metric="l2", n=1500, d=12, m=8, efc=48, seed=3, nthreads=None,
                dtype="float32"):
    X = clustered(n, d, dtype, metric, seed)
    lev = hnsw.random_levels(n, 1.0 / np.log(m), seed)
    adj, ent = oracle.build_hnsw(X, metric, m, efc, 1.0, lev, nthreads=nthreads)
    return X, lev, adj, ent


def model_of(X, lev, adj, ent, m, efc):
    return hnsw.HNSWModel(X.shape[1], float(1.0 / np.log(m)), m, efc, 1.0, lev,
                          np.arange(len(X), dtype=np.uint32), adj, np.array([ent], np.uint32))
