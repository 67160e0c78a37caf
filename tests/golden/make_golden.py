"""Writes the committed golden vectors under tests/golden/.

Inputs are small seeded datasets; graphs and expected search outputs come from
the oracle (oracle/pann_oracle.c, a restatement of beamSearch.h:33-646 and
vamana/index.h:63-319 pinned by tests/test_oracle_pins.py).  The reference
itself cannot be built here (parlaylib is absent), so these fixtures pin the
restatement against regressions and give the device path a fixed target.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O  # noqa: E402


def case(name, X, Q, metric, R, L, alpha, beams, k=10):
    rng = np.random.default_rng(99)
    perm = rng.permutation(len(X)).astype(np.uint32)[None, :]
    adj = O.build_vamana(X, metric, R, L, alpha, perm, nthreads=1)
    out = {"X": X, "Q": Q, "adj": adj, "perm": perm, "R": R, "L": L, "alpha": alpha,
           "beams": np.array(beams)}
    for B in beams:
        r = O.search_batch(adj, X, metric, Q, k, B, nthreads=1)
        assert not r["hash_grew"].any()
        out[f"ids_{B}"] = r["ids"]
        out[f"dists_{B}"] = r["dists"]
        out[f"visited_{B}"] = r["visited"]
        out[f"cmps_{B}"] = r["dist_cmps"]
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


def main():
    rng = np.random.default_rng(2024)
    C = rng.uniform(0, 255, (20, 32))
    lab = rng.integers(0, 20, 1200)
    Xu = np.clip(np.rint(C[lab] + 20 * rng.standard_normal((1200, 32))), 0, 255).astype(np.uint8)
    Qu = np.clip(np.rint(C[rng.integers(0, 20, 40)] + 20 * rng.standard_normal((40, 32))), 0,
                 255).astype(np.uint8)
    case("u8_l2", Xu, Qu, "l2", 16, 32, 1.2, [10, 32])
    Cf = rng.uniform(-1, 1, (20, 32))
    Xf = (Cf[rng.integers(0, 20, 1200)] + 0.1 * rng.standard_normal((1200, 32))).astype(np.float32)
    Qf = (Cf[rng.integers(0, 20, 40)] + 0.1 * rng.standard_normal((40, 32))).astype(np.float32)
    case("f32_l2", Xf, Qf, "l2", 16, 32, 1.2, [10, 32])
    Xi = rng.standard_normal((1200, 32)).astype(np.float32)
    Xi /= np.linalg.norm(Xi, axis=1, keepdims=True)
    Qi = rng.standard_normal((40, 32)).astype(np.float32)
    Qi /= np.linalg.norm(Qi, axis=1, keepdims=True)
    case("f32_ip", Xi, Qi, "ip", 16, 32, 1.0, [10, 32])


if __name__ == "__main__":
    main()
