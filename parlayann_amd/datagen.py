"""Seeded synthetic workloads (no datasets can be fetched).

Each generator is deterministic in its seeds and chunked, so the same bytes
come out on any host regardless of memory.  Shapes and laws:

- ``gaussian_mixture_f32``: centers ~ U[-1,1]^d, x = center[c] + sigma*N(0,I),
  c ~ U{0..centers-1} (100k x 128, 100 centers, sigma 0.1 for the FP32 L2
  workload; 1M x 200, 500 centers for the HNSW serving workload);
- ``gaussian_mixture_u8``: centers ~ U[0,255]^d, x = rint(center + 20*N(0,I))
  clipped to [0,255] (1M x 128, 1000 centers);
- ``anisotropic_unit``: sigma_j = 10^U[-1,0], x ~ N(0, diag sigma^2), x/|x|
  (10M x 96 inner-product workload).
"""
from __future__ import annotations

import numpy as np

_CHUNK = 1 << 16


def _rng(seed: int, chunk: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64([seed, chunk]))


def mixture_centers(num_centers: int, d: int, seed: int, low=-1.0, high=1.0) -> np.ndarray:
    return _rng(seed).uniform(low, high, size=(num_centers, d))


def gaussian_mixture_f32(n: int, centers: np.ndarray, sigma: float, seed: int,
                         shift: np.ndarray | None = None) -> np.ndarray:
    """Rows center[c] + sigma*N(0, I) (+ shift), float32."""
    k, d = centers.shape
    out = np.empty((n, d), dtype=np.float32)
    for ci, s in enumerate(range(0, n, _CHUNK)):
        e = min(n, s + _CHUNK)
        r = _rng(seed, ci)
        c = r.integers(0, k, size=e - s)
        x = centers[c] + sigma * r.standard_normal((e - s, d))
        if shift is not None:
            x += shift
        out[s:e] = x
    return out


def gaussian_mixture_u8(n: int, centers: np.ndarray, sigma: float, seed: int) -> np.ndarray:
    k, d = centers.shape
    out = np.empty((n, d), dtype=np.uint8)
    for ci, s in enumerate(range(0, n, _CHUNK)):
        e = min(n, s + _CHUNK)
        r = _rng(seed, ci)
        c = r.integers(0, k, size=e - s)
        x = np.rint(centers[c] + sigma * r.standard_normal((e - s, d)))
        out[s:e] = np.clip(x, 0, 255).astype(np.uint8)
    return out


def anisotropic_sigmas(d: int, seed: int) -> np.ndarray:
    return 10.0 ** _rng(seed).uniform(-1.0, 0.0, size=d)


def anisotropic_unit(n: int, sigmas: np.ndarray, seed: int) -> np.ndarray:
    d = sigmas.shape[0]
    out = np.empty((n, d), dtype=np.float32)
    for ci, s in enumerate(range(0, n, _CHUNK)):
        e = min(n, s + _CHUNK)
        x = _rng(seed, ci).standard_normal((e - s, d)) * sigmas
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[s:e] = x
    return out


def random_unit_vector(d: int, seed: int) -> np.ndarray:
    v = _rng(seed).standard_normal(d)
    return v / np.linalg.norm(v)


def insertion_orders(n: int, num_passes: int, seed: int) -> np.ndarray:
    """Insertion permutations for Vamana passes (the reference draws
    parlay::random_permutation per batch_insert call, vamana/index.h:212)."""
    return np.stack([_rng(seed, p).permutation(n).astype(np.uint32) for p in range(num_passes)])


# ---------------------------------------------------------------------------
# the named workloads
# ---------------------------------------------------------------------------

def workload_f32_l2(n=100_000, nq=10_000, d=128, centers=100, sigma=0.1):
    """Plain FP32 L2 workload: 100k x 128 from a 100-center mixture."""
    C = mixture_centers(centers, d, 101)
    return gaussian_mixture_f32(n, C, sigma, 102), gaussian_mixture_f32(nq, C, sigma, 103)


def workload_u8_l2(n=1_000_000, nq=10_000, d=128, centers=1000, sigma=20.0):
    """Exact-integer workload: 1M x 128 uint8 from a 1000-center mixture."""
    C = mixture_centers(centers, d, 201, 0.0, 255.0)
    return gaussian_mixture_u8(n, C, sigma, 202), gaussian_mixture_u8(nq, C, sigma, 203)


def workload_aniso_ip(n=10_000_000, nq=100_000, d=96):
    """Batched inner-product workload: unit-normalised anisotropic Gaussian."""
    s = anisotropic_sigmas(d, 301)
    return anisotropic_unit(n, s, 302), anisotropic_unit(nq, s, 303)


def workload_hnsw_ip(n=1_000_000, nq=10_000, d=200, centers=500, sigma=0.1, shift_sigmas=0.5):
    """Serving workload: 500-center mixture, queries mean-shifted by 0.5 sigma."""
    C = mixture_centers(centers, d, 501)
    u = random_unit_vector(d, 504)
    return (gaussian_mixture_f32(n, C, sigma, 502),
            gaussian_mixture_f32(nq, C, sigma, 503, shift=shift_sigmas * sigma * u))
