"""The per-row bounds the BF16 rerank prunes with (no GPU).

BF16 rows x^ are x rounded to nearest even (exact in FP32); the search sums
the FP32 query against them with FMA.  The library keeps per row
(launch_bf16_row_bound in parlayann_amd/csrc/quantize.hip):
  * inner product: b = |x - x^| + 5.8e-6 g (|x^| + |x|), g = max(1, dim/96),
    with |D16 - D32| <= |q| b;
  * L2: b = |x - x^| + 2^-16 |x^|, with the kernel's bounds on sqrt(D32) as
    for FP8 L2 (tests/test_fp8_row_bound.py).
D32 is the oracle's float loop (the reference's rerank distance).  The
queries lie along a row's rounding error (inner product) or on the line
through x^ and x (L2), where the bounds are reached; leaving out |x - x^|
fails.
"""
import numpy as np


def _bf16(X):
    u = np.ascontiguousarray(X, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _fma_sum(terms_a, terms_b, order):
    """sum of a_j * b_j by FMA in FP32 (sequential, or per-lane chains of 8 then pairwise)."""
    prod = terms_a.astype(np.float64) * terms_b.astype(np.float64)

    def chain(p):
        acc = np.float32(0)
        for v in p:
            acc = np.float32(np.float64(acc) + v)
        return acc
    if order == "sequential":
        return chain(prod)
    a = np.array([chain(prod[i:i + 8]) for i in range(0, len(prod), 8)], np.float32)
    while len(a) > 1:
        if len(a) % 2:
            a = np.append(a, np.float32(0))
        a = (a[0::2] + a[1::2]).astype(np.float32)
    return a[0]


def _rows(dim, seed):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((48, dim)) * 10.0 ** rng.uniform(-3, 2, (48, dim))).astype(np.float32)
    X[:2] = 0.0
    return X, _bf16(X), rng


def _bound(X, Xh, metric, drop=False):
    dim = X.shape[1]
    g = dim / 96.0 if dim > 96 else 1.0
    x, h = X.astype(np.float64), Xh.astype(np.float64)
    r = 0.0 if drop else np.linalg.norm(x - h, axis=1)
    if metric == "ip":
        v = r + 5.8e-6 * g * (np.linalg.norm(h, axis=1) + np.linalg.norm(x, axis=1))
    else:
        v = r + 2.0 ** -16 * np.linalg.norm(h, axis=1)
    return (v * (1.0 + 1e-6)).astype(np.float32)


def _ip_worst(oracle, dim, seed, drop=False):
    X, Xh, rng = _rows(dim, seed)
    b = _bound(X, Xh, "ip", drop).astype(np.float64)
    Q = [rng.standard_normal(dim) for _ in range(4)]
    for r in rng.choice(np.arange(2, len(X)), 6, replace=False):
        e = X[r].astype(np.float64) - Xh[r].astype(np.float64)
        Q.append(e / np.linalg.norm(e) * 3.0)
    Q.append(rng.standard_normal(dim) * 10.0 ** rng.uniform(-8, 3, dim))
    worst = 0.0
    for q in (np.asarray(v, np.float32) for v in Q):
        qa = float(np.float32(np.sqrt(np.float32(np.sum(q.astype(np.float32) ** 2)))) *
                   np.float32(1 + 2.0 ** -16))
        d32 = np.array([oracle.distance(x, q, "ip") for x in X], np.float64)
        for order in ("sequential", "pairwise"):
            d16 = np.array([-_fma_sum(q, h, order) for h in Xh], np.float64)
            lim = qa * b
            ok = lim > 0
            assert np.all(np.abs(d16 - d32)[~ok] == 0)
            worst = max(worst, float(np.max(np.abs(d16 - d32)[ok] / lim[ok])))
    return worst


def _l2_worst(oracle, dim, seed, drop=False):
    X, Xh, rng = _rows(dim, seed)
    b = _bound(X, Xh, "l2", drop)
    eta = np.float32((dim + 8) * 2.0 ** -24)
    Q = [X[r] + rng.standard_normal(dim).astype(np.float32) for r in range(2, 5)]
    for r in rng.choice(np.arange(2, len(X)), 6, replace=False):
        e = X[r].astype(np.float64) - Xh[r].astype(np.float64)
        for t in (0.0, 1.0, 3.0):
            Q.append(X[r].astype(np.float64) + t * e)
    worst = 0.0
    for q in (np.asarray(v, np.float32) for v in Q):
        d32 = np.array([oracle.distance(x, q, "l2") for x in X], np.float64)
        for order in ("sequential", "pairwise"):
            d16 = np.array([_fma_sum((q - h).astype(np.float32), (q - h).astype(np.float32),
                                     order) for h in Xh], np.float32)
            u = np.sqrt(d16).astype(np.float32)
            lo = (u * np.float32((1 - 2.0 ** -15) * (1 - 2 * eta))).astype(np.float32) \
                - (b * np.float32(1 + 2.0 ** -20)).astype(np.float32)
            hi = ((u * np.float32(1 + 2.0 ** -15)).astype(np.float32)
                  + (b * np.float32(1 + 2.0 ** -20)).astype(np.float32)) * np.float32(1 + 2 * eta)
            t = np.sqrt(d32)
            if not drop:
                assert np.all(lo.astype(np.float64) <= t) and np.all(t <= hi.astype(np.float64))
            lim = b.astype(np.float64) + 2.0 ** -16 * u
            ok = lim > 0
            worst = max(worst, float(np.max(np.abs(u.astype(np.float64) - t)[ok] / lim[ok])))
    return worst


def test_bf16_row_bounds_hold(oracle):
    for dim in (16, 96, 200):
        w = _ip_worst(oracle, dim, 3 + dim)
        assert 0.3 < w <= 1.0, ("ip", dim, w)
        w = _l2_worst(oracle, dim, 5 + dim)
        assert 0.5 < w <= 1.0, ("l2", dim, w)


def test_bf16_row_bounds_need_the_rounding_term(oracle):
    assert _ip_worst(oracle, 96, 31, drop=True) > 1.0
    assert _l2_worst(oracle, 96, 37, drop=True) > 1.0
