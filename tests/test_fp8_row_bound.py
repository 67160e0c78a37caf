"""The per-row bound the FP8 inner-product rerank prunes with (no GPU).

The library keeps, per FP8 row x, a value b (launch_fp8_row_bound in
parlayann_amd/csrc/quantize.hip) such that |D8(q, x) - D32(q, x)| <= |q| b for
every query q:
  * D8: the search's FP8 inner-product distance.  q_j s_j is rounded to FP32,
    scaled by 2^e (e from the query's largest |q_j s_j|) and rounded to FP16;
    it is multiplied by the decoded e4m3 code and summed in FP32 in some order;
    the sum is scaled back by 2^-e and negated;
  * D32: the reference's float loop (beamSearch.h's rerank distance), which
    the oracle restates (oracle.distance).
An entry whose D8 - |q| b lies above the first k entries' D8 + |q| b cannot
reach the first k by D32, so the kernel leaves its FP32 row unread; a bound
that is too small would change results.  Here the kernel's D8 arithmetic is
emulated exactly (two summation orders) on rows and queries chosen to make
the terms of the bound large: rows with e4m3-subnormal codes, queries along a
row's quantization error, queries spanning many binades (FP16-subnormal query
terms).  The worst case comes within 2 % of the bound, and dropping its
quantization term or its FP16 query term fails.

For L2 the search's FP8 distance is the FP32 sum of (q_j - s_j x8_j)^2, and
the library keeps b = |x - x^| + 2^-16 |x^| per row; the kernel's bounds on
sqrt(D32) are lo = sqrt(d) (1 - 2^-15)(1 - 2 eta) - b (1 + 2^-20) and hi =
(sqrt(d) (1 + 2^-15) + b (1 + 2^-20))(1 + 2 eta), eta = (dim + 8) 2^-24, in
float.  The L2 tests emulate that sum (fused and unfused differences, two
summation orders) against the oracle's L2 loop, with queries on the line
through a row and its decoded row, where ||q - x^| - |q - x|| = |x - x^|.
"""
import numpy as np


def _e4m3_values():
    v = np.zeros(256)
    for c in range(256):
        s = -1.0 if c & 0x80 else 1.0
        e, m = (c >> 3) & 15, c & 7
        if e == 15 and m == 7:
            v[c] = np.nan
        elif e == 0:
            v[c] = s * m / 8 * 2.0 ** -6
        else:
            v[c] = s * (1 + m / 8) * 2.0 ** (e - 7)
    return v


def _encode(y, tab):
    """Nearest e4m3 code of each float32 in y, ties to even mantissa,
    saturating at 448 (the store's encoder)."""
    finite = np.where(np.isnan(tab), np.inf, tab)
    d = np.abs(finite[None, :] - y.reshape(-1, 1).astype(np.float64))
    best = d.min(1, keepdims=True)
    cand = d == best
    even = cand & ((np.arange(256) & 1) == 0)[None, :]
    pick = np.where(even.any(1), even.argmax(1), cand.argmax(1))
    return pick.reshape(y.shape)


def _d8(q, s, codes, tab, order):
    """The kernel's FP8 inner-product distance of query q to the coded rows."""
    v = (q.astype(np.float32) * s.astype(np.float32)).astype(np.float32)
    mx = float(np.abs(v).max())
    e = 0 if mx == 0 else int(np.clip(15 - np.frexp(mx)[1], -120, 120))
    h = (v.astype(np.float64) * 2.0 ** e).astype(np.float16).astype(np.float64)
    out = np.empty(len(codes), np.float32)
    for r, c in enumerate(codes):
        prod = (h * tab[c]).astype(np.float32)  # FP16 x e4m3 products are exact in FP32
        if order == "sequential":
            acc = np.float32(0)
            for t in prod:
                acc = np.float32(acc + t)
        else:  # pairwise, as the 8-lane group's reduction ends
            a = prod
            while len(a) > 1:
                if len(a) % 2:
                    a = np.append(a, np.float32(0))
                a = (a[0::2] + a[1::2]).astype(np.float32)
            acc = a[0]
        out[r] = np.float32(-(np.float32(acc) * np.float32(2.0 ** -e)))
    return out


def _bound(X, s, codes, tab, drop=None):
    """b per row, as the library computes it (double, raised 1e-6, to float)."""
    dim = X.shape[1]
    g = dim / 96.0 if dim > 96 else 1.0
    xh = s.astype(np.float64)[None, :] * tab[codes]
    r = np.linalg.norm(X.astype(np.float64) - xh, axis=1)
    terms = {"quant": r,
             "fp16": (4.884e-4 + 5.73e-6 * g) * np.linalg.norm(xh, axis=1),
             "fp32": 5.8e-6 * g * np.linalg.norm(X.astype(np.float64), axis=1),
             "sub": np.full(len(X), 8e-8 * g * float(s.max()))}
    v = sum(t for k, t in terms.items() if k != drop)
    return (v * (1.0 + 1e-6)).astype(np.float32)


def _case(dim, seed):
    rng = np.random.default_rng(seed)
    n = 96
    X = rng.standard_normal((n, dim)) * 10.0 ** rng.uniform(-3, 1, (n, dim))
    X[: n // 4] *= 1e-3  # rows of e4m3-subnormal codes
    X[n // 4: n // 4 + 2, :] = 0.0
    X = X.astype(np.float32)
    s = (np.abs(X).max(0) / np.float32(448.0)).astype(np.float32)
    s[s == 0] = np.float32(1.0)
    tab = _e4m3_values()
    codes = _encode((X / s[None, :]).astype(np.float32), tab)
    xh = s.astype(np.float64)[None, :] * tab[codes]
    Q = [rng.standard_normal(dim) for _ in range(6)]
    for r in rng.choice(n, 6, replace=False):  # along a row's quantization error
        err = X[r].astype(np.float64) - xh[r]
        if np.linalg.norm(err) > 0:
            Q.append(err / np.linalg.norm(err))
    Q.append(rng.standard_normal(dim) * 10.0 ** rng.uniform(-12, 2, dim))  # many binades
    Q.append(np.sign(rng.standard_normal(dim)) * 1e3)
    return X, s, tab, codes, [np.asarray(q, np.float32) for q in Q]


def _worst_ratio(oracle, X, s, tab, codes, Q, drop=None):
    b = _bound(X, s, codes, tab, drop)
    worst = 0.0
    for q in Q:
        qn = float(np.linalg.norm(q.astype(np.float64)))
        d32 = np.array([oracle.distance(x, q, "ip") for x in X], np.float64)
        for order in ("sequential", "pairwise"):
            d8 = _d8(q, s, codes, tab, order).astype(np.float64)
            lim = qn * b.astype(np.float64)
            ok = lim > 0
            worst = max(worst, float(np.max(np.abs(d8 - d32)[ok] / lim[ok])))
            assert np.all(np.abs(d8 - d32)[~ok] == 0)
    return worst


def test_fp8_row_bound_holds(oracle):
    for dim in (16, 96, 200):
        X, s, tab, codes, Q = _case(dim, 3 + dim)
        worst = _worst_ratio(oracle, X, s, tab, codes, Q)
        assert worst <= 1.0, (dim, worst)
        assert worst > 0.3, (dim, worst)  # the cases come close to the bound


def test_fp8_row_bound_needs_its_terms(oracle):
    """Without the quantization term, or without the FP16 query term, the
    emulated distances leave the bound."""
    X, s, tab, codes, Q = _case(96, 99)
    assert _worst_ratio(oracle, X, s, tab, codes, Q, drop="quant") > 1.0
    X, s, tab, codes, Q = _case(16, 19)
    assert _worst_ratio(oracle, X, s, tab, codes, Q, drop="fp16") > 1.0


def _d8_l2(q, s, codes, tab, order, fused):
    """The kernel's FP8 L2 distance: t_j = q_j - s_j x8_j in FP32 (one rounding
    when fused, two otherwise), t_j^2 summed by FMA."""
    out = np.empty(len(codes), np.float32)
    q32 = q.astype(np.float32)
    for r, c in enumerate(codes):
        p = tab[c]
        if fused:
            t = (q32.astype(np.float64) - s.astype(np.float64) * p).astype(np.float32)
        else:
            t = (q32 - (s * p.astype(np.float32)).astype(np.float32)).astype(np.float32)
        sq = t.astype(np.float64) ** 2
        if order == "sequential":
            acc = np.float32(0)
            for v in sq:
                acc = np.float32(np.float64(acc) + v)
        else:  # per-lane FMA chains of 16, then pairwise
            parts = []
            for i in range(0, len(sq), 16):
                acc = np.float32(0)
                for v in sq[i:i + 16]:
                    acc = np.float32(np.float64(acc) + v)
                parts.append(acc)
            a = np.array(parts, np.float32)
            while len(a) > 1:
                if len(a) % 2:
                    a = np.append(a, np.float32(0))
                a = (a[0::2] + a[1::2]).astype(np.float32)
            acc = a[0]
        out[r] = acc
    return out


def _bound_l2(X, s, codes, tab, drop=False):
    xh = s.astype(np.float64)[None, :] * tab[codes]
    r = np.linalg.norm(X.astype(np.float64) - xh, axis=1)
    v = (0.0 if drop else r) + 2.0 ** -16 * np.linalg.norm(xh, axis=1)
    return (v * (1.0 + 1e-6)).astype(np.float32)


def _l2_case(dim, seed):
    X, s, tab, codes, _ = _case(dim, seed)
    rng = np.random.default_rng(seed + 1)
    xh = s.astype(np.float64)[None, :] * tab[codes]
    Q = [X[r] + rng.standard_normal(dim).astype(np.float32) * 0.1 for r in range(3)]
    for r in rng.choice(len(X), 8, replace=False):  # on the line through x^ and x
        e = X[r].astype(np.float64) - xh[r]
        for t in (0.0, 1.0, 3.0):
            Q.append(X[r].astype(np.float64) + t * e)
    Q.append(np.zeros(dim))
    Q.append(rng.standard_normal(dim) * 10.0 ** rng.uniform(-3, 1, dim))
    return X, s, tab, codes, [np.asarray(q, np.float32) for q in Q]


def _l2_worst(oracle, X, s, tab, codes, Q, drop=False):
    """Largest |sqrt(d) - sqrt(D32)| / (b + 2^-16 sqrt(d)); asserts lo <= sqrt(D32) <= hi."""
    b = _bound_l2(X, s, codes, tab, drop).astype(np.float32)
    eta = np.float32((X.shape[1] + 8) * 2.0 ** -24)
    worst = 0.0
    for q in Q:
        d32 = np.array([oracle.distance(x, q, "l2") for x in X], np.float64)
        for order in ("sequential", "pairwise"):
            for fused in (False, True):
                d8 = _d8_l2(q, s, codes, tab, order, fused)
                u = np.sqrt(np.maximum(d8, 0)).astype(np.float32)
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


def test_fp8_l2_row_bound_holds(oracle):
    for dim in (16, 128, 200):
        X, s, tab, codes, Q = _l2_case(dim, 5 + dim)
        worst = _l2_worst(oracle, X, s, tab, codes, Q)
        assert worst <= 1.0, (dim, worst)
        assert worst > 0.5, (dim, worst)  # the collinear queries reach the bound


def test_fp8_l2_row_bound_needs_its_terms(oracle):
    """Without |x - x^| the FP8 distances leave the bound."""
    X, s, tab, codes, Q = _l2_case(128, 71)
    assert _l2_worst(oracle, X, s, tab, codes, Q, drop=True) > 1.0
