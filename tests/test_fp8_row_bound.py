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
