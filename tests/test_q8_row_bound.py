"""The per-row bound the rerank of the reference's 8-bit inner-product codes
prunes with (no GPU).

For the Quantized_Mips_Point<8, true> codes (scale s, codes clipped to
[-127, 127]) the search's distance is d = -<code(q), code(x)>, an exact
integer, and the rerank's is the reference's float loop D32 (beamSearch.h's
rerank distance, restated by oracle.distance).  The library keeps per row
(launch_q8_row_bound in parlayann_amd/csrc/quantize.hip)
  R = s (|x - code(x)/s| + 5.8e-6 g |x|),  N = |code(x)|,  g = max(1, dim/96),
and per query the kernel forms
  A = |s q| (1 + 2^-16),  B = (|s q - code(q)| + 2^-20 A) (1 + 2^-16)
in float, so that |d - s^2 D32| <= A R + B N for every row; an entry whose
d minus that bound lies above the first k's largest d plus bound cannot
reach the first k by D32, and its FP32 row is left unread.  A bound that is
too small would change results.  Here both distances come from the oracle
(the codes from the oracle's own translation) on rows and queries chosen to
make each term large: a query along a row's quantization error, a query whose
own rounding error lines up with a row's codes, clipped coordinates, rows
far below the scale.  The worst case comes close to the bound, and leaving
out the row's quantization term or the query's rounding term fails.
"""
import numpy as np


def _row_terms(X, codes, s, drop=None):
    """(R, N) per row, as the library computes them (double, raised 1e-6,
    rounded to float)."""
    dim = X.shape[1]
    g = dim / 96.0 if dim > 96 else 1.0
    x = X.astype(np.float64)
    c = codes.astype(np.float64)
    quant = np.linalg.norm(x - c / s, axis=1)
    if drop == "quant":
        quant = 0.0 * quant
    R = (s * (quant + 5.8e-6 * g * np.linalg.norm(x, axis=1)) * (1 + 1e-6)).astype(np.float32)
    N = (np.linalg.norm(c, axis=1) * (1 + 1e-6)).astype(np.float32)
    return R, N


def _query_terms(q, cq, s, drop=None):
    """(A, B) as the kernel forms them: float32 products and sums, lanes of 64
    each summing every 64th element, then a butterfly."""
    v = (np.float32(s) * q.astype(np.float32)).astype(np.float32)
    w = (v - cq.astype(np.float32)).astype(np.float32)

    def lanesum(t):
        acc = np.zeros(64, np.float32)
        for j in range(len(t)):
            acc[j % 64] = np.float32(acc[j % 64] + t[j])
        off = 32
        while off >= 1:
            acc = (acc + acc[np.arange(64) ^ off]).astype(np.float32)
            off //= 2
        return acc[0]

    a = np.float32(np.sqrt(lanesum((v * v).astype(np.float32)))) * np.float32(1 + 2.0 ** -16)
    b = (np.float32(np.sqrt(lanesum((w * w).astype(np.float32)))) + np.float32(2.0 ** -20) * a) \
        * np.float32(1 + 2.0 ** -16)
    if drop == "query":
        b = np.float32(0)
    return float(a), float(b)


def _case(oracle, dim, seed):
    rng = np.random.default_rng(seed)
    n = 64
    X = rng.standard_normal((n, dim)) * np.exp(-np.arange(dim) / (dim / 4.0))[None, :]
    X[:4] *= 1e-3  # rows far below the scale: codes mostly 0
    X[4:6, :3] *= 40.0  # outliers past the trimmed maximum: clipped codes
    X = X.astype(np.float32)
    prm = oracle.q8_params(X, "ip")
    s = float(prm.scale)
    codes = oracle.q8_translate(prm, "ip", X)
    Q = [rng.standard_normal(dim) for _ in range(4)]
    for r in (4, 7, 9):  # along a row's quantization error (rows 4 and 7 clipped or not)
        e = X[r].astype(np.float64) - codes[r].astype(np.float64) / s
        Q.append(e / np.linalg.norm(e) * 3.0)
    for r in (8, 12):  # the query's own rounding error lined up with a row's codes
        c = codes[r].astype(np.float64)
        q = (np.round(rng.standard_normal(dim) * 20) + 0.49 * np.sign(c)) / s
        Q.append(q)
    Q.append(rng.standard_normal(dim) * 10.0 ** rng.uniform(-6, 0, dim))
    Q = [np.asarray(q, np.float32) for q in Q]
    return X, s, prm, codes, Q


def _worst_ratio(oracle, X, s, prm, codes, Q, drop=None):
    R, N = _row_terms(X, codes, s, drop)
    worst = 0.0
    for q in Q:
        cq = oracle.q8_translate(prm, "ip", q[None, :])[0]
        a, b = _query_terms(q, cq, s, drop)
        d = np.array([oracle.distance(c, cq, "ip") for c in codes], np.float64)
        d32 = np.array([oracle.distance(x, q, "ip") for x in X], np.float64)
        lim = a * R.astype(np.float64) + b * N.astype(np.float64)
        err = np.abs(d - s * s * d32)
        ok = lim > 0
        assert np.all(err[~ok] == 0)
        worst = max(worst, float(np.max(err[ok] / lim[ok])))
    return worst


def test_q8_codes_are_the_searched_distance(oracle):
    """The search distance is the integer -<code(q), code(x)>."""
    X, s, prm, codes, Q = _case(oracle, 96, 5)
    cq = oracle.q8_translate(prm, "ip", Q[0][None, :])[0]
    want = -(codes.astype(np.int64) @ cq.astype(np.int64))
    got = np.array([oracle.distance(c, cq, "ip") for c in codes])
    assert np.array_equal(got, want.astype(np.float64))
    assert np.abs(codes.astype(np.int64)).max() == 127  # the outliers clip


def test_q8_row_bound_holds(oracle):
    for dim in (16, 96, 200):
        X, s, prm, codes, Q = _case(oracle, dim, 7 + dim)
        worst = _worst_ratio(oracle, X, s, prm, codes, Q)
        assert worst <= 1.0, (dim, worst)
        assert worst > 0.3, (dim, worst)  # the cases come close to the bound


def test_q8_row_bound_needs_its_terms(oracle):
    """Without the row's quantization term, or without the query's rounding
    term, the distances leave the bound."""
    X, s, prm, codes, Q = _case(oracle, 96, 11)
    assert _worst_ratio(oracle, X, s, prm, codes, Q, drop="quant") > 1.0
    assert _worst_ratio(oracle, X, s, prm, codes, Q, drop="query") > 1.0
