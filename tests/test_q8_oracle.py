"""Pins the oracle's restatement of the reference's 8-bit scalar quantization
(Euclidian_Point<uint8_t>, euclidian_point.h:98-107, 188-243;
Quantized_Mips_Point<8, true>, mips_point.h:292-470) on cases whose answer the
definitions fix: the all-integer identity, exactly representable slopes and
offsets, the trimmed order statistic against numpy's sort, and rounding half
away from zero (2.5 -> 3, which round-half-even would make 2)."""
import numpy as np


def test_l2_all_integer_identity(oracle, rng):
    X = rng.integers(0, 200, (300, 16)).astype(np.float32)
    p = oracle.q8_params(X, "l2")
    assert p.slope == 1.0 and p.offset == 0
    assert np.array_equal(oracle.q8_translate(p, "l2", X), X.astype(np.uint8))
    # a non-integer value keeps slope 1 / offset 0 only by accident of range:
    # the direct copy truncates (translate_point's slope == 1 branch)
    Y = np.array([[0.0, 255.0, 3.7]], np.float32)
    assert oracle.q8_translate(p, "l2", Y).tolist() == [[0, 255, 3]]


def test_l2_affine_closed_form(oracle):
    # values span [-1, 3]: slope = 255 / 4 = 63.75, offset = round(-63.75) = -64
    X = np.array([[-1.0, 3.0, 0.0, 1.0], [-0.5, 2.0, 0.25, 1.5]], np.float32)
    p = oracle.q8_params(X, "l2")
    assert p.slope == np.float32(63.75) and p.offset == -64
    codes = oracle.q8_translate(p, "l2", np.array([[-1, 3, 0, 1, -0.5, -5, 9]], np.float32))
    # round(x * 63.75) + 64, clamped: -63.75 -> -64; 191.25 -> 191; 63.75 -> 64;
    # -31.875 -> -32; out-of-range values clamp to 0 / 255
    assert codes.tolist() == [[0, 255, 64, 128, 32, 0, 255]]


def test_mips_trimmed_bound_matches_sort(oracle, rng):
    n = 20000
    X = np.zeros((n, 4), np.float32)
    X[:, 0] = rng.permutation(n).astype(np.float32) * np.float32(0.01)   # per-point maxima
    X[:, 1] = -rng.permutation(n).astype(np.float32) * np.float32(0.02)  # per-point minima
    p = oracle.q8_params(X, "ip")
    maxs = np.sort(np.maximum(X.max(1), 0))
    mins = np.sort(np.minimum(X.min(1), 0))
    hi = maxs[int((1.0 - 1e-4) * (n - 1))]          # rank 19997
    lo = mins[int(1e-4 * n)]                        # rank 2
    assert p.max_val == max(hi, -lo)
    assert p.scale == np.float32(127) / np.float32(p.max_val)


def test_mips_rounding_and_clamp(oracle):
    # one point whose trimmed bound is exactly 127 -> scale 1: codes = round(x)
    X = np.array([[127.0, -3.0]], np.float32)
    p = oracle.q8_params(X, "ip")
    assert p.max_val == 127.0 and p.scale == 1.0
    q = np.array([[2.5, -2.5, 62.5, -62.5, 0.49, 127.5, -200.0, 126.6]], np.float32)
    assert oracle.q8_translate(p, "ip", q).tolist() == [[3, -3, 63, -63, 0, 127, -127, 127]]
