"""The oracle's float distances against the reference's float loops as the
reference's compiler builds them.

The reference sums `result += (q[i] - p[i]) * (q[i] - p[i])` and `result +=
q[i] * p[i]` in a plain loop (euclidian_point.h:83-90, mips_point.h:60-66),
compiled with g++ -O3 -march=native (algorithms/bench/parallelDefsANN:8).
This test compiles that loop, restated below, with those flags and compares
its results bit for bit with the oracle's at every dimension from 1 to 300,
for both metrics.  It fixes which terms are fused: g++ 11.4 keeps the loop's
order, adds separately rounded products through an 8-wide and then a 4-wide
vector body, and contracts the last d % 4 terms into fused multiply-adds.  A
model without the fused tail (or with a fused body) fails here at dimensions
that are not multiples of 4.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

LOOP_SRC = r"""
#include <cstdio>
#include <cstdlib>
#include <vector>
// The reference's float loops (euclidian_point.h:83-90, mips_point.h:60-66).
__attribute__((noinline)) float l2(const float *p, const float *q, unsigned d) {
  float result = 0.0;
  for (int i = 0; i < d; i++) result += (q[i] - p[i]) * (q[i] - p[i]);
  return result;
}
__attribute__((noinline)) float ip(const float *p, const float *q, unsigned d) {
  float result = 0;
  for (int i = 0; i < d; i++) result += q[i] * p[i];
  return -result;
}
// stdin: u32 n, u32 d, then n pairs of rows (p, q); stdout: n l2, n ip floats
int main() {
  unsigned h[2];
  if (fread(h, 4, 2, stdin) != 2) return 2;
  std::vector<float> v((size_t)h[0] * 2 * h[1]);
  if (fread(v.data(), 4, v.size(), stdin) != v.size()) return 2;
  std::vector<float> out(2 * (size_t)h[0]);
  for (unsigned i = 0; i < h[0]; i++) {
    const float *p = &v[(size_t)i * 2 * h[1]], *q = p + h[1];
    out[i] = l2(p, q, h[1]);
    out[h[0] + i] = ip(p, q, h[1]);
  }
  fwrite(out.data(), 4, out.size(), stdout);
  return 0;
}
"""


def _has_fma():
    try:
        with open("/proc/cpuinfo") as f:
            return " fma " in f.read()
    except OSError:
        return False


@pytest.fixture(scope="module")
def loop_bins(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    if not _has_fma():
        pytest.skip("host has no FMA: -march=native would compile the loop without it")
    d = tmp_path_factory.mktemp("refloop")
    src = d / "loop.cpp"
    src.write_text(LOOP_SRC)
    bins = []
    for march in ("native", "x86-64-v3"):
        out = str(d / f"loop_{march}")
        subprocess.run(["g++", "-O3", "-std=c++17", f"-march={march}", "-DNDEBUG", str(src),
                        "-o", out], check=True)
        bins.append(out)
    return bins


def _run(binary, P, Q):
    n, dim = P.shape
    buf = np.empty((n, 2, dim), np.float32)
    buf[:, 0], buf[:, 1] = P, Q
    r = subprocess.run([binary], input=np.array([n, dim], np.uint32).tobytes() + buf.tobytes(),
                       capture_output=True, check=True)
    out = np.frombuffer(r.stdout, np.float32)
    return out[:n], out[n:]


def test_oracle_equals_compiled_reference_loop(oracle, loop_bins):
    rng = np.random.default_rng(77)
    for dim in range(1, 301):
        P = rng.standard_normal((40, dim)).astype(np.float32)
        Q = rng.standard_normal((40, dim)).astype(np.float32)
        want_l2 = np.array([oracle.distance(P[i], Q[i], "l2") for i in range(40)], np.float32)
        want_ip = np.array([oracle.distance(P[i], Q[i], "ip") for i in range(40)], np.float32)
        for b in loop_bins:
            l2, ip = _run(b, P, Q)
            assert np.array_equal(l2.view(np.uint32), want_l2.view(np.uint32)), (b, dim)
            assert np.array_equal(ip.view(np.uint32), want_ip.view(np.uint32)), (b, dim)


def test_fused_tail_is_visible(oracle):
    """The fused tail changes results at d % 4 != 0 and nothing at d % 4 == 0:
    against a separately rounded in-order sum emulated in numpy, d = 5 and 99
    differ for some rows and d = 96 and 100 never do."""
    rng = np.random.default_rng(5)

    def plain(p, q):
        r = np.float32(0)
        for i in range(len(p)):
            t = np.float32(q[i] - p[i])
            r = np.float32(r + np.float32(t * t))
        return r

    for dim, must_differ in ((5, True), (99, True), (96, False), (100, False)):
        diff = 0
        for _ in range(200):
            p = rng.standard_normal(dim).astype(np.float32)
            q = rng.standard_normal(dim).astype(np.float32)
            diff += oracle.distance(p, q, "l2") != plain(p, q)
        assert (diff > 0) == must_differ, (dim, diff)
