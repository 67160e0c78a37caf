// bruteforce.hip -- brute-force nearest-neighbour candidates on MFMA (K7).
//
// The groundtruth computation of compute_groundtruth.cpp:22-60 (every query
// against every point, the k nearest kept) as a GEMM on the matrix cores
// with a fused selection.  Distances are FP32:
//   L2  : |x|^2 - 2 q.x   (|q|^2 is the same for all of a query's points and
//                          does not change its order)
//   MIPS: -q.x
// Each query keeps its kk smallest; the exact FP64 re-scoring and the final
// (distance, id) order of the first k are the caller's, so kk carries a
// margin over k for FP32 rounding.
//
// k_bf_topk: one wave per (32-query tile, point range).  The query tile and
// each 32-point tile are staged in LDS (integer points widened to float) and
// multiplied with v_mfma_f32_32x32x2f32: lane l supplies A[l%32][l/32] and
// B[l/32][l%32], register r of lane l receives C[(r/4)*8+(l/32)*4+r%4][l%32].
// A distance below its query's threshold is appended to the query's buffer
// (capacity >= kk + 32: room for one more tile); when a buffer could
// overflow it is sorted and cut to kk, and the kk-th distance becomes the
// threshold.  k_bf_merge sorts each query's per-range lists into its kk best.
#include "kernels.h"
#include "device_common.h"

#include <float.h>

namespace pann {

typedef float f32x16 __attribute__((ext_vector_type(16)));

// ascending bitonic sort of K[0..np) (np a power of two, <= 4096) in LDS
__device__ void lds_bitonic(uint64_t* K, int np, int lane) {
  for (int size = 2; size <= np; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < np / 2; t += PANN_WAVE) {
        const int lo = (t / stride) * (stride * 2) + (t % stride);
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t a = K[lo], b = K[hi];
        if ((a > b) == up) {
          K[lo] = b;
          K[hi] = a;
        }
      }
      wave_lds_sync();
    }
}

template <typename PT>
__global__ __launch_bounds__(PANN_WAVE) void k_bf_topk(BFArgs a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x;
  const uint32_t dp = a.dim_pad;  // dim rounded up to 2
  const uint32_t ld = dp + 1;     // odd LDS row stride: the 32 rows fall in distinct banks
  float* QL = (float*)smem;                 // 32 x ld
  float* PL = QL + 32 * ld;                 // 32 x ld
  float* NR = PL + 32 * ld;                 // 32 norms
  float* thr = NR + 32;                     // 32 thresholds
  uint32_t* cnt = (uint32_t*)(thr + 32);    // 32 counters
  uint64_t* S = (uint64_t*)(cnt + 32);      // sort scratch, cap keys

  const uint32_t qt = blockIdx.x / a.nranges, rg = blockIdx.x % a.nranges;
  const uint32_t q0 = qt * 32;
  const uint64_t pb = (uint64_t)rg * a.range_len;
  const uint64_t pe = pb + a.range_len < a.n ? pb + a.range_len : a.n;
  for (uint32_t e = lane; e < 32 * dp; e += PANN_WAVE) {
    const uint32_t i = e / dp, j = e % dp;
    QL[i * ld + j] = (q0 + i < a.nq && j < a.dim) ? a.Q[(uint64_t)(q0 + i) * a.qs + j] : 0.0f;
  }
  if (lane < 32) {
    thr[lane] = FLT_MAX;
    cnt[lane] = 0;
  }
  wave_lds_sync();
  const int li = lane & 31, lk = lane >> 5;
  uint64_t* bufq = a.buf + ((uint64_t)q0 * a.nranges + rg) * a.cap;  // query q0 + i: + i*nranges*cap

  auto compact = [&](uint32_t i) {  // sort query i's buffer, keep kk
    uint64_t* B = bufq + (uint64_t)i * a.nranges * a.cap;
    const uint32_t c = cnt[i];
    for (uint32_t t = lane; t < a.cap; t += PANN_WAVE) S[t] = t < c ? B[t] : ~0ull;
    wave_lds_sync();
    lds_bitonic(S, (int)a.cap, lane);
    const uint32_t keep = c < a.kk ? c : a.kk;
    for (uint32_t t = lane; t < keep; t += PANN_WAVE) B[t] = S[t];
    wave_lds_sync();
    if (lane == 0) {
      cnt[i] = keep;
      if (keep == a.kk) thr[i] = key_dist(S[a.kk - 1]);
    }
    wave_lds_sync();
  };

  for (uint64_t t0 = pb; t0 < pe; t0 += 32) {
    for (uint32_t e = lane; e < 32 * dp; e += PANN_WAVE) {
      const uint32_t i = e / dp, j = e % dp;
      PL[i * ld + j] =
          (t0 + i < pe && j < a.dim) ? (float)((const PT*)a.X)[(t0 + i) * a.xs + j] : 0.0f;
    }
    if (lane < 32) NR[lane] = (a.xnorm && t0 + lane < pe) ? a.xnorm[t0 + lane] : 0.0f;
    wave_lds_sync();
    f32x16 acc = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
                  0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t s = 0; s < dp / 2; s++) {
      const float av = QL[li * ld + 2 * s + lk];
      const float bv = PL[li * ld + 2 * s + lk];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
    }
    const uint64_t pid = t0 + (uint64_t)li;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const uint32_t i = (uint32_t)((r >> 2) * 8 + lk * 4 + (r & 3));
      if (pid < pe && q0 + i < a.nq) {
        const float d = a.metric == 0 ? NR[li] - 2.0f * acc[r] : -acc[r];
        if (d < thr[i]) {
          const uint32_t pos = atomicAdd(&cnt[i], 1u);
          bufq[(uint64_t)i * a.nranges * a.cap + pos] = make_key(d, (uint32_t)pid);
        }
      }
    }
    wave_lds_sync();
    // room for one more tile: a buffer past cap - 32 is cut back to kk
    for (uint32_t i = 0; i < 32; i++)
      if (cnt[i] > a.cap - 32) compact(i);
  }
  for (uint32_t i = 0; i < 32; i++)
    if (q0 + i < a.nq && cnt[i] > 0) compact(i);
  if (lane < 32 && q0 + lane < a.nq) a.counts[(uint64_t)(q0 + lane) * a.nranges + rg] = cnt[lane];
}

// one wave per query: the sorted per-range lists into the query's kk best
__global__ __launch_bounds__(PANN_WAVE) void k_bf_merge(BFArgs a, uint32_t* out_ids,
                                                        float* out_dists) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  uint64_t* S = (uint64_t*)smem;
  const int lane = threadIdx.x;
  const uint32_t q = blockIdx.x;
  if (q >= a.nq) return;
  const uint32_t tot = a.nranges * a.kk;
  uint32_t np = 2;
  while (np < tot) np <<= 1;
  for (uint32_t t = lane; t < np; t += PANN_WAVE) {
    uint64_t v = ~0ull;
    if (t < tot) {
      const uint32_t rg = t / a.kk, j = t % a.kk;
      if (j < a.counts[(uint64_t)q * a.nranges + rg])
        v = a.buf[((uint64_t)q * a.nranges + rg) * a.cap + j];
    }
    S[t] = v;
  }
  wave_lds_sync();
  lds_bitonic(S, (int)np, lane);
  for (uint32_t t = lane; t < a.kk; t += PANN_WAVE) {
    const uint64_t v = S[t];
    out_ids[(uint64_t)q * a.kk + t] = v == ~0ull ? PANN_SENTINEL : key_id(v);
    out_dists[(uint64_t)q * a.kk + t] = v == ~0ull ? FLT_MAX : key_dist(v);
  }
}

// |x|^2 per row, in FP32
template <typename PT>
__global__ void k_row_sqnorm(const PT* X, uint64_t xs, uint32_t n, uint32_t dim, float* out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t r = w; r < n; r += nw) {
    float s = 0.0f;
    for (uint32_t j = lane; j < dim; j += 64) {
      const float v = (float)X[(uint64_t)r * xs + j];
      s += v * v;
    }
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[r] = s;
  }
}

uint32_t bf_lds_bytes(uint32_t dim_pad, uint32_t cap) {
  return 2 * 32 * (dim_pad + 1) * 4 + 32 * 4 * 3 + ((cap * 8 + 15) & ~15u) + 16;
}

hipError_t launch_bf(const BFArgs& a, int point_dtype, uint32_t* out_ids, float* out_dists,
                     hipStream_t s) {
  const uint32_t lds = bf_lds_bytes(a.dim_pad, a.cap);
  const uint32_t blocks = ((a.nq + 31) / 32) * a.nranges;
  hipError_t e;
  if (point_dtype == 0) {
    e = hipFuncSetAttribute((const void*)k_bf_topk<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bf_topk<float>, dim3(blocks), dim3(PANN_WAVE), lds, s, a);
  } else if (point_dtype == 1) {
    e = hipFuncSetAttribute((const void*)k_bf_topk<uint8_t>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bf_topk<uint8_t>, dim3(blocks), dim3(PANN_WAVE), lds, s, a);
  } else {
    e = hipFuncSetAttribute((const void*)k_bf_topk<int8_t>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bf_topk<int8_t>, dim3(blocks), dim3(PANN_WAVE), lds, s, a);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  uint32_t np = 2;
  while (np < a.nranges * a.kk) np <<= 1;
  hipLaunchKernelGGL(k_bf_merge, dim3(a.nq), dim3(PANN_WAVE), np * 8, s, a, out_ids, out_dists);
  return hipGetLastError();
}

hipError_t launch_row_sqnorm(const void* X, int point_dtype, uint64_t xs, uint32_t n,
                             uint32_t dim, float* out, hipStream_t s) {
  const uint32_t blocks = (uint32_t)((n + 3) / 4 < 65536 ? (n + 3) / 4 : 65536);
  if (point_dtype == 0)
    hipLaunchKernelGGL(k_row_sqnorm<float>, dim3(blocks), dim3(256), 0, s, (const float*)X, xs,
                       n, dim, out);
  else if (point_dtype == 1)
    hipLaunchKernelGGL(k_row_sqnorm<uint8_t>, dim3(blocks), dim3(256), 0, s,
                       (const uint8_t*)X, xs, n, dim, out);
  else
    hipLaunchKernelGGL(k_row_sqnorm<int8_t>, dim3(blocks), dim3(256), 0, s, (const int8_t*)X,
                       xs, n, dim, out);
  return hipGetLastError();
}

}  // namespace pann
