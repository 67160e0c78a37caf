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
// k_bf_topk: one wave per (32 or 64 queries, point range), multiplying with
// v_mfma_f32_32x32x2f32 (lane l supplies A[l%32][l/32] and B[l/32][l%32],
// register r of lane l receives C[(r/4)*8+(l/32)*4+r%4][l%32]).  The K index
// is permuted so that each lane streams a contiguous half row of its point
// into the B operand and keeps the matching half of its queries in
// registers; integer points are widened to float on the way.  A distance
// below its query's threshold is appended to the query's buffer (capacity
// >= kk + 32: room for one more tile); when a buffer could overflow it is
// sorted and cut to kk, and the kk-th distance becomes the threshold.
// k_bf_merge sorts each query's per-range lists into its kk best.
#include "kernels.h"
#include "device_common.h"

#include <float.h>

#include <type_traits>

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

// A lane's half row: lane l holds, of point row0 + l % 32, the dp / 2
// contiguous elements [h * dp / 2, (h + 1) * dp / 2), h = l / 32, as nw
// 16-byte words (rows are zero padded to 128 bytes, and dp is dim rounded up
// to 32 bytes, so every word lies inside its row).
template <typename PT, int NV>
__device__ __forceinline__ void load_half_rows(uint4 (&w)[NV], const PT* X, uint64_t xs,
                                               uint64_t row0, uint64_t rows_end, uint32_t half,
                                               uint32_t nw, int lane) {
  const uint64_t row = row0 + (uint64_t)(lane & 31);
  const bool live = row < rows_end;
  const uint4* p = (const uint4*)(X + row * xs + (uint64_t)(lane >> 5) * half);
#pragma unroll
  for (int t = 0; t < NV; t++)
    w[t] = (live && (uint32_t)t < nw) ? p[t] : make_uint4(0u, 0u, 0u, 0u);
}

template <typename PT>
__device__ __forceinline__ float word_elem(const uint4& w, int e) {
  if constexpr (sizeof(PT) == 4) {
    return __uint_as_float(e == 0 ? w.x : e == 1 ? w.y : e == 2 ? w.z : w.w);
  } else {
    const uint32_t word = (e >> 2) == 0 ? w.x : (e >> 2) == 1 ? w.y : (e >> 2) == 2 ? w.z : w.w;
    const uint32_t byte = (word >> (8 * (e & 3))) & 0xFFu;
    return std::is_same<PT, int8_t>::value ? (float)(int8_t)byte : (float)byte;
  }
}

// One wave per (32 * QS queries, point range).  The dot products are an
// MFMA GEMM whose K index is permuted: at step s, lane l (h = l / 32) feeds
// dimension h * dp / 2 + s of query / point l % 32 as its A / B element, so
// each lane streams its own contiguous half row from global memory straight
// into the matrix core (no LDS staging) and holds its queries' matching
// halves in registers for the whole range.  The next tile's words are in
// flight while the current tile is multiplied.
// MINW: the waves per SIMD the register allocation must allow (2 where the
// compiler would otherwise settle on 1 but can fit 256 registers unspilled)
template <typename PT, int NV, int QS, int MINW>
__global__ __launch_bounds__(PANN_WAVE) __attribute__((amdgpu_waves_per_eu(MINW, 2))) void k_bf_topk(BFArgs a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  constexpr int EPW = 16 / (int)sizeof(PT);  // elements per 16-byte word
  constexpr int SMAX = NV * EPW;              // steps per lane, upper bound
  constexpr int NQ = 32 * QS;                 // queries per wave
  const int lane = threadIdx.x;
  const int li = lane & 31, h = lane >> 5;
  float* thr = (float*)smem;                   // NQ thresholds
  uint32_t* cnt = (uint32_t*)(thr + NQ);       // NQ counters
  uint64_t* S = (uint64_t*)(cnt + NQ);         // sort scratch, cap keys
  const uint32_t half = a.dim_pad / 2;
  const uint32_t nw = half / EPW;

  const uint32_t qt = blockIdx.x / a.nranges, rg = blockIdx.x % a.nranges;
  const uint32_t q0 = qt * NQ;
  const uint64_t pb = (uint64_t)rg * a.range_len;
  const uint64_t pe = pb + a.range_len < a.n ? pb + a.range_len : a.n;

  // this lane's halves of its queries (zero past dim)
  float qa[QS][SMAX];
#pragma unroll
  for (int g = 0; g < QS; g++) {
    const uint32_t row = q0 + (uint32_t)(g * 32 + li);
    const float* qr = a.Q + (uint64_t)(row < a.nq ? row : 0) * a.qs;
#pragma unroll
    for (int s = 0; s < SMAX; s++) {
      const uint32_t j = (uint32_t)h * half + (uint32_t)s;
      qa[g][s] = (row < a.nq && (uint32_t)s < half && j < a.dim) ? qr[j] : 0.0f;
    }
  }
  for (int i = lane; i < NQ; i += PANN_WAVE) {
    thr[i] = FLT_MAX;
    cnt[i] = 0;
  }
  wave_lds_sync();
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

  // lane j < NQ holds query j's threshold; the selection fetches it with
  // v_readlane at a compile-time lane, off the LDS pipe (thresholds change
  // only when a buffer is cut, which reloads them)
  static_assert(NQ <= PANN_WAVE, "one threshold per lane");
  float thr_reg = FLT_MAX;
  const PT* X = (const PT*)a.X;
  uint4 wc[NV], wn[NV];
  load_half_rows<PT, NV>(wc, X, a.xs, pb, pe, half, nw, lane);
  float nc = (a.xnorm && pb + li < pe) ? a.xnorm[pb + li] : 0.0f, nn = 0.0f;
  for (uint64_t t0 = pb; t0 < pe; t0 += 32) {
    if (t0 + 32 < pe) {
      load_half_rows<PT, NV>(wn, X, a.xs, t0 + 32, pe, half, nw, lane);
      nn = (a.xnorm && t0 + 32 + li < pe) ? a.xnorm[t0 + 32 + li] : 0.0f;
    }
    f32x16 acc[QS];
#pragma unroll
    for (int g = 0; g < QS; g++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[g][r] = 0.0f;
#pragma unroll
    for (int t = 0; t < NV; t++) {  // words past nw are zero, as are their query halves
#pragma unroll
      for (int e = 0; e < EPW; e++) {
        const float b = word_elem<PT>(wc[t], e);
#pragma unroll
        for (int g = 0; g < QS; g++)
          acc[g] = __builtin_amdgcn_mfma_f32_32x32x2f32(qa[g][t * EPW + e], b, acc[g], 0, 0, 0);
      }
    }
    const uint64_t pid = t0 + (uint64_t)li;
    if (pid < pe) {
#pragma unroll
      for (int g = 0; g < QS; g++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int i0 = g * 32 + (r >> 2) * 8 + (r & 3);  // query of this element, h = 0
          const uint32_t i = (uint32_t)(i0 + h * 4);
          const float t0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(thr_reg), i0));
          const float t1 =
              __int_as_float(__builtin_amdgcn_readlane(__float_as_int(thr_reg), i0 + 4));
          const float d = a.metric == 0 ? nc - 2.0f * acc[g][r] : -acc[g][r];
          if (q0 + i < a.nq && d < (h ? t1 : t0)) {
            const uint32_t pos = atomicAdd(&cnt[i], 1u);
            bufq[(uint64_t)i * a.nranges * a.cap + pos] = make_key(d, (uint32_t)pid);
          }
        }
    }
    wave_lds_sync();
    // room for one more tile: a buffer past cap - 32 is cut back to kk
    uint64_t full = __ballot(lane < NQ && cnt[lane < NQ ? lane : 0] > a.cap - 32);
    if (full) {
      while (full) {
        const uint32_t i = (uint32_t)__ffsll((long long)full) - 1;
        full &= full - 1;
        compact(i);
      }
      thr_reg = lane < NQ ? thr[lane] : FLT_MAX;
    }
#pragma unroll
    for (int t = 0; t < NV; t++) wc[t] = wn[t];
    nc = nn;
  }
  // the merge reads each buffer's first kk keys: cut longer ones
  uint64_t over = __ballot(lane < NQ && q0 + lane < a.nq && cnt[lane < NQ ? lane : 0] > a.kk);
  while (over) {
    const uint32_t i = (uint32_t)__ffsll((long long)over) - 1;
    over &= over - 1;
    compact(i);
  }
  for (int i = lane; i < NQ; i += PANN_WAVE)
    if (q0 + i < a.nq) a.counts[(uint64_t)(q0 + i) * a.nranges + rg] = cnt[i];
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

uint32_t bf_lds_bytes(uint32_t cap) { return 64 * 4 * 2 + cap * 8; }

uint32_t bf_dim_pad(uint32_t dim, int point_dtype) {
  return point_dtype == 0 ? (dim + 7) & ~7u : (dim + 31) & ~31u;  // 32 bytes
}

template <typename PT, int NV, int QS, int MINW = 1>
static hipError_t launch_topk_t(const BFArgs& a, hipStream_t s) {
  const uint32_t lds = bf_lds_bytes(a.cap);
  const uint32_t blocks = ((a.nq + 32 * QS - 1) / (32 * QS)) * a.nranges;
  hipLaunchKernelGGL((k_bf_topk<PT, NV, QS, MINW>), dim3(blocks), dim3(PANN_WAVE), lds, s, a);
  return hipGetLastError();
}

// words per lane nw = dim_pad / 2 / (16 / sizeof(PT)).  One query tile (32
// queries) per wave where that leaves room for two waves per SIMD, so one
// wave's selection overlaps the other's multiplies (measured: f32 d = 64 and
// 96, u8 d = 128 gain 12-33 % over two tiles per wave at one wave per SIMD);
// two tiles for the smallest rows, which reach two waves per SIMD with them
template <typename PT>
static hipError_t launch_topk_nv(const BFArgs& a, hipStream_t s) {
  const uint32_t nw = a.dim_pad / 2 / (16 / (uint32_t)sizeof(PT));
  if constexpr (sizeof(PT) == 4) {
    if (nw <= 4) return launch_topk_t<PT, 4, 2>(a, s);
    if (nw <= 8) return launch_topk_t<PT, 8, 1>(a, s);
    if (nw <= 12) return launch_topk_t<PT, 12, 1>(a, s);
    if (nw <= 16) return launch_topk_t<PT, 16, 1, 2>(a, s);
    if (nw <= 24) return launch_topk_t<PT, 24, 1>(a, s);
    if (nw <= 32) return launch_topk_t<PT, 32, 1>(a, s);
  } else {
    if (nw <= 2) return launch_topk_t<PT, 2, 2>(a, s);
    if (nw <= 4) return launch_topk_t<PT, 4, 1>(a, s);
    if (nw <= 8) return launch_topk_t<PT, 8, 1>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_bf(const BFArgs& a, int point_dtype, uint32_t* out_ids, float* out_dists,
                     hipStream_t s) {
  hipError_t e = point_dtype == 0 ? launch_topk_nv<float>(a, s)
                 : point_dtype == 1 ? launch_topk_nv<uint8_t>(a, s)
                                    : launch_topk_nv<int8_t>(a, s);
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


// ---------------------------------------------------------------------------
// The FP8 residual store's per-query centroid table T = Q . C^T (nq x ncent,
// row-major; Q: nq rows of qs floats, dim used; C: ncent rows of cs floats,
// zero past dim up to a multiple of 8, 16-byte aligned) on v_mfma_f32_32x32x2f32.  One wave
// per tile of 32 queries x 32 * QC_TN centroids.  The K index is permuted as
// in k_bf_topk: over a block of 8 dimensions, MFMA step s takes dimension
// k0 + (l / 32) * 4 + s in lane l, so each lane loads the four it needs with
// one 16-byte load of its query row (A) and one of each of its centroid rows
// (B).  The QC_TN accumulators are independent, so their MFMAs overlap.
// ---------------------------------------------------------------------------
constexpr int QC_TN = 4;

__global__ __launch_bounds__(PANN_WAVE) void k_qc_table(const float* Q, uint64_t qs, uint32_t nq,
                                                       uint32_t dim, const float* C,
                                                       uint64_t cs, uint32_t ncent, float* T) {
  const int lane = threadIdx.x;
  const uint32_t q0 = blockIdx.x * 32u, c0 = blockIdx.y * 32u * QC_TN;
  const uint32_t qa = q0 + (uint32_t)(lane & 31);
  const uint32_t half = ((uint32_t)lane >> 5) * 4u;
  const float* qrow = Q + (uint64_t)(qa < nq ? qa : nq - 1) * qs;
  const float* crow[QC_TN];
  bool cok[QC_TN];
#pragma unroll
  for (int t = 0; t < QC_TN; t++) {
    const uint32_t c = c0 + (uint32_t)t * 32u + (uint32_t)(lane & 31);
    cok[t] = c < ncent;
    crow[t] = C + (uint64_t)(cok[t] ? c : 0u) * cs;
  }
  f32x16 acc[QC_TN];
#pragma unroll
  for (int t = 0; t < QC_TN; t++)
#pragma unroll
    for (int r = 0; r < 16; r++) acc[t][r] = 0.0f;
  // 16-byte query loads when every row starts 16-byte aligned (uniform)
  const bool qvec = (qs & 3u) == 0 && ((uintptr_t)Q & 15u) == 0;
  const uint32_t dfull = qvec ? dim & ~7u : 0u;
  for (uint32_t k0 = 0; k0 < dim; k0 += 8) {
    float a[4];
    if (k0 < dfull) {  // uniform: a whole block of 8 inside the query row
      const float4 v = *(const float4*)(qrow + k0 + half);
      a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
    } else {  // unaligned rows, or the last dim % 8 dimensions: the row ends at dim
#pragma unroll
      for (int e = 0; e < 4; e++) a[e] = k0 + half + e < dim ? qrow[k0 + half + e] : 0.0f;
    }
#pragma unroll
    for (int t = 0; t < QC_TN; t++) {
      // centroid rows are zero from dim to the next multiple of 8
      const float4 v = cok[t] ? *(const float4*)(crow[t] + k0 + half) : make_float4(0, 0, 0, 0);
      const float b[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 4; e++)
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[e], b[e], acc[t], 0, 0, 0);
    }
  }
  // register r of lane l holds C[(r / 4) * 8 + (l / 32) * 4 + r % 4][l % 32]
#pragma unroll
  for (int t = 0; t < QC_TN; t++)
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const uint32_t q = q0 + (uint32_t)((r >> 2) * 8 + (lane >> 5) * 4 + (r & 3));
      const uint32_t c = c0 + (uint32_t)t * 32u + (uint32_t)(lane & 31);
      if (q < nq && c < ncent) T[(uint64_t)q * ncent + c] = acc[t][r];
    }
}

hipError_t launch_qc_table(const float* Q, uint64_t qs, uint32_t nq, uint32_t dim,
                           const float* C, uint64_t cs, uint32_t ncent, float* T,
                           hipStream_t s) {
  if (nq == 0 || ncent == 0) return hipSuccess;
  // the centroid rows' 16-byte loads reach dim rounded up to 8
  if (cs < ((dim + 7u) & ~7u) || (cs & 3u) || ((uintptr_t)C & 15u)) return hipErrorInvalidValue;
  const dim3 grid((nq + 31u) / 32u, (ncent + 32u * QC_TN - 1u) / (32u * QC_TN));
  hipLaunchKernelGGL(k_qc_table, grid, dim3(PANN_WAVE), 0, s, Q, qs, nq, dim, C, cs, ncent, T);
  return hipGetLastError();
}

}  // namespace pann
