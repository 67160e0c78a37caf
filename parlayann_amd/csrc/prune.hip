// prune.hip -- Vamana insertion steps on the device (K5).
//
// k_robust_prune restates knn_index::robustPrune (vamana/index.h:63-137) for
// one point per 64-lane wave:
//   candidates (+ p's current out-neighbours when `add`, :72-77) are sorted by
//   (dist, id) (:79-83) with a bitonic network in LDS, de-duplicated by id
//   (:85-88), then the greedy loop (:95-116) takes the next surviving
//   candidate p* (skipping p itself) and, in parallel over the remaining
//   candidates p', removes those with alpha * d(p*, p') <= d(p, p'); the
//   comparison is done in double exactly as the reference writes it (:111).
// Distances are the reference's (euclidian_point.h:45-90, mips_point.h:44-66):
// exact int32 for uint8/int8, in-order FP32 sums otherwise.
//
// The same kernel runs HNSW's two neighbour-selection steps (HNSW.hpp):
//   add = 4: prune_heuristic (:430-484) on a search frontier: candidates
//     sorted by distance, then c is dropped when some already selected r has
//     d(c, r) < alpha * d(c, u), compared in float; no dedup, no self skip;
//   add = 5: the reverse-edge step of insert (:952-983): the added ids are
//     appended to the current list when both fit in R, otherwise the union is
//     sorted by distance to the point and cut to R.
#include <algorithm>

#include "kernels.h"
#include "dist.h"

namespace pann {

// Rows evaluated per block of the pruning steps: their ids, distances and
// sort positions stay in LDS.
constexpr int PRUNE_BLK = 128;

// Integer rows (uint8 / int8): distances of rows NL[0..n) to the row QL (LDS,
// zero-padded to whole 128-byte lines) into DL, exact int32 as in
// euclidian_point.h:45-56 / mips_point.h:44-50.  Rows are gathered 8 at a
// time, one 128-byte line per 8-lane group and instruction, as the search
// gathers them; L2 is sum q^2 + sum p^2 - 2 sum q.p (mod 2^32: the exact
// value, which is below 2^31).
template <int DT, int MT>
__device__ __forceinline__ void eval_int_lds(const char* QL, int row_bytes, const char* base,
                                             uint64_t stride, const uint32_t* NL, int n,
                                             float* DL, int lane) {
  constexpr int UP = 4;
  const int sub = lane & 7, grp = lane >> 3;
  const int nl = (row_bytes + 127) >> 7;
  for (int r0 = 0; r0 < n; r0 += 8 * UP) {
    const char* row[UP];
#pragma unroll
    for (int u = 0; u < UP; u++) {
      const int r = r0 + u * 8 + grp;
      row[u] = base + (uint64_t)NL[r < n ? r : 0] * stride;
    }
    uint32_t pq[UP] = {}, pp[UP] = {}, qq = 0u;
    for (int c = 0; c < nl; c++) {  // uniform
      const uint4 qv = *(const uint4*)(QL + c * 128 + sub * 16);
      const uint32_t qw[4] = {qv.x, qv.y, qv.z, qv.w};
      uint4 raw[UP];
#pragma unroll
      for (int u = 0; u < UP; u++) raw[u] = ld16(row[u] + c * 128 + sub * 16);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        if constexpr (DT == ST_U8) {
          if (MT == M_L2) qq = __builtin_amdgcn_udot4(qw[e], qw[e], qq, false);
        } else {
          if (MT == M_L2) qq = (uint32_t)__builtin_amdgcn_sdot4((int)qw[e], (int)qw[e], (int)qq, false);
        }
      }
#pragma unroll
      for (int u = 0; u < UP; u++) {
        const uint32_t w[4] = {raw[u].x, raw[u].y, raw[u].z, raw[u].w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
          if constexpr (DT == ST_U8) {
            if (MT == M_L2) pp[u] = __builtin_amdgcn_udot4(w[e], w[e], pp[u], false);
            pq[u] = __builtin_amdgcn_udot4(w[e], qw[e], pq[u], false);
          } else {
            if (MT == M_L2)
              pp[u] = (uint32_t)__builtin_amdgcn_sdot4((int)w[e], (int)w[e], (int)pp[u], false);
            pq[u] = (uint32_t)__builtin_amdgcn_sdot4((int)w[e], (int)qw[e], (int)pq[u], false);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < UP; u++) {
      const int r = r0 + u * 8 + grp;
      const int v = MT == M_L2 ? (int)(qq + pp[u] - 2u * pq[u]) : (int)pq[u];
      const int s = group8_sum(v);
      if (sub == 0 && r < n) DL[r] = MT == M_L2 ? (float)s : -((float)s);
    }
  }
}

// Distances of the stored rows NL[0..n) to the row QL (LDS) into DL, the
// reference's arithmetic: in-order FP32 sums (the search's 8-lane gather and
// DPP chain, bit-identical to the reference's float loop) or exact int32.
template <int DT, int MT>
__device__ __forceinline__ void eval_rows_lds(const char* QL, int dim, const char* base,
                                              uint64_t stride, const uint32_t* NL, int n,
                                              float* DL, int lane) {
  if constexpr (DT == ST_F32)
    eval_group_inorder_lds<MT, 4, 2>((const float*)QL, dim, base, stride, NL, n, DL, lane);
  else
    eval_int_lds<DT, MT>(QL, dim * store_elem_bytes(DT), base, stride, NL, n, DL, lane);
}

// Orders this wave's accesses to its sort buffer (LDS, or global scratch
// that only this wave touches: workgroup scope suffices, the wave's CU L1
// serves its own stores).
template <bool KG>
__device__ __forceinline__ void wave_sync_k() {
  if constexpr (KG) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    wave_lds_sync();
  }
}

// ascending bitonic sort of K[0..np) (np a power of two) by one wave
template <bool KG>
__device__ void wave_bitonic_sort(uint64_t* K, int np, int lane) {
  for (int size = 2; size <= np; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < np / 2; t += PANN_WAVE) {
        int lo = (t / stride) * (stride * 2) + (t % stride);
        int hi = lo + stride;
        bool up = ((lo & size) == 0);
        uint64_t a = K[lo], b = K[hi];
        if ((a > b) == up) {
          K[lo] = b;
          K[hi] = a;
        }
      }
      wave_sync_k<KG>();
    }
  }
}



// One stored row (whole 128-byte lines: its zero padding included) into LDS.
__device__ __forceinline__ void row_to_lds(char* dst, const char* row, uint32_t row_lines,
                                           int lane) {
  for (uint32_t j = lane; j < row_lines * 8u; j += PANN_WAVE)
    *(uint4*)(dst + j * 16u) = ld16(row + j * 16u);
}

template <int DT, int MT, bool KG>
__device__ void prune_one(const PruneArgs& a, uint32_t i, int lane, char* QP, char* QS,
                          uint64_t* K, uint8_t* PR, uint32_t* NL, float* DL, uint32_t* TI,
                          uint64_t& evals) {
  const uint32_t p = a.points[i];
  const int dim = (int)a.dim;
  const uint32_t row_lines = (a.dim * (uint32_t)store_elem_bytes(DT) + 127u) >> 7;

  row_to_lds(QP, a.pts + (uint64_t)p * a.stride, row_lines, lane);
  wave_lds_sync();

  uint32_t nc;
  const uint32_t* cid;
  const float* cdist = nullptr;
  if (a.cand_offsets) {  // CSR groups
    nc = a.cand_offsets[i + 1] - a.cand_offsets[i];
    cid = a.cand_ids + a.cand_offsets[i];
  } else {
    nc = a.cand_count[i];
    cid = a.cand_ids + (uint64_t)i * a.cand_stride;
    if (a.cand_dists) cdist = a.cand_dists + (uint64_t)i * a.cand_stride;
  }
  if (a.add == 5) {
    const uint32_t* row = a.nbrs + (uint64_t)p * a.nbr_stride;
    uint32_t deg = 0;
    for (uint32_t t = 0; t < a.max_degree; t += PANN_WAVE) {
      uint32_t tt = t + lane;
      deg += __popcll(ballot(tt < a.max_degree && row[tt] != PANN_SENTINEL));
    }
    if (nc + deg <= a.R) {  // nbh_v.insert(end, adds) (:982)
      uint32_t* out = a.out_ids + (uint64_t)i * a.R;
      for (uint32_t t = lane; t < deg; t += PANN_WAVE) out[t] = row[t];
      for (uint32_t t = lane; t < nc; t += PANN_WAVE) out[deg + t] = cid[t];
      if (lane == 0) a.out_count[i] = deg + nc;
      wave_sync_k<KG>();
      return;
    }
  }
  if (a.add == 2) {
    // reverse-edge rule of batch_insert (vamana/index.h:292-303): when the
    // candidates fit next to the current list, append without repeats
    // (add_neighbors_without_repeats, :140-146: candidates first, then the
    // current neighbours that are not candidates); otherwise robustPrune.
    const uint32_t* row = a.nbrs + (uint64_t)p * a.nbr_stride;
    uint32_t deg = 0;
    for (uint32_t t = 0; t < a.max_degree; t += PANN_WAVE) {
      uint32_t tt = t + lane;
      deg += __popcll(ballot(tt < a.max_degree && row[tt] != PANN_SENTINEL));
    }
    if (nc + deg <= a.R) {
      uint32_t* out = a.out_ids + (uint64_t)i * a.R;
      for (uint32_t t = lane; t < nc; t += PANN_WAVE) out[t] = cid[t];
      uint32_t w = nc;
      for (uint32_t t0 = 0; t0 < deg; t0 += PANN_WAVE) {
        uint32_t t = t0 + lane;
        bool keep = false;
        uint32_t x = 0;
        if (t < deg) {
          x = row[t];
          keep = true;
          for (uint32_t j = 0; j < nc; j++)
            if (cid[j] == x) {
              keep = false;
              break;
            }
        }
        uint64_t m = ballot(keep);
        if (keep) out[w + mbcnt(m)] = x;
        w += (uint32_t)__popcll(m);
      }
      if (lane == 0) a.out_count[i] = w;
      wave_sync_k<KG>();
      return;
    }
  }
  // keys (d(p, c), c) of the candidates and, with `add`, of p's current
  // out-neighbours (:72-77), the distances evaluated PRUNE_BLK rows at a time
  uint32_t total = nc;
  if (cdist) {
    for (uint32_t t = lane; t < nc; t += PANN_WAVE) K[t] = make_key(cdist[t], cid[t]);
  } else {
    for (uint32_t t0 = 0; t0 < nc; t0 += PRUNE_BLK) {
      const int m = (int)(nc - t0 < (uint32_t)PRUNE_BLK ? nc - t0 : (uint32_t)PRUNE_BLK);
      for (int t = lane; t < m; t += PANN_WAVE) NL[t] = cid[t0 + t];
      wave_lds_sync();
      eval_rows_lds<DT, MT>(QP, dim, a.pts, a.stride, NL, m, DL, lane);
      evals += (uint64_t)m;
      wave_lds_sync();
      for (int t = lane; t < m; t += PANN_WAVE) K[t0 + t] = make_key(DL[t], NL[t]);
      wave_lds_sync();
    }
  }
  if (a.add != 0 && a.add != 4) {  // :72-77 (robustPrune's add, also the reverse-edge rules)
    const uint32_t* row = a.nbrs + (uint64_t)p * a.nbr_stride;
    // the row's ids before its first sentinel, in order
    for (uint32_t t0 = 0; t0 < a.max_degree; t0 += PRUNE_BLK) {
      const int m = (int)(a.max_degree - t0 < (uint32_t)PRUNE_BLK ? a.max_degree - t0
                                                                   : (uint32_t)PRUNE_BLK);
      int cnt = 0;
      for (int b = 0; b < m; b += PANN_WAVE) {
        const int t = b + lane;
        const uint32_t c = t < m ? row[t0 + t] : PANN_SENTINEL;
        const bool v = c != PANN_SENTINEL;
        const uint64_t bm = ballot(v);
        if (v) NL[cnt + (int)mbcnt(bm)] = c;
        cnt += __popcll(bm);
      }
      wave_lds_sync();
      eval_rows_lds<DT, MT>(QP, dim, a.pts, a.stride, NL, cnt, DL, lane);
      evals += (uint64_t)cnt;
      wave_lds_sync();
      for (int t = lane; t < cnt; t += PANN_WAVE) K[total + t] = make_key(DL[t], NL[t]);
      total += (uint32_t)cnt;
      wave_lds_sync();
    }
  }
  int np2 = 1;
  while ((uint32_t)np2 < total) np2 <<= 1;
  if (np2 < 2) np2 = 2;
  for (uint32_t t = total + lane; t < (uint32_t)np2; t += PANN_WAVE) K[t] = ~0ull;
  wave_sync_k<KG>();
  wave_bitonic_sort<KG>(K, np2, lane);  // :83
  if (a.add == 3 || a.add == 5) {  // sort, then cut to R: no pruning
    uint32_t* out = a.out_ids + (uint64_t)i * a.R;
    const uint32_t m = total < a.R ? total : a.R;
    for (uint32_t t = lane; t < m; t += PANN_WAVE) out[t] = key_id(K[t]);
    if (lane == 0) a.out_count[i] = m;
    wave_sync_k<KG>();
    return;
  }
  const bool hnsw = a.add == 4;
  const float alpha_f = (float)a.alpha;
  // unique by id (:86-88): equal ids sit next to each other (same distance);
  // prune_heuristic keeps repeats
  for (uint32_t t = lane; t < total; t += PANN_WAVE)
    PR[t] = (!hnsw && t > 0 && key_id(K[t]) == key_id(K[t - 1])) ? 1 : 0;
  wave_sync_k<KG>();

  uint32_t nout = 0;
  uint32_t* out = a.out_ids + (uint64_t)i * a.R;
  for (uint32_t idx = 0; idx < total && nout < a.R; idx++) {  // :95-116
    if (PR[idx]) continue;
    const uint32_t ps = key_id(K[idx]);
    if (!hnsw && ps == p) continue;
    if (lane == 0) out[nout] = ps;
    nout++;
    row_to_lds(QS, a.pts + (uint64_t)ps * a.stride, row_lines, lane);
    wave_sync_k<KG>();
    // the surviving candidates after idx, PRUNE_BLK at a time: their ids into
    // NL (positions in TI), d(p*, p') into DL, then the pruning test (:108-113)
    for (uint32_t tb = idx + 1; tb < total;) {
      int cnt = 0;
      while (tb < total && cnt <= PRUNE_BLK - PANN_WAVE) {
        const uint32_t t = tb + lane;
        const bool v = t < total && !PR[t];
        const uint64_t bm = ballot(v);
        if (v) {
          const int o = cnt + (int)mbcnt(bm);
          NL[o] = key_id(K[t]);
          TI[o] = t;
        }
        cnt += __popcll(bm);
        tb += PANN_WAVE;
      }
      wave_lds_sync();
      eval_rows_lds<DT, MT>(QS, dim, a.pts, a.stride, NL, cnt, DL, lane);
      evals += (uint64_t)cnt;
      wave_lds_sync();
      for (int j = lane; j < cnt; j += PANN_WAVE) {
        const uint32_t t = TI[j];
        const float d_sp = DL[j];
        const float d_pc = key_dist(K[t]);
        if (hnsw) {
          if (d_sp < d_pc * alpha_f) PR[t] = 1;  // HNSW.hpp:460-470
        } else if (a.alpha * (double)d_sp <= (double)d_pc) {
          PR[t] = 1;
        }
      }
      wave_sync_k<KG>();
    }
  }
  if (lane == 0) a.out_count[i] = nout;
}


// KG: sort keys and pruned flags live in a per-block global scratch (groups too
// large for LDS); otherwise in LDS.  Blocks loop over points.
template <int DT, int MT, bool KG>
__global__ __launch_bounds__(PANN_WAVE) void k_robust_prune(PruneArgs a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x;
  const uint32_t row = ((a.dim * (uint32_t)store_elem_bytes(DT) + 127u) >> 7) * 128u;
  char* QP = smem;       // row of p (whole lines)
  char* QS = QP + row;   // row of p*
  uint32_t* NL = (uint32_t*)(QS + row);  // a block of rows to evaluate ...
  float* DL = (float*)(NL + PRUNE_BLK);   // ... their distances ...
  uint32_t* TI = (uint32_t*)(DL + PRUNE_BLK);  // ... and sort positions
  char* tail = (char*)(TI + PRUNE_BLK);
  uint64_t* K = KG ? a.gkeys + (uint64_t)blockIdx.x * a.sort_cap : (uint64_t*)tail;
  uint8_t* PR = KG ? a.gflags + (uint64_t)blockIdx.x * a.sort_cap
                   : (uint8_t*)(tail + a.sort_cap * 8);
  uint64_t evals = 0;  // uniform over the wave
  for (uint32_t i = blockIdx.x; i < a.np; i += gridDim.x) {
    prune_one<DT, MT, KG>(a, i, lane, QP, QS, K, PR, NL, DL, TI, evals);
  }
  if (a.evals && lane == 0 && evals) atomicAdd(a.evals, (unsigned long long)evals);
}

// acc[0] += the sum of v[0..n): one wave per block, one atomic per block
__global__ __launch_bounds__(PANN_WAVE) void k_sum_u32(const uint32_t* v, uint32_t n,
                                                       unsigned long long* acc) {
  unsigned long long s = 0;
  for (uint32_t i = blockIdx.x * PANN_WAVE + threadIdx.x; i < n; i += gridDim.x * PANN_WAVE)
    s += v[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (threadIdx.x == 0 && s) atomicAdd(acc, s);
}
// Sorts each query's visited keys and splits them into ids / dists.
__global__ __launch_bounds__(PANN_WAVE) void k_sort_visited(const uint64_t* vis, uint32_t nq,
                                                           uint32_t cap, const uint32_t* counts,
                                                           uint32_t* ids, float* dists) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  uint64_t* K = (uint64_t*)smem;
  const int lane = threadIdx.x;
  const uint32_t q = blockIdx.x;
  if (q >= nq) return;
  uint32_t n = counts[q] < cap ? counts[q] : cap;
  int np2 = 2;
  while ((uint32_t)np2 < n) np2 <<= 1;
  for (uint32_t t = lane; t < (uint32_t)np2; t += PANN_WAVE)
    K[t] = t < n ? vis[(uint64_t)q * cap + t] : ~0ull;
  wave_lds_sync();
  wave_bitonic_sort<false>(K, np2, lane);
  for (uint32_t t = lane; t < n; t += PANN_WAVE) {
    ids[(uint64_t)q * cap + t] = key_id(K[t]);
    dists[(uint64_t)q * cap + t] = key_dist(K[t]);
  }
}

// edgeRange::update_neighbors (graph.h:77-88) on the sentinel-padded rows
__global__ void k_update_neighbors(uint32_t* nbrs, uint32_t nbr_stride, uint32_t max_degree,
                                   const uint32_t* points, uint32_t np, const uint32_t* lists,
                                   const uint32_t* counts, uint32_t R) {
  uint32_t i = blockIdx.x;
  if (i >= np) return;
  uint32_t p = points[i];
  uint32_t c = counts[i];
  for (uint32_t t = threadIdx.x; t < max_degree; t += blockDim.x)
    nbrs[(uint64_t)p * nbr_stride + t] = t < c ? lists[(uint64_t)i * R + t] : PANN_SENTINEL;
}

uint32_t prune_lds_bytes(uint32_t sort_cap, uint32_t dim, int dtype) {
  const uint32_t row = ((dim * (uint32_t)store_elem_bytes(dtype) + 127u) >> 7) * 128u;
  return 2u * row + 12u * (uint32_t)PRUNE_BLK + (sort_cap ? sort_cap * 9u : 0u);
}

hipError_t launch_robust_prune(const PruneArgs& a, int dtype, int metric, bool keys_global,
                               uint32_t grid, hipStream_t s) {
  uint32_t lds = prune_lds_bytes(keys_global ? 0 : a.sort_cap, a.dim, dtype);
  if (a.np == 0) return hipSuccess;
#define PANN_PRUNE(DT, MT)                                                                    \
  {                                                                                           \
    auto kern = keys_global ? k_robust_prune<DT, MT, true> : k_robust_prune<DT, MT, false>;   \
    hipError_t e = hipFuncSetAttribute((const void*)kern,                                     \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                            \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(PANN_WAVE), lds, s, a);                         \
    return hipGetLastError();                                                                 \
  }
  if (dtype == ST_F32) {
    if (metric == M_L2) PANN_PRUNE(ST_F32, M_L2) else PANN_PRUNE(ST_F32, M_IP)
  } else if (dtype == ST_U8) {
    if (metric == M_L2) PANN_PRUNE(ST_U8, M_L2) else PANN_PRUNE(ST_U8, M_IP)
  } else if (dtype == ST_I8) {
    if (metric == M_L2) PANN_PRUNE(ST_I8, M_L2) else PANN_PRUNE(ST_I8, M_IP)
  }
#undef PANN_PRUNE
  return hipErrorInvalidValue;
}

hipError_t launch_sort_visited(const uint64_t* vis, uint32_t nq, uint32_t cap,
                               const uint32_t* counts, uint32_t* ids, float* dists,
                               hipStream_t s) {
  uint32_t np2 = 2;
  while (np2 < cap) np2 <<= 1;
  uint32_t lds = np2 * 8;
  hipError_t e = hipFuncSetAttribute((const void*)k_sort_visited,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_visited, dim3(nq), dim3(PANN_WAVE), lds, s, vis, nq, cap, counts, ids,
                     dists);
  return hipGetLastError();
}

hipError_t launch_sum_u32(const uint32_t* v, uint32_t n, unsigned long long* acc, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t grid = std::min<uint32_t>((n + PANN_WAVE - 1) / PANN_WAVE, 1024u);
  hipLaunchKernelGGL(k_sum_u32, dim3(grid), dim3(PANN_WAVE), 0, s, v, n, acc);
  return hipGetLastError();
}

hipError_t launch_update_neighbors(uint32_t* nbrs, uint32_t nbr_stride, uint32_t max_degree,
                                   const uint32_t* points, uint32_t np, const uint32_t* lists,
                                   const uint32_t* counts, uint32_t R, hipStream_t s) {
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_update_neighbors, dim3(np), dim3(64), 0, s, nbrs, nbr_stride, max_degree,
                     points, np, lists, counts, R);
  return hipGetLastError();
}

}  // namespace pann
