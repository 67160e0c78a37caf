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
#include "kernels.h"
#include "dist.h"

namespace pann {

// distance between two stored rows, the reference's arithmetic
template <int DT, int MT>
__device__ __forceinline__ float row_dist(const char* qrow_lds, const char* row, int dim) {
  if constexpr (DT == ST_F32) {
    return seq_f32_row<MT>((const float*)qrow_lds, (const float*)row, dim);
  } else {
    int32_t acc = 0;
    for (int j = 0; j < dim; j++) {
      int32_t a, b;
      if constexpr (DT == ST_U8) {
        a = (int32_t)((const uint8_t*)qrow_lds)[j];
        b = (int32_t)((const uint8_t*)row)[j];
      } else {
        a = (int32_t)((const int8_t*)qrow_lds)[j];
        b = (int32_t)((const int8_t*)row)[j];
      }
      if constexpr (MT == M_L2) acc += (a - b) * (a - b);
      else acc += a * b;
    }
    return MT == M_L2 ? (float)acc : -((float)acc);
  }
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



template <int DT, int MT, bool KG>
__device__ void prune_one(const PruneArgs& a, uint32_t i, int lane, char* QP, char* QS,
                          uint64_t* K, uint8_t* PR) {
  const int esz = store_elem_bytes(DT);
  const uint32_t p = a.points[i];

  const char* prow = a.pts + (uint64_t)p * a.stride;
  for (uint32_t j = lane; j < a.dim * esz; j += PANN_WAVE) QP[j] = prow[j];
  wave_sync_k<KG>();

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
  for (uint32_t t = lane; t < nc; t += PANN_WAVE) {
    uint32_t c = cid[t];
    float d = cdist ? cdist[t] : row_dist<DT, MT>(QP, a.pts + (uint64_t)c * a.stride, (int)a.dim);
    K[t] = make_key(d, c);
  }
  uint32_t total = nc;
  if (a.add != 0 && a.add != 4) {  // :72-77 (robustPrune's add, also the reverse-edge rules)
    const uint32_t* row = a.nbrs + (uint64_t)p * a.nbr_stride;
    for (uint32_t t = lane; t < a.max_degree; t += PANN_WAVE) {
      uint32_t c = row[t];
      if (c != PANN_SENTINEL)
        K[nc + t] = make_key(row_dist<DT, MT>(QP, a.pts + (uint64_t)c * a.stride, (int)a.dim), c);
    }
    uint32_t deg = 0;
    for (uint32_t t = 0; t < a.max_degree; t += PANN_WAVE) {
      uint32_t tt = t + lane;
      deg += __popcll(ballot(tt < a.max_degree && row[tt] != PANN_SENTINEL));
    }
    total = nc + deg;
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
    const char* srow = a.pts + (uint64_t)ps * a.stride;
    for (uint32_t j = lane; j < a.dim * esz; j += PANN_WAVE) QS[j] = srow[j];
    wave_sync_k<KG>();
    for (uint32_t t = idx + 1 + lane; t < total; t += PANN_WAVE) {
      if (!PR[t]) {
        uint64_t kk = K[t];
        float d_sp = row_dist<DT, MT>(QS, a.pts + (uint64_t)key_id(kk) * a.stride, (int)a.dim);
        if (hnsw) {
          const float d_cu = key_dist(kk) * alpha_f;  // HNSW.hpp:460-470
          if (d_sp < d_cu) PR[t] = 1;
        } else if (a.alpha * (double)d_sp <= (double)key_dist(kk)) {
          PR[t] = 1;
        }
      }
    }
    wave_sync_k<KG>();
  }
  if (lane == 0) a.out_count[i] = nout;
}


// KG: sort keys and pruned flags live in a per-block global scratch (groups too
// large for LDS); otherwise in LDS.  Blocks loop over points.
template <int DT, int MT, bool KG>
__global__ __launch_bounds__(PANN_WAVE) void k_robust_prune(PruneArgs a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int lane = threadIdx.x;
  const int esz = store_elem_bytes(DT);
  char* QP = smem;                                             // row of p
  char* QS = QP + ((a.dim * esz + 15) & ~15u);                 // row of p*
  char* tail = QS + ((a.dim * esz + 15) & ~15u);
  uint64_t* K = KG ? a.gkeys + (uint64_t)blockIdx.x * a.sort_cap : (uint64_t*)tail;
  uint8_t* PR = KG ? a.gflags + (uint64_t)blockIdx.x * a.sort_cap
                   : (uint8_t*)(tail + a.sort_cap * 8);
  for (uint32_t i = blockIdx.x; i < a.np; i += gridDim.x) {
    prune_one<DT, MT, KG>(a, i, lane, QP, QS, K, PR);
  }
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
  uint32_t row = (dim * (uint32_t)store_elem_bytes(dtype) + 15u) & ~15u;
  return 2u * row + (sort_cap ? sort_cap * 9u : 0u);
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

hipError_t launch_update_neighbors(uint32_t* nbrs, uint32_t nbr_stride, uint32_t max_degree,
                                   const uint32_t* points, uint32_t np, const uint32_t* lists,
                                   const uint32_t* counts, uint32_t R, hipStream_t s) {
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_update_neighbors, dim3(np), dim3(64), 0, s, nbrs, nbr_stride, max_degree,
                     points, np, lists, counts, R);
  return hipGetLastError();
}

}  // namespace pann
