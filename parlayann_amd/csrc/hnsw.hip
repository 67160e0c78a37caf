// hnsw.hip -- the last step of HNSW::search (HNSW.hpp:1568-1613) on the device.
//
// The layer-0 beam search leaves each query's frontier sorted by (dist, id)
// in ids/dists with its size in count.  HNSW::search sorts it by distance and,
// when it holds more than k entries, cuts it after the last entry whose
// distance equals the k-th one (std::upper_bound with `farthest`, :1590-1596);
// result ids are the points' ids (U::get_id of the node's data, :1608-1609).
#include "kernels.h"

namespace pann {

__global__ void k_hnsw_cut(uint32_t* ids, float* dists, uint32_t stride, uint32_t* count,
                           uint32_t nq, uint32_t k_all, const int32_t* k_pq,
                           const uint32_t* node_ids) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const uint32_t k = k_pq ? (uint32_t)k_pq[q] : k_all;
  uint32_t* I = ids + (uint64_t)q * stride;
  const float* D = dists + (uint64_t)q * stride;
  uint32_t m = count[q] < stride ? count[q] : stride;
  if (m > k) {
    uint32_t kk = k;
    if (k > 0) {
      const float dk = D[k - 1];
      while (kk < m && !(dk < D[kk])) kk++;
    }
    m = kk;
  }
  count[q] = m;
  if (node_ids)
    for (uint32_t i = 0; i < m; i++) I[i] = node_ids[I[i]];
}

hipError_t launch_hnsw_cut(uint32_t* ids, float* dists, uint32_t stride, uint32_t* count,
                           uint32_t nq, uint32_t k, const int32_t* k_pq,
                           const uint32_t* node_ids, hipStream_t s) {
  hipLaunchKernelGGL(k_hnsw_cut, dim3((nq + 255) / 256), dim3(256), 0, s, ids, dists, stride,
                     count, nq, k, k_pq, node_ids);
  return hipGetLastError();
}

}  // namespace pann
