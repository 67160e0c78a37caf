// beam_search.hip -- host-side launch of the beam-search kernels
// (beam_search_kernel.h), whose instantiations are spread over one
// translation unit per store (bs_*.hip).
#include "beam_search_kernel.h"

namespace pann {

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

uint32_t search_lds_bytes(const SearchArgs& a) {
  return wave_lds_bytes(a.fcap, a.ccap, a.dim, a.hslots_log2, a.rc_log2, a.fb_log2, a.ncent,
                        a.q_own, a.wpq);
}

static int nch_for(int row_bytes) {
  int c = (row_bytes + 127) / 128;
  if (c <= 1) return 1;
  if (c <= 2) return 2;
  if (c <= 3) return 3;
  if (c <= 4) return 4;
  if (c <= 8) return 8;
  return -1;
}

hipError_t launch_beam_search(const SearchArgs& a, int store, int metric, bool sequential,
                              uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, a.work_counter, 1u);
  int nch = nch_for((int)a.row_bytes);
  if (sequential) {
    if (store != ST_F32) return hipErrorInvalidValue;
    if (nch < 0) nch = 8;  // rows past 8 lines: 8 lines at a time, query in LDS
    return launch_bs_f32_inorder(a, metric, nch, grid, s);
  }
  switch (store) {
    case ST_F32: return launch_bs_f32(a, metric, nch, grid, s);
    case ST_U8: return launch_bs_u8(a, metric, nch, grid, s);
    case ST_I8: return launch_bs_i8(a, metric, nch, grid, s);
    case ST_BF16: return launch_bs_bf16(a, metric, nch, grid, s);
    case ST_FP8: return launch_bs_fp8(a, metric, nch, grid, s);
    case ST_FP8R:  // rows of dim + 4 <= 384 bytes, inner product only
      if (metric != M_IP || !a.qtable || a.ncent == 0) return hipErrorInvalidValue;
      return launch_bs_fp8r(a, metric, nch, grid, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pann
