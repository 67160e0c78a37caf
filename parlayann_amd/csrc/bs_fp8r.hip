// bs_fp8r.hip -- beam-search kernels of the FP8 residual store (inner
// product, rows of 1 to 3 lines): one translation unit per store, so the build
// compiles the stores' kernels in parallel.
#include "beam_search_kernel.h"

namespace pann {

hipError_t launch_bs_fp8r(const SearchArgs& a, int metric, int nch, uint32_t grid,
                          hipStream_t s) {
  (void)metric;
  switch (nch) {
    case 1: return launch_t<ST_FP8R, M_IP, 1, false>(a, grid, s);
    case 2: return launch_t<ST_FP8R, M_IP, 2, false>(a, grid, s);
    case 3: return launch_t<ST_FP8R, M_IP, 3, false>(a, grid, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pann
