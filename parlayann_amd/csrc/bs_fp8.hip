// bs_fp8.hip -- beam-search kernels of the FP8 e4m3 store: one translation unit per
// store, so the build compiles the stores' kernels in parallel.
#include "beam_search_kernel.h"

namespace pann {

PANN_DEFINE_STORE_LAUNCH(launch_bs_fp8, ST_FP8, false)

}  // namespace pann
