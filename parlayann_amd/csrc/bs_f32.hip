// bs_f32.hip -- beam-search kernels of the FP32 store, tree (DPP) summation: one translation unit per
// store, so the build compiles the stores' kernels in parallel.
#include "beam_search_kernel.h"

namespace pann {

PANN_DEFINE_STORE_LAUNCH(launch_bs_f32, ST_F32, false)

}  // namespace pann
