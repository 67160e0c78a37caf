// bs_f32_inorder.hip -- beam-search kernels of the FP32 store, in-order summation (the reference's arithmetic): one translation unit per
// store, so the build compiles the stores' kernels in parallel.
#include "beam_search_kernel.h"

namespace pann {

PANN_DEFINE_STORE_LAUNCH(launch_bs_f32_inorder, ST_F32, true)

}  // namespace pann
