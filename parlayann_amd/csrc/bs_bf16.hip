// bs_bf16.hip -- beam-search kernels of the BF16 store: one translation unit per
// store, so the build compiles the stores' kernels in parallel.
#include "beam_search_kernel.h"

namespace pann {

PANN_DEFINE_STORE_LAUNCH(launch_bs_bf16, ST_BF16, false)

}  // namespace pann
