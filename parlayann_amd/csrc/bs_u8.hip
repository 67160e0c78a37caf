// bs_u8.hip -- beam-search kernels of the uint8 store (and the reference's 8-bit L2 codes): one translation unit per
// store, so the build compiles the stores' kernels in parallel.
#include "beam_search_kernel.h"

namespace pann {

PANN_DEFINE_STORE_LAUNCH(launch_bs_u8, ST_U8, false)

}  // namespace pann
