// bs_i8.hip -- beam-search kernels of the int8 store (and the reference's 8-bit MIPS codes): one translation unit per
// store, so the build compiles the stores' kernels in parallel.
#include "beam_search_kernel.h"

namespace pann {

PANN_DEFINE_STORE_LAUNCH(launch_bs_i8, ST_I8, false)

}  // namespace pann
