// wave_ge_m12.hip -- instantiation of the wave-per-instance solve under a global-error
// controller, and of the global-error estimate alone, for M = 1 and 2.
#include "mirk_wave_impl.h"

namespace bvpdev {
#define INST(m, me)                                                                   \
    template hipError_t launch_solve_ge_t<m, me>(const BatchDev &, int, hipStream_t); \
    template hipError_t launch_global_error_t<m, me>(const BatchDev &, int, hipStream_t);
INST(1, 4) INST(1, 6) INST(2, 4) INST(2, 6)
#undef INST
}  // namespace bvpdev
