// wave_ge_m34.hip -- instantiation of the wave-per-instance solve under a global-error
// controller, and of the global-error estimate alone, for M = 3 and 4.
#include "mirk_wave_impl.h"

namespace bvpdev {
#define INST(m, me)                                                                   \
    template hipError_t launch_solve_ge_t<m, me>(const BatchDev &, int, hipStream_t); \
    template hipError_t launch_global_error_t<m, me>(const BatchDev &, int, hipStream_t);
INST(3, 4) INST(3, 6) INST(4, 4) INST(4, 6)
#undef INST
}  // namespace bvpdev
