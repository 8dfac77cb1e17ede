// wave_f32_m2.hip -- the M = 2 persistent solve kernels with the chunk
// elimination factors stored in FP32 (mixed precision), MIRK4 and MIRK6 (the
// variants at 2 and 3 waves per SIMD are in wave_f32_m2v.hip).
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<2, 4, 1, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 1, -1, true>();
template hipError_t launch_solve_t<2, 6, 1, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 6, 1, -1, true>();
}  // namespace bvpdev
