// wave_f32_m134.hip -- the M = 1, 3, 4 persistent solve kernels with the chunk
// elimination factors stored in FP32 (mixed precision).
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<1, 4, 1, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<1, 4, 1, -1, true>();
template hipError_t launch_solve_t<1, 6, 1, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<1, 6, 1, -1, true>();
template hipError_t launch_solve_t<3, 4, 1, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<3, 4, 1, -1, true>();
template hipError_t launch_solve_t<3, 6, 1, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<3, 6, 1, -1, true>();
template hipError_t launch_solve_t<4, 4, 1, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<4, 4, 1, -1, true>();
template hipError_t launch_solve_t<4, 6, 1, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<4, 6, 1, -1, true>();
}  // namespace bvpdev
