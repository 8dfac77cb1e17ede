// wave_f32_m2v.hip -- the M = 2 persistent solve kernels with FP32 storage of
// the chunk elimination factors, register allocation held to 2 and 3 waves per
// SIMD (variants the autotuner measures).
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<2, 4, 2, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 2, -1, true>();
template hipError_t launch_solve_t<2, 4, 3, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 3, -1, true>();
template hipError_t launch_solve_t<2, 6, 2, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 6, 2, -1, true>();
template hipError_t launch_solve_t<2, 6, 3, -1, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 6, 3, -1, true>();
}  // namespace bvpdev
