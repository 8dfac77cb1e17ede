// wave_m2.hip -- instantiation of the wave-per-instance kernels for M = 2 (the
// solve kernel variants at 2 and 3 waves per SIMD are in wave_m2v.hip).
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<2, 4, 1>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 1>();
template hipError_t launch_solve_t<2, 6, 1>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 6, 1>();
template hipError_t launch_component_t<2, 4>(const BatchDev &, int, int, hipStream_t);
template hipError_t launch_component_t<2, 6>(const BatchDev &, int, int, hipStream_t);
}  // namespace bvpdev
