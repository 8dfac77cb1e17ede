// wave_m4.hip -- instantiation of the wave-per-instance kernels for M = 4.
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<4, 4, 1>(const BatchDev &, int, hipStream_t);
template int occupancy_t<4, 4, 1>();
template hipError_t launch_solve_t<4, 6, 1>(const BatchDev &, int, hipStream_t);
template int occupancy_t<4, 6, 1>();
template hipError_t launch_component_t<4, 4>(const BatchDev &, int, int, hipStream_t);
template hipError_t launch_component_t<4, 6>(const BatchDev &, int, int, hipStream_t);
}  // namespace bvpdev
