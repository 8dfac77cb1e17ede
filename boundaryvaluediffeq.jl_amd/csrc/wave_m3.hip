// wave_m3.hip -- instantiation of the wave-per-instance kernels for M = 3.
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<3, 4, 1>(const BatchDev &, int, hipStream_t);
template int occupancy_t<3, 4, 1>();
template hipError_t launch_solve_t<3, 6, 1>(const BatchDev &, int, hipStream_t);
template int occupancy_t<3, 6, 1>();
template hipError_t launch_component_t<3, 4>(const BatchDev &, int, int, hipStream_t);
template hipError_t launch_component_t<3, 6>(const BatchDev &, int, int, hipStream_t);
}  // namespace bvpdev
