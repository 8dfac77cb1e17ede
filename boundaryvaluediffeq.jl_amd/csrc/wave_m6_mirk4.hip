// wave_m6_mirk4.hip -- instantiation of the wave-per-instance kernels for M = 6,
// MIRK4 (FP64 records; the component kernels of the parity entry points).
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<6, 4, 1>(const BatchDev &, int, hipStream_t);
template int occupancy_t<6, 4, 1>();
template hipError_t launch_component_t<6, 4>(const BatchDev &, int, int, hipStream_t);
}  // namespace bvpdev
