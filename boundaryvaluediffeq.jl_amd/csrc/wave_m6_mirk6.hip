// wave_m6_mirk6.hip -- instantiation of the wave-per-instance kernels for M = 6,
// MIRK6 (FP64 records; the component kernels of the parity entry points).
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<6, 6, 1>(const BatchDev &, int, hipStream_t);
template int occupancy_t<6, 6, 1>();
template hipError_t launch_component_t<6, 6>(const BatchDev &, int, int, hipStream_t);
}  // namespace bvpdev
