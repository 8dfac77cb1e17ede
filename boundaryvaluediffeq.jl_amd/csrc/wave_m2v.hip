// wave_m2v.hip -- the M = 2 persistent solve kernel with its register allocation
// held to 2 and 3 waves per SIMD (variants the autotuner measures).
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<2, 4, 2>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 2>();
template hipError_t launch_solve_t<2, 4, 3>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 3>();
template hipError_t launch_solve_t<2, 6, 2>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 6, 2>();
template hipError_t launch_solve_t<2, 6, 3>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 6, 3>();
}  // namespace bvpdev
