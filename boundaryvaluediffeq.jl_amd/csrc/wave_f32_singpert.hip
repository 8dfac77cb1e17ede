// wave_f32_singpert.hip -- the M = 2 MIRK4 persistent solve kernel specialised
// to batches of the singular-perturbation family only, with the chunk
// elimination factors stored in FP32 (mixed precision); its three
// register-allocation variants.
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<2, 4, 1, BVP_FAMILY_SINGPERT, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 1, BVP_FAMILY_SINGPERT, true>();
template hipError_t launch_solve_t<2, 4, 2, BVP_FAMILY_SINGPERT, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 2, BVP_FAMILY_SINGPERT, true>();
template hipError_t launch_solve_t<2, 4, 3, BVP_FAMILY_SINGPERT, true>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 3, BVP_FAMILY_SINGPERT, true>();
}  // namespace bvpdev
