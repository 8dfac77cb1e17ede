// wave_m2_singpert.hip -- the M = 2 MIRK4 persistent solve kernel specialised to
// batches of the singpert family only (its three register-allocation variants).
#include "mirk_wave_impl.h"

namespace bvpdev {
template hipError_t launch_solve_t<2, 4, 1, BVP_FAMILY_SINGPERT>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 1, BVP_FAMILY_SINGPERT>();
template hipError_t launch_solve_t<2, 4, 2, BVP_FAMILY_SINGPERT>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 2, BVP_FAMILY_SINGPERT>();
template hipError_t launch_solve_t<2, 4, 3, BVP_FAMILY_SINGPERT>(const BatchDev &, int, hipStream_t);
template int occupancy_t<2, 4, 3, BVP_FAMILY_SINGPERT>();
}  // namespace bvpdev
