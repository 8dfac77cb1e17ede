// firk_m6a.hip -- the FIRK collocation kernels (expanded and nested forms) for state dimension
// M = 6, stage counts 1-3 (firk_m6b.hip: 4, 5, 7; two files so that they compile in parallel).
#include "firk_wave_impl.h"

namespace bvpdev {
#define FIRK_INST(S)                                                                                         \
    template hipError_t firk_launch_solve_t<6, S, 1>(const BatchDev &, int, hipStream_t);                   \
    template hipError_t firk_launch_solve_t<6, S, 2>(const BatchDev &, int, hipStream_t);                   \
    template hipError_t firk_launch_component_t<6, S>(const BatchDev &, int, int, hipStream_t);             \
    template hipError_t firk_launch_interp_t<6, S>(const BatchDev &, const int *, const double *, double *,  \
                                                     double *, long long, hipStream_t);                       \
    template hipError_t firk_launch_gather_t<6, S>(const BatchDev &, const long long *, double *, hipStream_t); \
    template int firk_occupancy_t<6, S, 1>(int);                                                              \
    template int firk_occupancy_t<6, S, 2>(int);
FIRK_INST(1)
FIRK_INST(2)
FIRK_INST(3)
#undef FIRK_INST
}  // namespace bvpdev
