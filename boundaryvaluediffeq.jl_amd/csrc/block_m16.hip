// block_m16.hip -- instantiation of the block kernels for M = 16.
#include "mirk_block_impl.h"

namespace bvpdev {
#define INST(me, LT)                                                                          \
    template hipError_t launch_bsolve_t<16, me, LT>(const BatchDev &, int, hipStream_t);       \
    template hipError_t launch_bcomponent_t<16, me, LT>(const BatchDev &, int, int, hipStream_t); \
    template int boccupancy_t<16, me, LT>();
INST(4, double) INST(6, double) INST(4, float) INST(6, float)
#undef INST
}  // namespace bvpdev
