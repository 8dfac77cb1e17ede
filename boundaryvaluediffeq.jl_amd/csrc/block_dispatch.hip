// block_dispatch.hip -- dispatch of the block kernels over (M, method, precision);
// the kernels are instantiated in block_m8/16/32.hip.
#include "mirk_block_impl.h"

namespace bvpdev {
#define DECL(m, me, LT)                                                                             \
    extern template hipError_t launch_bsolve_t<m, me, LT>(const BatchDev &, int, hipStream_t);        \
    extern template hipError_t launch_bcomponent_t<m, me, LT>(const BatchDev &, int, int, hipStream_t); \
    extern template int boccupancy_t<m, me, LT>();
#define DECL_M(m) DECL(m, 4, double) DECL(m, 6, double) DECL(m, 4, float) DECL(m, 6, float)
DECL_M(8) DECL_M(16) DECL_M(32)
#undef DECL_M
#undef DECL
}  // namespace bvpdev

using namespace bvpdev;

#define BLOCK_DISPATCH(M_, METHOD_, PREC_, CALL)                                  \
    do {                                                                           \
        const bool six_ = (METHOD_) == 6, f32_ = (PREC_) == BVP_PREC_FP32_REFINE64; \
        switch (M_) {                                                              \
        case 8:                                                                    \
            if (six_) { if (f32_) CALL(8, 6, float); else CALL(8, 6, double); }    \
            else { if (f32_) CALL(8, 4, float); else CALL(8, 4, double); }         \
            break;                                                                 \
        case 16:                                                                   \
            if (six_) { if (f32_) CALL(16, 6, float); else CALL(16, 6, double); }  \
            else { if (f32_) CALL(16, 4, float); else CALL(16, 4, double); }       \
            break;                                                                 \
        case 32:                                                                   \
            if (six_) { if (f32_) CALL(32, 6, float); else CALL(32, 6, double); }  \
            else { if (f32_) CALL(32, 4, float); else CALL(32, 4, double); }       \
            break;                                                                 \
        }                                                                          \
    } while (0)

bool block_supports(int M, int family) {
    return (M == 8 || M == 16 || M == 32) && family == BVP_FAMILY_LINEAR_TANH;
}

hipError_t block_launch_solve(const BatchDev &B, int method, int precision, int blocks, hipStream_t s) {
#define C1(m, me, LT) return launch_bsolve_t<m, me, LT>(B, blocks, s)
    BLOCK_DISPATCH(B.M, method, precision, C1);
#undef C1
    return hipErrorInvalidValue;
}

hipError_t block_launch_component(const BatchDev &B, int method, int precision, int op, int blocks, hipStream_t s) {
#define C2(m, me, LT) return launch_bcomponent_t<m, me, LT>(B, op, blocks, s)
    BLOCK_DISPATCH(B.M, method, precision, C2);
#undef C2
    return hipErrorInvalidValue;
}

int block_blocks_per_cu(int M, int method, int precision) {
#define C3(m, me, LT) return boccupancy_t<m, me, LT>()
    BLOCK_DISPATCH(M, method, precision, C3);
#undef C3
    return 1;
}

size_t block_record_values(int M) {
    switch (M) {
    case 8: return BRec<8>::SIZE;
    case 16: return BRec<16>::SIZE;
    case 32: return BRec<32>::SIZE;
    }
    return 0;
}
