// firk_dispatch.hip -- run-time dispatch of the FIRK kernels on (M, stages).
#include "firk_wave_impl.h"

namespace bvpdev {
#define FIRK_DECL(M, S)                                                                                      \
    extern template hipError_t firk_launch_solve_t<M, S, 1>(const BatchDev &, int, hipStream_t);             \
    extern template hipError_t firk_launch_solve_t<M, S, 2>(const BatchDev &, int, hipStream_t);             \
    extern template hipError_t firk_launch_component_t<M, S>(const BatchDev &, int, int, hipStream_t);       \
    extern template hipError_t firk_launch_interp_t<M, S>(const BatchDev &, const int *, const double *,      \
                                                           double *, double *, long long, hipStream_t);       \
    extern template hipError_t firk_launch_gather_t<M, S>(const BatchDev &, const long long *, double *,      \
                                                           hipStream_t);                                      \
    extern template int firk_occupancy_t<M, S, 1>(int);                                                         \
    extern template int firk_occupancy_t<M, S, 2>(int);
#define FIRK_DECL_M(M) FIRK_DECL(M, 1) FIRK_DECL(M, 2) FIRK_DECL(M, 3) FIRK_DECL(M, 4) FIRK_DECL(M, 5) FIRK_DECL(M, 7)
FIRK_DECL_M(1)
FIRK_DECL_M(2)
FIRK_DECL_M(3)
FIRK_DECL_M(4)
FIRK_DECL_M(6)
#undef FIRK_DECL_M
#undef FIRK_DECL
}  // namespace bvpdev

using namespace bvpdev;

bool firk_supports(int M, int S) {
    return ((M >= 1 && M <= 4) || M == 6) && (S == 1 || S == 2 || S == 3 || S == 4 || S == 5 || S == 7);
}

// one case per (M, S): CALL is the launcher call with MM, SS in scope
#define FIRK_SWITCH(CALL)                                                                       \
    switch (B.M * 16 + S) {                                                                     \
    FIRK_CASES_M(1, CALL) FIRK_CASES_M(2, CALL) FIRK_CASES_M(3, CALL) FIRK_CASES_M(4, CALL)   \
    FIRK_CASES_M(6, CALL)                                                                       \
    default: return hipErrorInvalidValue;                                                       \
    }
#define FIRK_CASE(MM_, SS_, CALL)   \
    case MM_ * 16 + SS_: {          \
        constexpr int MM = MM_;     \
        constexpr int SS = SS_;     \
        return CALL;                \
    }
#define FIRK_CASES_M(MM_, CALL)                                                                      \
    FIRK_CASE(MM_, 1, CALL) FIRK_CASE(MM_, 2, CALL) FIRK_CASE(MM_, 3, CALL) FIRK_CASE(MM_, 4, CALL) \
    FIRK_CASE(MM_, 5, CALL) FIRK_CASE(MM_, 7, CALL)

hipError_t firk_launch_solve(const BatchDev &B, int S, int variant, int blocks, hipStream_t s) {
    if (variant == 1) {
        FIRK_SWITCH((firk_launch_solve_t<MM, SS, 2>(B, blocks, s)))
    }
    FIRK_SWITCH((firk_launch_solve_t<MM, SS, 1>(B, blocks, s)))
}
hipError_t firk_launch_component(const BatchDev &B, int S, int op, int blocks, hipStream_t s) {
    FIRK_SWITCH((firk_launch_component_t<MM, SS>(B, op, blocks, s)))
}
hipError_t firk_launch_interp(const BatchDev &B, int S, const int *qinst, const double *tq, double *u, double *du,
                              long long nq, hipStream_t s) {
    FIRK_SWITCH((firk_launch_interp_t<MM, SS>(B, qinst, tq, u, du, nq, s)))
}
hipError_t firk_launch_gather(const BatchDev &B, int S, const long long *out_off, double *out, hipStream_t s) {
    FIRK_SWITCH((firk_launch_gather_t<MM, SS>(B, out_off, out, s)))
}
int firk_blocks_per_cu(int M, int S, int variant, int nested) {
#define FIRK_OCC(MM, SS)                                                                   \
    case MM * 16 + SS:                                                                     \
        return variant == 1 ? firk_occupancy_t<MM, SS, 2>(nested) : firk_occupancy_t<MM, SS, 1>(nested);
#define FIRK_OCC_M(MM) FIRK_OCC(MM, 1) FIRK_OCC(MM, 2) FIRK_OCC(MM, 3) FIRK_OCC(MM, 4) FIRK_OCC(MM, 5) FIRK_OCC(MM, 7)
    switch (M * 16 + S) {
        FIRK_OCC_M(1) FIRK_OCC_M(2) FIRK_OCC_M(3) FIRK_OCC_M(4) FIRK_OCC_M(6)
    default: return 4;
    }
#undef FIRK_OCC_M
#undef FIRK_OCC
}
