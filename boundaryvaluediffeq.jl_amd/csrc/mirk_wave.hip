// mirk_wave.hip -- dispatch of the wave-per-instance kernels over (M, method);
// the kernels themselves are instantiated in wave_m1..wave_m4.hip.
#include "mirk_wave_impl.h"

namespace bvpdev {
#define DECL(m, me)                                                                              \
    extern template hipError_t launch_component_t<m, me>(const BatchDev &, int, int, hipStream_t); \
    extern template hipError_t launch_solve_ge_t<m, me>(const BatchDev &, int, hipStream_t);      \
    extern template hipError_t launch_global_error_t<m, me>(const BatchDev &, int, hipStream_t);
#define DECLV(m, me, w, f32)                                                                     \
    extern template hipError_t launch_solve_t<m, me, w, -1, f32>(const BatchDev &, int, hipStream_t); \
    extern template int occupancy_t<m, me, w, -1, f32>();
#define DECLP(f32)                                                                               \
    DECLV(1, 4, 1, f32) DECLV(1, 6, 1, f32) DECLV(3, 4, 1, f32) DECLV(3, 6, 1, f32)             \
    DECLV(4, 4, 1, f32) DECLV(4, 6, 1, f32) DECLV(2, 4, 1, f32) DECLV(2, 6, 1, f32)             \
    DECLV(2, 4, 2, f32) DECLV(2, 6, 2, f32) DECLV(2, 4, 3, f32) DECLV(2, 6, 3, f32)
DECL(1, 4) DECL(1, 6) DECL(2, 4) DECL(2, 6) DECL(3, 4) DECL(3, 6) DECL(4, 4) DECL(4, 6)
DECLP(false) DECLP(true)
// M = 6: FP64 records and the component kernels only (wave_m6_mirk4/6.hip)
#define DECL6(me)                                                                                \
    extern template hipError_t launch_component_t<6, me>(const BatchDev &, int, int, hipStream_t); \
    DECLV(6, me, 1, false)
DECL6(4) DECL6(6)
#undef DECL6
#define DECLF(w, f, f32)                                                                         \
    extern template hipError_t launch_solve_t<2, 4, w, f, f32>(const BatchDev &, int, hipStream_t); \
    extern template int occupancy_t<2, 4, w, f, f32>();
#define DECLFP(f32)                                                                              \
    DECLF(1, BVP_FAMILY_BRATU, f32) DECLF(2, BVP_FAMILY_BRATU, f32) DECLF(3, BVP_FAMILY_BRATU, f32) \
    DECLF(1, BVP_FAMILY_SINGPERT, f32) DECLF(2, BVP_FAMILY_SINGPERT, f32) DECLF(3, BVP_FAMILY_SINGPERT, f32)
DECLFP(false) DECLFP(true)
#undef DECL
#undef DECLV
#undef DECLP
#undef DECLF
#undef DECLFP
}  // namespace bvpdev
using namespace bvpdev;

__global__ void k_fill_u32(unsigned *p, unsigned v) { *p = v; }

// Plain streaming copy of n doubles: the calibration reference for HBM byte
// counters (it reads 8n and writes 8n bytes).
__global__ __launch_bounds__(256) void k_copy_probe(const double *__restrict__ src, double *__restrict__ dst,
                                                    size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy_probe(const double *src, double *dst, size_t n, hipStream_t s) {
    k_copy_probe<<<4096, 256, 0, s>>>(src, dst, n);
    return hipGetLastError();
}

hipError_t wave_reset_counter(unsigned *counter, hipStream_t s) {
    k_fill_u32<<<1, 1, 0, s>>>(counter, 0u);
    return hipGetLastError();
}

#define DISPATCH_M_METHOD(M_, METHOD_, CALL)                                                  \
    do {                                                                                       \
        if ((M_) == 1 && (METHOD_) == 4) { CALL(1, 4); }                                       \
        else if ((M_) == 1 && (METHOD_) == 6) { CALL(1, 6); }                                  \
        else if ((M_) == 2 && (METHOD_) == 4) { CALL(2, 4); }                                  \
        else if ((M_) == 2 && (METHOD_) == 6) { CALL(2, 6); }                                  \
        else if ((M_) == 3 && (METHOD_) == 4) { CALL(3, 4); }                                  \
        else if ((M_) == 3 && (METHOD_) == 6) { CALL(3, 6); }                                  \
        else if ((M_) == 4 && (METHOD_) == 4) { CALL(4, 4); }                                  \
        else if ((M_) == 4 && (METHOD_) == 6) { CALL(4, 6); }                                  \
        else return hipErrorInvalidValue;                                                      \
    } while (0)

int wave_variants(int M) { return M == 2 ? 3 : 1; }

static bool specialised(int M, int method, int fam) {
    return M == 2 && method == 4 && (fam == BVP_FAMILY_BRATU || fam == BVP_FAMILY_SINGPERT);
}

// F32: the chunk factors are stored in FP32 (BatchDev::f32rec); each storage
// precision is its own kernel, with its own register allocation.
template <bool F32>
static hipError_t launch_solve_p(const BatchDev &B, int method, int variant, int blocks, int fam, hipStream_t s) {
    if (specialised(B.M, method, fam)) {
        const int w = variant + 1;
        if (fam == BVP_FAMILY_BRATU)
            return w == 1 ? launch_solve_t<2, 4, 1, BVP_FAMILY_BRATU, F32>(B, blocks, s)
                 : w == 2 ? launch_solve_t<2, 4, 2, BVP_FAMILY_BRATU, F32>(B, blocks, s)
                          : launch_solve_t<2, 4, 3, BVP_FAMILY_BRATU, F32>(B, blocks, s);
        return w == 1 ? launch_solve_t<2, 4, 1, BVP_FAMILY_SINGPERT, F32>(B, blocks, s)
             : w == 2 ? launch_solve_t<2, 4, 2, BVP_FAMILY_SINGPERT, F32>(B, blocks, s)
                      : launch_solve_t<2, 4, 3, BVP_FAMILY_SINGPERT, F32>(B, blocks, s);
    }
    if (B.M == 6) {
        if constexpr (F32) return hipErrorInvalidValue;  // FP64 records only at M = 6
        else return method == 6 ? launch_solve_t<6, 6, 1, -1, false>(B, blocks, s) : launch_solve_t<6, 4, 1, -1, false>(B, blocks, s);
    }
    if (B.M == 2 && variant == 1)
        return method == 6 ? launch_solve_t<2, 6, 2, -1, F32>(B, blocks, s) : launch_solve_t<2, 4, 2, -1, F32>(B, blocks, s);
    if (B.M == 2 && variant == 2)
        return method == 6 ? launch_solve_t<2, 6, 3, -1, F32>(B, blocks, s) : launch_solve_t<2, 4, 3, -1, F32>(B, blocks, s);
#define CALL_SOLVE(m, me) return launch_solve_t<m, me, 1, -1, F32>(B, blocks, s)
    DISPATCH_M_METHOD(B.M, method, CALL_SOLVE);
#undef CALL_SOLVE
}

hipError_t wave_launch_solve(const BatchDev &B, int method, int variant, int blocks, int fam, hipStream_t s) {
    return B.f32rec ? launch_solve_p<true>(B, method, variant, blocks, fam, s)
                    : launch_solve_p<false>(B, method, variant, blocks, fam, s);
}

hipError_t wave_launch_component(const BatchDev &B, int method, int op, int blocks, hipStream_t s) {
    if (B.M == 6) return method == 6 ? launch_component_t<6, 6>(B, op, blocks, s) : launch_component_t<6, 4>(B, op, blocks, s);
#define CALL_COMP(m, me) return launch_component_t<m, me>(B, op, blocks, s)
    DISPATCH_M_METHOD(B.M, method, CALL_COMP);
#undef CALL_COMP
}

hipError_t wave_launch_solve_ge(const BatchDev &B, int method, int blocks, hipStream_t s) {
#define CALL_GE(m, me) return launch_solve_ge_t<m, me>(B, blocks, s)
    DISPATCH_M_METHOD(B.M, method, CALL_GE);
#undef CALL_GE
}

hipError_t wave_launch_global_error(const BatchDev &B, int method, int blocks, hipStream_t s) {
#define CALL_GEST(m, me) return launch_global_error_t<m, me>(B, blocks, s)
    DISPATCH_M_METHOD(B.M, method, CALL_GEST);
#undef CALL_GEST
}

hipError_t wave_launch_gather(const BatchDev &B, const long long *out_off, double *out_mesh, double *out_y,
                              int max_nodes, hipStream_t s) {
    (void)max_nodes;
    dim3 grid(B.n);
    switch (B.M) {
    case 1: k_gather<1><<<grid, 256, 0, s>>>(B, out_off, out_mesh, out_y); break;
    case 2: k_gather<2><<<grid, 256, 0, s>>>(B, out_off, out_mesh, out_y); break;
    case 3: k_gather<3><<<grid, 256, 0, s>>>(B, out_off, out_mesh, out_y); break;
    case 4: k_gather<4><<<grid, 256, 0, s>>>(B, out_off, out_mesh, out_y); break;
    case 6: k_gather<6><<<grid, 256, 0, s>>>(B, out_off, out_mesh, out_y); break;
    case 8: k_gather<8><<<grid, 256, 0, s>>>(B, out_off, out_mesh, out_y); break;
    case 16: k_gather<16><<<grid, 256, 0, s>>>(B, out_off, out_mesh, out_y); break;
    case 32: k_gather<32><<<grid, 256, 0, s>>>(B, out_off, out_mesh, out_y); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t wave_launch_interp(const BatchDev &B, int method, const int *qinst, const double *tq, double *u,
                              double *du, long long nq, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    const dim3 grid((unsigned)((nq + 255) / 256));
    if (B.M == 6) {
        if (method == 6) k_interp<6, 6><<<grid, 256, 0, s>>>(B, qinst, tq, u, du, nq);
        else k_interp<6, 4><<<grid, 256, 0, s>>>(B, qinst, tq, u, du, nq);
        return hipGetLastError();
    }
#define CALL_INTERP(m, me)                                                      \
    k_interp<m, me><<<grid, 256, 0, s>>>(B, qinst, tq, u, du, nq);              \
    return hipGetLastError()
    DISPATCH_M_METHOD(B.M, method, CALL_INTERP);
#undef CALL_INTERP
}

bool wave_supports(int M, int family) {
    switch (M) {
    case 1: return family_supported<1>(family);
    case 2: return family_supported<2>(family);
    case 3: return family_supported<3>(family);
    case 4: return family_supported<4>(family);
    case 6: return family_supported<6>(family);
    }
    return false;
}

int wave_rel_fields(int M) { return 2 * M * M + M; }


template <bool F32>
static int blocks_per_cu_p(int M, int method, int variant, int fam) {
    if (specialised(M, method, fam)) {
        const int w = variant + 1;
        if (fam == BVP_FAMILY_BRATU)
            return w == 1 ? occupancy_t<2, 4, 1, BVP_FAMILY_BRATU, F32>()
                 : w == 2 ? occupancy_t<2, 4, 2, BVP_FAMILY_BRATU, F32>()
                          : occupancy_t<2, 4, 3, BVP_FAMILY_BRATU, F32>();
        return w == 1 ? occupancy_t<2, 4, 1, BVP_FAMILY_SINGPERT, F32>()
             : w == 2 ? occupancy_t<2, 4, 2, BVP_FAMILY_SINGPERT, F32>()
                      : occupancy_t<2, 4, 3, BVP_FAMILY_SINGPERT, F32>();
    }
    const bool six = (method == 6);
    if (M == 2 && variant == 1) return six ? occupancy_t<2, 6, 2, -1, F32>() : occupancy_t<2, 4, 2, -1, F32>();
    if (M == 2 && variant == 2) return six ? occupancy_t<2, 6, 3, -1, F32>() : occupancy_t<2, 4, 3, -1, F32>();
    switch (M) {
    case 1: return six ? occupancy_t<1, 6, 1, -1, F32>() : occupancy_t<1, 4, 1, -1, F32>();
    case 2: return six ? occupancy_t<2, 6, 1, -1, F32>() : occupancy_t<2, 4, 1, -1, F32>();
    case 3: return six ? occupancy_t<3, 6, 1, -1, F32>() : occupancy_t<3, 4, 1, -1, F32>();
    case 4: return six ? occupancy_t<4, 6, 1, -1, F32>() : occupancy_t<4, 4, 1, -1, F32>();
    case 6: return six ? occupancy_t<6, 6, 1, -1, false>() : occupancy_t<6, 4, 1, -1, false>();
    }
    return 4;
}

int wave_blocks_per_cu(int M, int method, int variant, int fam, int f32) {
    return f32 ? blocks_per_cu_p<true>(M, method, variant, fam) : blocks_per_cu_p<false>(M, method, variant, fam);
}

bool blockm_supports(int M, int family) { return block_supports(M, family); }
