// bvp_host.cpp -- host-side validation and packing of the C ABI's problem
// instances and options (bvp_host.h).
#include "bvp_host.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>

namespace bvphost {

namespace {

int fail(std::string &err, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
}

}  // namespace

int params_needed(int family, int M) {
    switch (family) {
    case BVP_FAMILY_LINEAR_TANH: return 1 + M * M;
    case BVP_FAMILY_SCALAR_TU: return 0;
    default: return 1;
    }
}

void uniform_mesh(double t0, double t1, int N, double *m) {
    for (int k = 0; k <= N; ++k)
        m[k] = (double)(((long double)t0 * (N - k) + (long double)t1 * k) / (long double)N);
    m[0] = t0;
    m[N] = t1;
}

int validate_instances(int32_t n, const bvp_instance *inst, int &M, SupportsFn supports, std::string &err) {
    if (n <= 0 || !inst) return fail(err, BVP_EINVAL, "empty batch");
    M = inst[0].M;
    for (int i = 0; i < n; ++i) {
        const bvp_instance &P = inst[i];
        if (P.M != M) return fail(err, BVP_EINVAL, "instance %d: all instances of a batch must share M", i);
        if (P.M < 1) return fail(err, BVP_EINVAL, "instance %d: M must be >= 1", i);
        if (!supports(M, P.family))
            return fail(err, BVP_ENOTSUP, "instance %d: family %d with M=%d is not supported", i, P.family, M);
        const int need = params_needed(P.family, M);
        if (P.n_params < need || P.n_params < 0)
            return fail(err, BVP_EINVAL, "instance %d: family %d needs %d parameters", i, P.family, need);
        if (P.n_params > 0 && !P.params) return fail(err, BVP_EINVAL, "instance %d: params is NULL", i);
        if (P.n_bc != M || !P.bc_t || !P.bc_comp || !P.bc_val)
            return fail(err, BVP_EINVAL, "instance %d: need exactly M boundary rows", i);
        if (P.bc_kind != BVP_BC_TWO_POINT && P.bc_kind != BVP_BC_STANDARD)
            return fail(err, BVP_EINVAL, "instance %d: bad bc_kind", i);
        if (!(P.t1 > P.t0) || !std::isfinite(P.t0) || !std::isfinite(P.t1))
            return fail(err, BVP_EINVAL, "instance %d: need finite t1 > t0", i);
        for (int k = 0; k < M; ++k) {
            if (P.bc_comp[k] < 0 || P.bc_comp[k] >= M)
                return fail(err, BVP_EINVAL, "instance %d: bc_comp out of range", i);
            if (P.bc_kind == BVP_BC_TWO_POINT && P.bc_t[k] != P.t0 && P.bc_t[k] != P.t1)
                return fail(err, BVP_EINVAL, "instance %d: two-point BC row %d not at t0 or t1", i, k);
            if (!(P.bc_t[k] >= P.t0 && P.bc_t[k] <= P.t1))
                return fail(err, BVP_EINVAL, "instance %d: BC point outside tspan", i);
        }
        if (P.N < 1) return fail(err, BVP_EINVAL, "instance %d: N must be >= 1", i);
        if (!P.guess && !P.u0) return fail(err, BVP_EINVAL, "instance %d: need a guess or u0", i);
        if (P.mesh) {
            for (int k = 0; k < P.N; ++k)
                if (!(P.mesh[k + 1] > P.mesh[k])) return fail(err, BVP_EINVAL, "instance %d: mesh not increasing", i);
        }
    }
    return BVP_OK;
}

int validate_opts(const bvp_opts *o, int32_t n, const bvp_instance *inst, int M, bool blk, std::string &err) {
    if (!o) return fail(err, BVP_EINVAL, "opts is NULL");
    if (o->method != BVP_MIRK4 && o->method != BVP_MIRK6) return fail(err, BVP_EINVAL, "method must be 4 or 6");
    if (o->precision != BVP_PREC_FP64 && o->precision != BVP_PREC_FP32_REFINE64)
        return fail(err, BVP_EINVAL, "precision must be BVP_PREC_FP64 or BVP_PREC_FP32_REFINE64");
    if (!(o->abstol > 0)) return fail(err, BVP_EINVAL, "abstol must be positive");
    if (o->max_num_subintervals < 1) return fail(err, BVP_EINVAL, "max_num_subintervals must be >= 1");
    if (o->max_newton_iters < 0) return fail(err, BVP_EINVAL, "max_newton_iters must be >= 0");
    if (o->refine_iters < 0) return fail(err, BVP_EINVAL, "refine_iters must be >= 0");
    if (o->transfer_mode != BVP_TRANSFER_REFERENCE && o->transfer_mode != BVP_TRANSFER_CLEAN)
        return fail(err, BVP_EINVAL, "transfer_mode must be BVP_TRANSFER_REFERENCE or BVP_TRANSFER_CLEAN");
    if (blk) {
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < M; ++k)
                if (inst[i].bc_t[k] != inst[i].t0 && inst[i].bc_t[k] != inst[i].t1)
                    return fail(err, BVP_ENOTSUP, "instance %d: for M=%d boundary rows must sit at t0 or t1", i, M);
    }
    if (o->warm_start == BVP_WARM_MLP) {
        if (blk || M != 2) return fail(err, BVP_ENOTSUP, "the learned warm start serves M = 2 problems");
        for (int i = 0; i < n; ++i) {
            if (!inst[i].u0) return fail(err, BVP_EINVAL, "warm start: every instance needs u0 (the fallback guess)");
            if (!inst[i].guess) return fail(err, BVP_EINVAL, "warm start: instance %d has no guess", i);
        }
    } else if (o->warm_start != BVP_WARM_NONE) {
        return fail(err, BVP_EINVAL, "unknown warm_start");
    }
    if (o->nlsolve_stages < 1 || o->nlsolve_stages > 7)
        return fail(err, BVP_EINVAL, "nlsolve_stages must be a non-empty mask of bits 0 (NewtonRaphson), "
                                     "1 (BackTracking), 2 (TrustRegion)");
    if (blk && !(o->nlsolve_stages & 3))
        return fail(err, BVP_ENOTSUP, "the block kernels (M >= 8) run the NewtonRaphson stages (bits 0 and 1) only");
    if (o->controller < BVP_CONTROL_DEFECT || o->controller > BVP_CONTROL_HYBRID)
        return fail(err, BVP_EINVAL, "controller must be one of BVP_CONTROL_*");
    if (o->controller != BVP_CONTROL_DEFECT) {
        if (blk || M > 4) return fail(err, BVP_ENOTSUP, "the global-error controllers run on the M <= 4 kernels");
        if (o->global_method != BVP_GLOBAL_HO && o->global_method != BVP_GLOBAL_RE)
            return fail(err, BVP_EINVAL, "global_method must be BVP_GLOBAL_HO or BVP_GLOBAL_RE");
        if (o->global_method == BVP_GLOBAL_HO && o->method != BVP_MIRK4)
            return fail(err, BVP_ENOTSUP, "HOErrorControl needs MIRK(p+2): MIRK4 only (no MIRK8)");
    }
    return BVP_OK;
}

void pack_instances(int32_t n, const bvp_instance *inst, int M, Packed &p) {
    p.fam.assign(n, 0);
    p.poff.assign(n, 0);
    p.pcnt.assign(n, 0);
    p.bck.assign(n, 0);
    p.N0.assign(n, 0);
    p.bcc.assign((size_t)n * M, 0);
    p.order.assign(n, 0);
    p.t0.assign(n, 0.0);
    p.t1.assign(n, 0.0);
    p.params.clear();
    p.bct.assign((size_t)n * M, 0.0);
    p.bcv.assign((size_t)n * M, 0.0);
    p.off.assign(n, 0);
    p.nodes = 0;
    for (int i = 0; i < n; ++i) {
        const bvp_instance &P = inst[i];
        p.fam[i] = P.family;
        p.t0[i] = P.t0;
        p.t1[i] = P.t1;
        p.poff[i] = (int)p.params.size();
        p.pcnt[i] = P.n_params;
        for (int k = 0; k < P.n_params; ++k) p.params.push_back(P.params[k]);
        p.bck[i] = P.bc_kind;
        for (int k = 0; k < M; ++k) {
            p.bct[(size_t)i * M + k] = P.bc_t[k];
            p.bcc[(size_t)i * M + k] = P.bc_comp[k];
            p.bcv[(size_t)i * M + k] = P.bc_val[k];
        }
        p.N0[i] = P.N;
        p.off[i] = p.nodes;
        p.nodes += P.N + 1;
    }
    p.mesh.assign((size_t)p.nodes, 0.0);
    p.y.assign((size_t)p.nodes * M, 0.0);
    p.u0c.assign((size_t)n * M, 0.0);
    for (int i = 0; i < n; ++i) {
        const bvp_instance &P = inst[i];
        if (P.u0)
            for (int c = 0; c < M; ++c) p.u0c[(size_t)i * M + c] = P.u0[c];
        double *m = &p.mesh[(size_t)p.off[i]];
        if (P.mesh) std::memcpy(m, P.mesh, sizeof(double) * (P.N + 1));
        else uniform_mesh(P.t0, P.t1, P.N, m);
        double *y = &p.y[(size_t)p.off[i] * M];
        if (P.guess) std::memcpy(y, P.guess, sizeof(double) * (size_t)M * (P.N + 1));
        else
            for (int k = 0; k <= P.N; ++k)
                for (int c = 0; c < M; ++c) y[(size_t)k * M + c] = P.u0[c];
    }
    // queue order: largest meshes first (longest-processing-time first)
    std::iota(p.order.begin(), p.order.end(), 0);
    std::stable_sort(p.order.begin(), p.order.end(), [&](int a, int b) { return p.N0[a] > p.N0[b]; });
}

}  // namespace bvphost
