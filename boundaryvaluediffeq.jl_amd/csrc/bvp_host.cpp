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
    if (is_firk(o->method)) {
        FirkTableau T;
        if (!firk_tableau(o->method, T))
            return fail(err, BVP_EINVAL, "unknown FIRK method %d (BVP_RADAU_IIA(1, 2, 3, 5, 7), "
                                         "BVP_LOBATTO_IIIA/B/C(2..5))", o->method);
        if (blk || (M > 4 && M != 6))
            return fail(err, BVP_ENOTSUP, "FIRK methods run on the wave kernels (M = 1..4, 6)");
        if (o->adaptive && !T.adaptive)
            return fail(err, BVP_ENOTSUP, "RadauIIa1, LobattoIIIb2 and LobattoIIIc2 do not adapt the mesh: "
                                          "set adaptive = 0");
        if (o->controller != BVP_CONTROL_DEFECT)
            return fail(err, BVP_ENOTSUP, "FIRK methods use DefectControl");
        if (o->warm_start != BVP_WARM_NONE) return fail(err, BVP_ENOTSUP, "no learned warm start for FIRK methods");
        if (!(o->nlsolve_stages & 7) || (o->nlsolve_stages & ~7))
            return fail(err, BVP_EINVAL, "nlsolve_stages must be a non-empty subset of bits 0-2");
        if (o->nlsolve_stages & BVP_STAGE_TRUST_REGION) {
            if (o->method >= BVP_FIRK_NESTED)
                return fail(err, BVP_ENOTSUP, "the TrustRegion stage runs on the expanded FIRK form only");
        }
    } else if (o->method != BVP_MIRK4 && o->method != BVP_MIRK6) {
        return fail(err, BVP_EINVAL, "method must be 4, 6 or a FIRK method code");
    }
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

namespace {

using ld = long double;

// Legendre P_n at x by the three-term recurrence; *dp receives P_n'(x) (|x| < 1)
ld legendre_p(int n, ld x, ld *dp) {
    ld pm = 1, p = x;
    if (n == 0) {
        if (dp) *dp = 0;
        return 1;
    }
    for (int k = 2; k <= n; ++k) {
        const ld pn = ((2 * k - 1) * x * p - (k - 1) * pm) / k;
        pm = p;
        p = pn;
    }
    if (dp) *dp = n * (pm - x * p) / (1 - x * x);
    return p;
}

// the roots of g in (-1, 1): sign changes on a grid of 8192 cells, refined by bisection
template <class G>
int roots(G &&g, ld *out, int max_out) {
    const int cells = 8192;
    int n = 0;
    ld a = -1 + (ld)2 / cells, ga = g(a);
    for (int k = 2; k < cells && n < max_out; ++k) {
        const ld b = -1 + (ld)2 * k / cells, gb = g(b);
        if (ga == 0) {
            out[n++] = a;
        } else if (gb != 0 && ((ga > 0) != (gb > 0))) {
            ld lo = a, hi = b, glo = ga;
            for (int it = 0; it < 128; ++it) {
                const ld m = (lo + hi) / 2, gm = g(m);
                if ((gm > 0) == (glo > 0)) { lo = m; glo = gm; } else hi = m;
            }
            out[n++] = (lo + hi) / 2;
        }
        a = b;
        ga = gb;
    }
    return n;
}

// x solves V x = r (n x n, row-major), Gaussian elimination with partial pivoting
void solve_ld(int n, ld *V, ld *r) {
    for (int j = 0; j < n; ++j) {
        int p = j;
        for (int i = j + 1; i < n; ++i) if (std::fabs(V[i * n + j]) > std::fabs(V[p * n + j])) p = i;
        for (int c = 0; c < n; ++c) std::swap(V[j * n + c], V[p * n + c]);
        std::swap(r[j], r[p]);
        for (int i = j + 1; i < n; ++i) {
            const ld m = V[i * n + j] / V[j * n + j];
            for (int c = j; c < n; ++c) V[i * n + c] -= m * V[j * n + c];
            r[i] -= m * r[j];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        ld v = r[i];
        for (int c = i + 1; c < n; ++c) v -= V[i * n + c] * r[c];
        r[i] = v / V[i * n + i];
    }
}

}  // namespace

// Nodes: Radau IIA the zeros of P_s(2t-1) - P_{s-1}(2t-1) (t = 1 among them), Lobatto
// 0, 1 and the zeros of P'_{s-1}(2t-1).  Radau IIA and Lobatto IIIA collocate:
// A_ij = int_0^{c_i} l_j, b_j = int_0^1 l_j, with the monomial coefficients of each
// Lagrange basis polynomial l_j expanded from its product form.  Lobatto IIIB:
// A_ij = b_j (1 - A^IIIA_ji / b_i); Lobatto IIIC: A_i1 = b_1 and the simplifying
// conditions sum_j A_ij c_j^(k-1) = c_i^k / k for k < s.  The reference tabulates the
// same methods (FIRK/radau_tableaus.jl, FIRK/lobatto_tableaus.jl), tau* included.
bool firk_tableau(int code, FirkTableau &T) {
    T = FirkTableau();
    if (code >= 2 * BVP_FIRK_NESTED) return false;
    const int base = code % BVP_FIRK_NESTED;  // the nested flag selects the form, not the tableau
    const int kind = base / 100, s = base % 100;
    if (kind == 1) {
        if (s != 1 && s != 2 && s != 3 && s != 5 && s != 7) return false;
        T.order = 2 * s - 1;
        T.adaptive = s != 1;
    } else if (kind >= 2 && kind <= 4) {
        if (s < 2 || s > 5) return false;
        T.order = 2 * s - 2;
        T.adaptive = !(s == 2 && kind != 2);
    } else {
        return false;
    }
    T.s = s;
    ld x[7];
    int nx = 0;
    if (kind == 1) {
        nx = roots([&](ld v) { return legendre_p(s, v, nullptr) - legendre_p(s - 1, v, nullptr); }, x, 6);
        x[nx++] = 1;
    } else {
        x[nx++] = -1;
        if (s > 2) {
            nx += roots([&](ld v) {
                ld d;
                legendre_p(s - 1, v, &d);
                return d;
            }, x + 1, 5);
        }
        x[nx++] = 1;
    }
    if (nx != s) return false;
    ld c[7], l[7][7] = {}, Acol[7][7], bcol[7];
    for (int i = 0; i < s; ++i) c[i] = (x[i] + 1) / 2;
    for (int j = 0; j < s; ++j) {
        // l_j = prod_{m != j} (t - c_m) / (c_j - c_m), coefficients l[j][k] of t^k
        ld poly[8] = {1};
        int deg = 0;
        for (int m = 0; m < s; ++m) {
            if (m == j) continue;
            const ld den = c[j] - c[m];
            for (int k = deg + 1; k >= 0; --k) poly[k] = ((k > 0 ? poly[k - 1] : 0) - c[m] * poly[k]) / den;
            ++deg;
        }
        for (int k = 0; k < s; ++k) l[j][k] = poly[k];
        bcol[j] = 0;
        for (int k = 0; k < s; ++k) bcol[j] += l[j][k] / (k + 1);
        for (int i = 0; i < s; ++i) {
            ld a = 0, p = c[i];
            for (int k = 0; k < s; ++k, p *= c[i]) a += l[j][k] * p / (k + 1);
            Acol[i][j] = a;
        }
    }
    ld A[7][7];
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j)
            A[i][j] = kind == 3 ? bcol[j] * (1 - Acol[j][i] / bcol[i]) : Acol[i][j];
    if (kind == 4) {
        const int n = s - 1;
        for (int i = 0; i < s; ++i) {
            ld V[36], r[6];
            for (int k = 0; k < n; ++k) {
                for (int j = 1; j < s; ++j) {
                    ld p = 1;
                    for (int q = 0; q < k; ++q) p *= c[j];
                    V[k * n + (j - 1)] = p;
                }
                ld ck = c[i];
                for (int q = 0; q < k; ++q) ck *= c[i];
                r[k] = ck / (k + 1) - (k == 0 ? bcol[0] : 0);
            }
            solve_ld(n, V, r);
            A[i][0] = bcol[0];
            for (int j = 1; j < s; ++j) A[i][j] = r[j - 1];
        }
    }
    for (int i = 0; i < s; ++i) {
        T.c[i] = (double)c[i];
        T.b[i] = (double)bcol[i];
        for (int j = 0; j < s; ++j) {
            T.A[i][j] = (double)A[i][j];
            T.Q[i][j] = (double)(l[j][i] / (i + 1));
        }
    }
    // tau* of each table (FIRK/radau_tableaus.jl: 0 for every Radau IIA method;
    // FIRK/lobatto_tableaus.jl: per stage count, IIIC at 1 - tau* of IIIA/IIIB for s = 3, 5)
    if (kind == 1) T.tau = 0.0;
    else if (s == 2 || s == 4) T.tau = 0.5;
    else if (s == 3) T.tau = kind == 4 ? 0.7886751345948129 : 0.21132486540518713;
    else T.tau = kind == 4 ? 0.6699905217924309 : 0.33000947820757126;
    return true;
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
