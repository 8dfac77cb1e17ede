// sanitize_main.cpp -- drives every entry point of the oracle (refcpu.cpp) once
// per problem family and option combination, and with invalid inputs, so that a
// build under -fsanitize=address,undefined checks the oracle's memory accesses and
// arithmetic.  TEST INFRASTRUCTURE ONLY (built by `make sanitize`, run by
// tests/test_oracle_sanitize.py).  Exit status 0 when every call returned what it
// should; the sanitizers abort on the first finding.
#include "../include/bvp_mirk.h"

#include <cmath>
#include <cstdio>
#include <vector>

extern "C" {
void ref_opts_default(bvp_opts *o);
int ref_solve(const bvp_opts *o, const bvp_instance *P, bvp_result *res);
int ref_residual(int method, const bvp_instance *P, double *resid);
int ref_jacobian_blocks(int method, const bvp_instance *P, double *S, double *R, int32_t *bc_node);
int ref_newton_step(int method, const bvp_instance *P, double *delta);
int ref_defect(int method, const bvp_instance *P, double *errors, double *defect_norm);
int ref_interpolate(int method, const bvp_instance *P, int nt, const double *t, double *u, double *du);
int ref_global_error(const bvp_opts *o, const bvp_instance *P, double *errors, double *norm);
int ref_mesh_select(const bvp_opts *o, const bvp_instance *P, int capacity, int32_t *new_N, int32_t *info,
                    double *new_mesh, double *new_y);
int ref_backtracking_poly(const double *coef, int ncoef, double *alpha, double *trace, int max_trace,
                          int *n_trace);
int ref_dogleg(int n, const double *J, const double *F, double radius, double *d, double *pred);
int ref_firk_tableau(int code, int *ints, double *dbl);
int ref_firk_residual(int code, const bvp_instance *P, const double *stages, double *resid);
int ref_firk_dense_system(int code, const bvp_instance *P, const double *stages, double *J, double *F);
int ref_firk_newton_step(int code, const bvp_instance *P, const double *stages, double *delta);
int ref_firk_defect(int code, const bvp_instance *P, const double *stages, double *errors, double *defect_norm);
int ref_firk_interpolate(int code, const bvp_instance *P, const double *stages, int nt, const double *t, double *u,
                         double *du);
int ref_firk_evalsol(int code, const bvp_instance *P, double tj, double tj1, double t, int comp, const double *yi,
                     const double *yj, const double *K, double *value, double *g);
}

namespace {

struct Problem {
    bvp_instance inst{};
    std::vector<double> params, bct, bcv, u0;
    std::vector<int32_t> bcc;
    void fix() {
        inst.params = params.data();
        inst.n_params = (int32_t)params.size();
        inst.bc_t = bct.data();
        inst.bc_comp = bcc.data();
        inst.bc_val = bcv.data();
        inst.n_bc = (int32_t)bct.size();
        inst.u0 = u0.data();
        inst.mesh = nullptr;
        inst.guess = nullptr;
    }
};

Problem make(int family, int M, double t0, double t1, std::vector<double> p, int kind,
             std::vector<double> bct, std::vector<int32_t> bcc, std::vector<double> bcv, int N) {
    Problem q;
    q.inst.family = family;
    q.inst.M = M;
    q.inst.t0 = t0;
    q.inst.t1 = t1;
    q.inst.bc_kind = kind;
    q.inst.N = N;
    q.params = std::move(p);
    q.bct = std::move(bct);
    q.bcc = std::move(bcc);
    q.bcv = std::move(bcv);
    q.u0.assign(M, 0.0);
    q.fix();
    return q;
}

int failures = 0;
void expect(bool ok, const char *what) {
    if (!ok) {
        std::printf("FAILED: %s\n", what);
        ++failures;
    }
}

void exercise(Problem &q, const char *name) {
    const bvp_instance &P = q.inst;
    const int M = P.M, N = P.N;
    for (int method : {BVP_MIRK4, BVP_MIRK6}) {
        std::vector<double> r((size_t)M * (N + 1)), S((size_t)N * M * M), R((size_t)N * M * M), d((size_t)M * (N + 1));
        std::vector<int32_t> node(M);
        expect(ref_residual(method, &P, r.data()) == 0, name);
        expect(ref_jacobian_blocks(method, &P, S.data(), R.data(), node.data()) == 0, name);
        ref_newton_step(method, &P, d.data());  // may report a singular matrix at the zero guess
        std::vector<double> err((size_t)N * M);
        double dn = 0;
        expect(ref_defect(method, &P, err.data(), &dn) == 0, name);
        const double tq[4] = {P.t0 - 0.1, P.t0, 0.5 * (P.t0 + P.t1), P.t1 + 0.1};
        std::vector<double> u(4 * M), du(4 * M);
        expect(ref_interpolate(method, &P, 4, tq, u.data(), du.data()) == 0, name);
        bvp_opts o;
        ref_opts_default(&o);
        o.method = method;
        o.abstol = 1e-6;
        o.max_num_subintervals = 400;
        const int cap = 401;
        int32_t nN = 0, info = 0;
        std::vector<double> nm(cap), ny((size_t)cap * M);
        expect(ref_mesh_select(&o, &P, cap, &nN, &info, nm.data(), ny.data()) == 0, name);
        for (int controller = 0; controller < 4; ++controller)
            for (int gm = 0; gm < 2; ++gm)
                for (int stages : {1, 2, 4, 7}) {
                    if (controller == 0 && gm == 1) continue;
                    o.controller = controller;
                    o.global_method = gm;
                    o.nlsolve_stages = stages;
                    o.max_newton_iters = 50;
                    std::vector<double> rm(cap), ru((size_t)cap * M);
                    bvp_result res{};
                    res.capacity = cap;
                    res.mesh = rm.data();
                    res.u = ru.data();
                    const int rc = ref_solve(&o, &P, &res);
                    const bool ho6 = controller != 0 && gm == BVP_GLOBAL_HO && method == BVP_MIRK6;
                    expect(ho6 ? rc == BVP_ENOTSUP : rc == 0, name);
                    if (controller == 1 && stages == 7 && !ho6) {
                        std::vector<double> ge((size_t)N * M);
                        double gn = 0;
                        expect(ref_global_error(&o, &P, ge.data(), &gn) == 0, name);
                    }
                }
    }
}

// the FIRK entry points (expanded and nested forms) for every method
void exercise_firk(Problem &q, const char *name) {
    const bvp_instance &P = q.inst;
    const int M = P.M, N = P.N;
    const int bases[] = {101, 102, 103, 105, 107, 202, 203, 204, 205, 302, 303, 304, 305, 402, 403, 404, 405};
    for (int nest : {0, 1000})
    for (int base : bases) {
        const int code = base + nest;
        int ints[3];
        double tab[113];
        expect(ref_firk_tableau(code, ints, tab) == 0, name);
        const int s = ints[0];
        const size_t nx = nest ? (size_t)M * (N + 1) : (size_t)M * ((size_t)N * (s + 1) + 1);
        std::vector<double> K((size_t)N * s * M, 0.1), F(nx), J(nx * nx), d(nx), err((size_t)N * M);
        expect(ref_firk_residual(code, &P, nullptr, F.data()) == 0, name);
        expect(ref_firk_residual(code, &P, K.data(), F.data()) == 0, name);
        expect(ref_firk_dense_system(code, &P, K.data(), J.data(), F.data()) == 0, name);
        ref_firk_newton_step(code, &P, K.data(), d.data());  // may be singular at this state
        double dn = 0;
        expect(ref_firk_defect(code, &P, K.data(), err.data(), &dn) == 0, name);
        const double tq[3] = {P.t0 - 0.1, 0.5 * (P.t0 + P.t1), P.t1 + 0.1};
        std::vector<double> u(3 * M), du(3 * M);
        expect(ref_firk_interpolate(code, &P, K.data(), 3, tq, u.data(), du.data()) == 0, name);
        bvp_opts o;
        ref_opts_default(&o);
        o.method = code;
        o.abstol = 1e-5;
        o.max_num_subintervals = 200;
        o.adaptive = ints[2];
        const int cap = 201;
        for (int stages : {1, 3, 4, 7}) {
            o.nlsolve_stages = stages;
            o.max_newton_iters = 30;
            std::vector<double> rm(cap), ru((size_t)cap * M), rs((size_t)cap * s * M);
            bvp_result res{};
            res.capacity = cap;
            res.mesh = rm.data();
            res.u = ru.data();
            res.stages = rs.data();
            expect(ref_solve(&o, &P, &res) == 0, name);
        }
    }
    int ints[3];
    double tab[113];
    expect(ref_firk_tableau(106, ints, tab) == BVP_EINVAL, "no RadauIIa6");
    expect(ref_firk_tableau(206, ints, tab) == BVP_EINVAL, "no LobattoIIIa6");
}

}  // namespace

int main() {
    const double pi = 3.14159265358979323846;
    std::vector<Problem> ps;
    ps.push_back(make(BVP_FAMILY_BRATU, 2, 0, 1, {2.0}, BVP_BC_TWO_POINT, {0, 1}, {0, 0}, {0, 0}, 12));
    ps.push_back(make(BVP_FAMILY_PENDULUM, 2, 0, pi / 2, {9.81}, BVP_BC_STANDARD, {pi / 4, pi / 2}, {0, 0},
                      {-pi / 2, pi / 2}, 16));
    ps[1].u0 = {pi / 2, pi / 2};
    ps[1].fix();
    ps.push_back(make(BVP_FAMILY_SINGPERT, 2, -1, 1, {0.05}, BVP_BC_TWO_POINT, {-1, 1}, {0, 0}, {1, 2}, 16));
    ps.push_back(make(BVP_FAMILY_TURNING, 2, -1, 1, {0.1}, BVP_BC_STANDARD, {-1, 1}, {0, 0}, {-2, 0}, 20));
    ps.push_back(make(BVP_FAMILY_SCALAR_TU, 1, 0, 1, {}, BVP_BC_STANDARD, {0.5}, {0}, {1.0}, 9));
    ps.push_back(make(BVP_FAMILY_LINEAR_TANH, 3, 0, 1, {0.3, 0.1, 1, 0, -1, 0.2, 0.5, 0.3, -0.4, -0.1},
                      BVP_BC_TWO_POINT, {0, 0, 1}, {0, 2, 1}, {1, -1, 0.5}, 7));
    ps.push_back(make(BVP_FAMILY_SWIRL, 6, 0, 1, {0.05}, BVP_BC_STANDARD, {0, 0, 0, 1, 1, 1}, {0, 2, 3, 0, 2, 3},
                      {-1, 0, 0, 1, 0, 0}, 10));
    const char *names[] = {"bratu", "pendulum", "singpert", "turning", "scalar_tu", "linear_tanh3", "swirl"};
    for (size_t i = 0; i < ps.size(); ++i) exercise(ps[i], names[i]);
    exercise_firk(ps[0], "firk bratu");
    Problem tu0 = make(BVP_FAMILY_SCALAR_TU, 1, 0, 1, {}, BVP_BC_STANDARD, {0.0}, {0}, {1.0}, 9);
    exercise_firk(tu0, "firk scalar_tu");
    exercise_firk(ps[4], "firk scalar_tu interior row");
    exercise_firk(ps[1], "firk pendulum interior row");
    exercise_firk(ps[5], "firk linear_tanh3");
    {
        double v, g[16];
        const double yi[2] = {0.1, 0.2}, yj[2] = {0.3, 0.4}, K[10] = {0};
        expect(ref_firk_evalsol(105, &ps[1].inst, 0.5, 0.6, 0.55, 1, yi, yj, K, &v, g) == 0, "firk evalsol");
        expect(ref_firk_evalsol(105, &ps[1].inst, 0.5, 0.6, 0.55, 2, yi, yj, K, &v, g) == BVP_EINVAL, "evalsol comp");
    }

    // invalid inputs: error codes, no out-of-bounds access
    Problem bad = ps[0];
    bad.inst.N = 0;
    bvp_opts o;
    ref_opts_default(&o);
    std::vector<double> m(8), u(16);
    bvp_result res{};
    res.capacity = 8;
    res.mesh = m.data();
    res.u = u.data();
    expect(ref_solve(&o, &bad.inst, &res) == BVP_EINVAL, "N = 0");
    bad = ps[0];
    bad.inst.family = 99;
    expect(ref_solve(&o, &bad.inst, &res) == BVP_EINVAL, "unknown family");
    bad = ps[0];
    bad.inst.n_bc = 1;
    expect(ref_solve(&o, &bad.inst, &res) == BVP_EINVAL, "too few BC rows");
    res.capacity = 2;  // smaller than the result
    expect(ref_solve(&o, &ps[0].inst, &res) == BVP_EINVAL, "capacity");

    // the nonlinear solver's building blocks on small inputs
    const double coef[3] = {1.0, -2.0, 1.5};
    double alpha = 0, trace[8];
    int nt = 0;
    expect(ref_backtracking_poly(coef, 3, &alpha, trace, 8, &nt) == 0, "backtracking");
    const double J[4] = {2, 1, 0, 3}, F[2] = {1, -1};
    double d[2], pred = 0;
    expect(ref_dogleg(2, J, F, 0.1, d, &pred) == 0, "dogleg");

    std::printf("%s: %d failures\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
