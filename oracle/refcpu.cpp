// refcpu.cpp -- CPU restatement of the reference MIRK collocation solve.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: the GPU library is
// checked against it, and the benchmark times it as the CPU baseline.  The
// product never links, loads or calls it.
//
// It restates, in plain FP64 C++, the algorithm of SciML/BoundaryValueDiffEq.jl
// on the MIRK path (paths relative to the reference root;
// MIRK/ = lib/BoundaryValueDiffEqMIRK/src/, Core/ = lib/BoundaryValueDiffEqCore/src/):
//   tableaus ............ MIRK/mirk_tableaus.jl:62-87, 120-152
//   interp weights ...... MIRK/interpolation.jl:481-499, 528-575
//   Phi (residual) ...... MIRK/collocation.jl:44-102
//   BC residual ......... Core/utils.jl:157-209; MIRK/interpolation.jl:214-242
//   residual layout ..... MIRK/mirk.jl:471-534; Core/utils.jl:28-52
//   Jacobian ............ exact chain rule for the blocks the reference gets
//                         from sparse ForwardDiff (MIRK/mirk.jl:697-862, 927-1002);
//                         BC rows with frozen stages (MIRK/mirk.jl:565-572)
//   linear solve ........ banded Gaussian elimination with partial pivoting
//                         (the reference hands the banded / almost-banded
//                         Jacobian to LinearSolve, Core/internal_problems.jl:107-169)
//   Newton .............. the default polyalgorithm NewtonRaphson -> NewtonRaphson
//                         + BackTracking -> TrustRegion (Core/default_internal_solve.jl:
//                         31-45), each stage from the same start; restated from the
//                         published algorithms of NonlinearSolveFirstOrder "1.2, 2" /
//                         LineSearch (BackTracking of order 3: Armijo 1e-4, quadratic
//                         then cubic steps clamped to [0.1, 0.5]) and a dogleg trust
//                         region ("Simple" radius update); stop when max|F| <= tol
//   dense output ........ MIRK/interpolation.jl:17-206 (sol(t), sol(t, Val{1}))
//   global error ........ MIRK/adaptivity.jl:77-243, 463-567; Core/calc_errors.jl:54-163
//   interp_setup! ....... MIRK/interpolation.jl:300-372
//   defect .............. MIRK/adaptivity.jl:370-461, 623-666
//   mesh_selector! ...... MIRK/adaptivity.jl:23-75
//   redistribute! ....... MIRK/adaptivity.jl:250-278
//   half_mesh! .......... MIRK/adaptivity.jl:287-304
//   transfer ............ MIRK/mirk.jl:361-373; MIRK/adaptivity.jl:6-13, 590-621
//   outer loop .......... MIRK/mirk.jl:286-388
//   mesh from dt ........ Core/utils.jl:354-362, 694-700
#include "../include/bvp_mirk.h"

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Tab {
    int s, sstar, order;
    double c[5], v[5], b[5], x[5][5];
    double cs[4], vs[4], xs[4][9];
    double tau;
};

// MIRK/mirk_tableaus.jl:62-87
Tab make_mirk4() {
    Tab t{};
    t.s = 3; t.sstar = 4; t.order = 4;
    t.c[0] = 0; t.c[1] = 1; t.c[2] = 1.0 / 2;
    t.v[0] = 0; t.v[1] = 1; t.v[2] = 1.0 / 2;
    t.b[0] = 1.0 / 6; t.b[1] = 1.0 / 6; t.b[2] = 2.0 / 3;
    t.x[2][0] = 1.0 / 8; t.x[2][1] = -1.0 / 8;
    t.cs[0] = 3.0 / 4; t.vs[0] = 27.0 / 32;
    t.xs[0][0] = 3.0 / 64; t.xs[0][1] = -9.0 / 64;
    t.tau = 0.226;
    return t;
}

// MIRK/mirk_tableaus.jl:120-152
Tab make_mirk6() {
    Tab t{};
    t.s = 5; t.sstar = 9; t.order = 6;
    const double c[5] = {0, 1, 1.0 / 4, 3.0 / 4, 1.0 / 2};
    const double v[5] = {0, 1, 5.0 / 32, 27.0 / 32, 1.0 / 2};
    const double b[5] = {7.0 / 90, 7.0 / 90, 16.0 / 45, 16.0 / 45, 2.0 / 15};
    for (int i = 0; i < 5; ++i) { t.c[i] = c[i]; t.v[i] = v[i]; t.b[i] = b[i]; }
    t.x[2][0] = 9.0 / 64;  t.x[2][1] = -3.0 / 64;
    t.x[3][0] = 3.0 / 64;  t.x[3][1] = -9.0 / 64;
    t.x[4][0] = -5.0 / 24; t.x[4][1] = 5.0 / 24; t.x[4][2] = 2.0 / 3; t.x[4][3] = -2.0 / 3;
    const double cs[4] = {7.0 / 16, 3.0 / 8, 9.0 / 16, 1.0 / 8};
    for (int i = 0; i < 4; ++i) { t.cs[i] = cs[i]; t.vs[i] = cs[i]; }
    const double xs[4][9] = {
        {1547.0 / 32768, -1225.0 / 32768, 749.0 / 4096, -287.0 / 2048, -861.0 / 16384, 0, 0, 0, 0},
        {83.0 / 1536, -13.0 / 384, 283.0 / 1536, -167.0 / 1536, -49.0 / 512, 0, 0, 0, 0},
        {1225.0 / 32768, -1547.0 / 32768, 287.0 / 2048, -749.0 / 4096, 861.0 / 16384, 0, 0, 0, 0},
        {233.0 / 3456, -19.0 / 1152, 0, 0, 0, -5.0 / 72, 7.0 / 72, -17.0 / 216, 0}};
    std::memcpy(t.xs, xs, sizeof(xs));
    t.tau = 0.7156;
    return t;
}

const Tab &tableau(int method) {
    static const Tab t4 = make_mirk4();
    static const Tab t6 = make_mirk6();
    return method == BVP_MIRK6 ? t6 : t4;
}

// Continuous-extension weights w(tau), w'(tau): MIRK/interpolation.jl:481-499 (MIRK4),
// :528-575 (MIRK6).  Rational coefficients are rounded to double once, as
// Rational * Float64 does in the reference.
void interp_weights(int method, double tau, double *w, double *wp) {
    if (method == BVP_MIRK4) {
        const double t2 = tau * tau, tm1 = tau - 1.0, t4m3 = tau * 4.0 - 3.0, t2m1 = tau * 2.0 - 1.0;
        w[0] = -tau * (2.0 * tau - 3.0) * (2.0 * t2 - 3.0 * tau + 2.0) / 6.0;
        w[1] = t2 * (12.0 * t2 - 20.0 * tau + 9.0) / 6.0;
        w[2] = 2.0 * t2 * (6.0 * t2 - 14.0 * tau + 9.0) / 3.0;
        w[3] = -16.0 * t2 * tm1 * tm1 / 3.0;
        wp[0] = -tm1 * t4m3 * t2m1 / 3.0;
        wp[1] = tau * t2m1 * t4m3;
        wp[2] = 4.0 * tau * t4m3 * tm1;
        wp[3] = -32.0 * tau * t2m1 * tm1 / 3.0;
        return;
    }
    const double t = tau, t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t, t6 = t5 * t;
    w[0] = t - 28607.0 / 7434 * t2 - 166210.0 / 33453 * t3 + 334780.0 / 11151 * t4 -
           1911296.0 / 55755 * t5 + 406528.0 / 33453 * t6;
    w[1] = 777.0 / 590 * t2 - 2534158.0 / 234171 * t3 + 2088580.0 / 78057 * t4 -
           10479104.0 / 390285 * t5 + 11328512.0 / 1170855 * t6;
    w[2] = -1008.0 / 59 * t2 + 222176.0 / 1593 * t3 - 180032.0 / 531 * t4 +
           876544.0 / 2655 * t5 - 180224.0 / 1593 * t6;
    w[3] = w[2];  // the reference lists the same polynomial for stages 3 and 4
    w[4] = -378.0 / 59 * t2 + 27772.0 / 531 * t3 - 22504.0 / 177 * t4 +
           109568.0 / 885 * t5 - 22528.0 / 531 * t6;
    w[5] = -95232.0 / 413 * t2 + 62384128.0 / 33453 * t3 - 49429504.0 / 11151 * t4 +
           46759936.0 / 11151 * t5 - 46661632.0 / 33453 * t6;
    w[6] = 896.0 / 5 * t2 - 4352.0 / 3 * t3 + 3456.0 * t4 - 16384.0 / 5 * t5 + 16384.0 / 15 * t6;
    w[7] = 50176.0 / 531 * t2 - 179554304.0 / 234171 * t3 + 143363072.0 / 78057 * t4 -
           136675328.0 / 78057 * t5 + 137363456.0 / 234171 * t6;
    w[8] = 16384.0 / 441 * t3 - 16384.0 / 147 * t4 + 16384.0 / 147 * t5 - 16384.0 / 441 * t6;
    wp[0] = 1 - 28607.0 / 3717 * t - 166210.0 / 11151 * t2 + 1339120.0 / 11151 * t3 -
            1911296.0 / 11151 * t4 + 813056.0 / 11151 * t5;
    wp[1] = 777.0 / 295 * t - 2534158.0 / 78057 * t2 + 8354320.0 / 78057 * t3 -
            10479104.0 / 78057 * t4 + 22657024.0 / 390285 * t5;
    wp[2] = -2016.0 / 59 * t + 222176.0 / 531 * t2 - 720128.0 / 531 * t3 +
            876544.0 / 531 * t4 - 360448.0 / 531 * t5;
    wp[3] = wp[2];
    wp[4] = -756.0 / 59 * t + 27772.0 / 177 * t2 - 90016.0 / 177 * t3 +
            109568.0 / 177 * t4 - 45056.0 / 177 * t5;
    wp[5] = -190464.0 / 413 * t + 62384128.0 / 11151 * t2 - 197718016.0 / 11151 * t3 +
            233799680.0 / 11151 * t4 - 93323264.0 / 11151 * t5;
    wp[6] = 1792.0 / 5 * t - 4352.0 * t2 + 13824.0 * t3 - 16384.0 * t4 + 32768.0 / 5 * t5;
    wp[7] = 100352.0 / 531 * t - 179554304.0 / 78057 * t2 + 573452288.0 / 78057 * t3 -
            683376640.0 / 78057 * t4 + 274726912.0 / 78057 * t5;
    wp[8] = 16384.0 / 147 * t2 - 65536.0 / 147 * t3 + 81920.0 / 147 * t4 - 32768.0 / 147 * t5;
}

// ---------------------------------------------------------------- families
void f_eval(const bvp_instance &P, double t, const double *y, double *out) {
    const double *p = P.params;
    switch (P.family) {
    case BVP_FAMILY_LINEAR_TANH: {
        const int M = P.M;
        const double c = p[0];
        const double *A = p + 1;
        for (int i = 0; i < M; ++i) {
            double acc = 0;
            for (int j = 0; j < M; ++j) acc += A[i * M + j] * y[j];
            out[i] = acc + c * std::tanh(y[i]);
        }
        break;
    }
    case BVP_FAMILY_PENDULUM:
        out[0] = y[1];
        out[1] = -p[0] * std::sin(y[0]);
        break;
    case BVP_FAMILY_BRATU:
        out[0] = y[1];
        out[1] = -p[0] * std::exp(y[0]);
        break;
    case BVP_FAMILY_SINGPERT:
        out[0] = y[1];
        out[1] = (y[0] - t * y[1]) / p[0];
        break;
    case BVP_FAMILY_SCALAR_TU:
        out[0] = t * y[0];
        break;
    case BVP_FAMILY_TURNING: {
        const double pi = 3.14159265358979323846;
        out[0] = y[1];
        out[1] = -t / p[0] * y[1] - pi * pi * std::cos(pi * t) - pi * t / p[0] * std::sin(pi * t);
        break;
    }
    case BVP_FAMILY_SWIRL: {
        // Swirling Flow III (MIRKt/Core/mirk_basic_tests.jl:318-327)
        const double e = p[0];
        out[0] = y[1];
        out[1] = (y[0] * y[3] - y[2] * y[1]) / e;
        out[2] = y[3];
        out[3] = y[4];
        out[4] = y[5];
        out[5] = (-y[2] * y[5] - y[0] * y[1]) / e;
        break;
    }
    }
}

// df/dy, M*M row-major (analytic, family by family)
void J_eval(const bvp_instance &P, double t, const double *y, double *J) {
    const double *p = P.params;
    const int M = P.M;
    switch (P.family) {
    case BVP_FAMILY_LINEAR_TANH: {
        const double c = p[0];
        for (int i = 0; i < M * M; ++i) J[i] = p[1 + i];
        for (int i = 0; i < M; ++i) {
            const double th = std::tanh(y[i]);
            J[i * M + i] += c * (1.0 - th * th);
        }
        break;
    }
    case BVP_FAMILY_PENDULUM:
        J[0] = 0; J[1] = 1; J[2] = -p[0] * std::cos(y[0]); J[3] = 0;
        break;
    case BVP_FAMILY_BRATU:
        J[0] = 0; J[1] = 1; J[2] = -p[0] * std::exp(y[0]); J[3] = 0;
        break;
    case BVP_FAMILY_SINGPERT:
        J[0] = 0; J[1] = 1; J[2] = 1.0 / p[0]; J[3] = -t / p[0];
        break;
    case BVP_FAMILY_SCALAR_TU:
        J[0] = t;
        break;
    case BVP_FAMILY_TURNING:
        J[0] = 0; J[1] = 1; J[2] = 0; J[3] = -t / p[0];
        break;
    case BVP_FAMILY_SWIRL: {
        const double e = p[0];
        for (int i = 0; i < 36; ++i) J[i] = 0.0;
        J[0 * 6 + 1] = 1.0;
        J[1 * 6 + 0] = y[3] / e; J[1 * 6 + 1] = -y[2] / e; J[1 * 6 + 2] = -y[1] / e; J[1 * 6 + 3] = y[0] / e;
        J[2 * 6 + 3] = 1.0;
        J[3 * 6 + 4] = 1.0;
        J[4 * 6 + 5] = 1.0;
        J[5 * 6 + 0] = -y[1] / e; J[5 * 6 + 1] = -y[0] / e; J[5 * 6 + 2] = -y[5] / e; J[5 * 6 + 5] = -y[2] / e;
        break;
    }
    }
}

bool family_ok(const bvp_instance &P) {
    switch (P.family) {
    case BVP_FAMILY_LINEAR_TANH: return P.M >= 1 && P.n_params >= 1 + P.M * P.M;
    case BVP_FAMILY_PENDULUM:
    case BVP_FAMILY_BRATU:
    case BVP_FAMILY_SINGPERT:
    case BVP_FAMILY_TURNING: return P.M == 2 && P.n_params >= 1;
    case BVP_FAMILY_SCALAR_TU: return P.M == 1;
    case BVP_FAMILY_SWIRL: return P.M == 6 && P.n_params >= 1;
    }
    return false;
}

// ---------------------------------------------------------------- mesh helpers
// range(t0; stop=t1, length=N+1) (Core/utils.jl:694-700).  The node is the
// correctly rounded value of (t0*(N-k) + t1*k)/N, computed in long double.
void uniform_mesh(double t0, double t1, int N, double *mesh) {
    for (int k = 0; k <= N; ++k) {
        long double v = ((long double)t0 * (N - k) + (long double)t1 * k) / (long double)N;
        mesh[k] = (double)v;
    }
    mesh[0] = t0;
    mesh[N] = t1;
}

// interval(mesh, t) = clamp(searchsortedfirst(mesh, t) - 1, 1, N)  (Core/utils.jl:119-121),
// returned 0-based.
int interval_of(const double *mesh, int N, double t) {
    const double *it = std::lower_bound(mesh, mesh + N + 1, t);
    int j1 = (int)(it - mesh) + 1;        // 1-based searchsortedfirst
    int iv = std::min(std::max(j1 - 1, 1), N);
    return iv - 1;
}

// ---------------------------------------------------------------- workspace
struct Ctx {
    const bvp_instance &P;
    const Tab &T;
    int method;
    int M;
    Ctx(const bvp_instance &p, int m) : P(p), T(tableau(m)), method(m), M(p.M) {}
};

// Stages of one interval: Y_r = (1-v_r) y_i + v_r y_{i+1} + h sum_{j<r} x_rj K_j,
// K_r = f(Y_r, t_i + c_r h)   (MIRK/collocation.jl:60-66)
void interval_stages(const Ctx &C, double ti, double h, const double *yi, const double *yj,
                     double *K, double *Y) {
    const int M = C.M, s = C.T.s;
    std::vector<double> acc(M);
    for (int r = 0; r < s; ++r) {
        double *Yr = Y + r * M;
        for (int c = 0; c < M; ++c) Yr[c] = (1 - C.T.v[r]) * yi[c] + C.T.v[r] * yj[c];
        if (r > 0) {
            for (int c = 0; c < M; ++c) {
                double a = 0;
                for (int j = 0; j < r; ++j) a += K[j * M + c] * C.T.x[r][j];
                Yr[c] += h * a;
            }
        }
        f_eval(C.P, ti + C.T.c[r] * h, Yr, K + r * M);
    }
}

// Phi_i = y_{i+1} - y_i - h sum_r b_r K_r   (MIRK/collocation.jl:69-70)
void interval_phi(const Ctx &C, double h, const double *yi, const double *yj, const double *K,
                  double *phi) {
    const int M = C.M, s = C.T.s;
    for (int c = 0; c < M; ++c) {
        double a = 0;
        for (int r = 0; r < s; ++r) a += K[r * M + c] * C.T.b[r];
        phi[c] = (yj[c] - yi[c]) - h * a;
    }
}

// Extra interpolation stages k*_r of one interval (MIRK/interpolation.jl:337-372)
void interval_interp_stages(const Ctx &C, double ti, double h, const double *yi, const double *yj,
                            const double *K, double *KI) {
    const int M = C.M, s = C.T.s, ns = C.T.sstar - C.T.s;
    std::vector<double> Y(M);
    for (int r = 0; r < ns; ++r) {
        for (int c = 0; c < M; ++c) {
            double a = 0;
            for (int j = 0; j < s; ++j) a += K[j * M + c] * C.T.xs[r][j];
            for (int j = 0; j < r; ++j) a += KI[j * M + c] * C.T.xs[r][s + j];
            Y[c] = a * h + (1 - C.T.vs[r]) * yi[c] + C.T.vs[r] * yj[c];
        }
        f_eval(C.P, ti + C.T.cs[r] * h, Y.data(), KI + r * M);
    }
}

// z = y_i + h (K w[0:s] + KI w[s:]),  z' = K w' + KI w'   (MIRK/adaptivity.jl:623-666)
void sum_stages(const Ctx &C, double h, const double *yi, const double *K, const double *KI,
                const double *w, const double *wp, double *z, double *zp) {
    const int M = C.M, s = C.T.s, ns = C.T.sstar - C.T.s;
    for (int c = 0; c < M; ++c) {
        double a = 0, ap = 0;
        for (int r = 0; r < s; ++r) { a += K[r * M + c] * w[r]; if (zp) ap += K[r * M + c] * wp[r]; }
        for (int r = 0; r < ns; ++r) { a += KI[r * M + c] * w[s + r]; if (zp) ap += KI[r * M + c] * wp[s + r]; }
        z[c] = a * h + yi[c];
        if (zp) zp[c] = ap;
    }
}

// Exact Jacobian blocks of Phi_i by the chain rule through the stage recursion.
//   dY_r/dy_i = (1-v_r) I + h sum_{j<r} x_rj dK_j/dy_i,  dK_r/dy_i = J_r dY_r/dy_i
//   S = -I - h sum_r b_r dK_r/dy_i,  R = I - h sum_r b_r dK_r/dy_{i+1}
void interval_jac(const Ctx &C, double ti, double h, const double *Y, double *S, double *R) {
    const int M = C.M, s = C.T.s, MM = M * M;
    std::vector<double> dKa(s * MM), dKb(s * MM), dYa(MM), dYb(MM), Jr(MM);
    for (int r = 0; r < s; ++r) {
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j) {
                double ea = (i == j) ? (1 - C.T.v[r]) : 0.0, eb = (i == j) ? C.T.v[r] : 0.0;
                double aa = 0, ab = 0;
                for (int q = 0; q < r; ++q) {
                    aa += C.T.x[r][q] * dKa[q * MM + i * M + j];
                    ab += C.T.x[r][q] * dKb[q * MM + i * M + j];
                }
                dYa[i * M + j] = ea + h * aa;
                dYb[i * M + j] = eb + h * ab;
            }
        J_eval(C.P, ti + C.T.c[r] * h, Y + r * M, Jr.data());
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j) {
                double aa = 0, ab = 0;
                for (int q = 0; q < M; ++q) {
                    aa += Jr[i * M + q] * dYa[q * M + j];
                    ab += Jr[i * M + q] * dYb[q * M + j];
                }
                dKa[r * MM + i * M + j] = aa;
                dKb[r * MM + i * M + j] = ab;
            }
    }
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            double aa = 0, ab = 0;
            for (int r = 0; r < s; ++r) {
                aa += C.T.b[r] * dKa[r * MM + i * M + j];
                ab += C.T.b[r] * dKb[r * MM + i * M + j];
            }
            S[i * M + j] = (i == j ? -1.0 : 0.0) - h * aa;
            R[i * M + j] = (i == j ? 1.0 : 0.0) - h * ab;
        }
}

// ---------------------------------------------------------------- BC rows
struct BcRow {
    int node;      // mesh node carrying the unit Jacobian entry
    int comp;
    bool side_b;   // TWO_POINT: row belongs to bcb (after Phi)
    bool direct;   // value read straight from a node (no interpolation)
    int iv;        // interval for the continuous extension (STANDARD, !direct)
    double tau;
};

// Resolve each BC row on the current mesh.  TWO_POINT: Core/utils.jl:157-209.
// STANDARD: EvalSol shortcut at the ends, else interval + tau
// (MIRK/interpolation.jl:214-226).
int resolve_bc(const bvp_instance &P, int N, const double *mesh, std::vector<BcRow> &rows) {
    rows.clear();
    for (int k = 0; k < P.n_bc; ++k) {
        BcRow r{};
        r.comp = P.bc_comp[k];
        if (r.comp < 0 || r.comp >= P.M) return BVP_EINVAL;
        const double t = P.bc_t[k];
        if (P.bc_kind == BVP_BC_TWO_POINT) {
            if (t == P.t0) { r.node = 0; r.side_b = false; }
            else if (t == P.t1) { r.node = N; r.side_b = true; }
            else return BVP_EINVAL;
            r.direct = true;
        } else {
            if (t == mesh[0]) { r.node = 0; r.direct = true; }
            else if (t == mesh[N]) { r.node = N; r.direct = true; }
            else {
                r.iv = interval_of(mesh, N, t);
                r.tau = (t - mesh[r.iv]) / (mesh[r.iv + 1] - mesh[r.iv]);
                r.node = r.iv;
                r.direct = false;
            }
        }
        rows.push_back(r);
    }
    return 0;
}

// Residual of one instance in the reference layout.  Also returns the stages.
struct Eval {
    std::vector<double> K, Y, KI, phi;    // per interval
    std::vector<double> bc;               // per BC row, in input order
};

void evaluate(const Ctx &C, int N, const double *mesh, const double *y,
              const std::vector<BcRow> &rows, Eval &E, bool need_interp_all) {
    const int M = C.M, s = C.T.s, ns = C.T.sstar - C.T.s;
    E.K.assign((size_t)N * s * M, 0.0);
    E.Y.assign((size_t)N * s * M, 0.0);
    E.phi.assign((size_t)N * M, 0.0);
    E.KI.assign((size_t)N * ns * M, 0.0);
    for (int i = 0; i < N; ++i) {
        const double h = mesh[i + 1] - mesh[i];
        interval_stages(C, mesh[i], h, y + i * M, y + (i + 1) * M, &E.K[(size_t)i * s * M],
                        &E.Y[(size_t)i * s * M]);
        interval_phi(C, h, y + i * M, y + (i + 1) * M, &E.K[(size_t)i * s * M], &E.phi[(size_t)i * M]);
    }
    if (need_interp_all)
        for (int i = 0; i < N; ++i) {
            const double h = mesh[i + 1] - mesh[i];
            interval_interp_stages(C, mesh[i], h, y + i * M, y + (i + 1) * M,
                                   &E.K[(size_t)i * s * M], &E.KI[(size_t)i * ns * M]);
        }
    E.bc.assign(rows.size(), 0.0);
    std::vector<double> w(9), wp(9), z(M);
    for (size_t k = 0; k < rows.size(); ++k) {
        const BcRow &r = rows[k];
        double val;
        if (r.direct) {
            val = y[r.node * M + r.comp];
        } else {
            const int i = r.iv;
            const double h = mesh[i + 1] - mesh[i];
            if (!need_interp_all)
                interval_interp_stages(C, mesh[i], h, y + i * M, y + (i + 1) * M,
                                       &E.K[(size_t)i * s * M], &E.KI[(size_t)i * ns * M]);
            interp_weights(C.method, r.tau, w.data(), wp.data());
            sum_stages(C, h, y + i * M, &E.K[(size_t)i * s * M], &E.KI[(size_t)i * ns * M],
                       w.data(), nullptr, z.data(), nullptr);
            val = z[r.comp];
        }
        E.bc[k] = val - C.P.bc_val[k];
    }
}

// Flatten to the reference residual layout.
void flatten_residual(const Ctx &C, int N, const std::vector<BcRow> &rows, const Eval &E,
                      double *out) {
    const int M = C.M;
    size_t o = 0;
    if (C.P.bc_kind == BVP_BC_TWO_POINT) {
        for (size_t k = 0; k < rows.size(); ++k) if (!rows[k].side_b) out[o++] = E.bc[k];
        for (size_t i = 0; i < (size_t)N * M; ++i) out[o++] = E.phi[i];
        for (size_t k = 0; k < rows.size(); ++k) if (rows[k].side_b) out[o++] = E.bc[k];
    } else {
        for (size_t k = 0; k < rows.size(); ++k) out[o++] = E.bc[k];
        for (size_t i = 0; i < (size_t)N * M; ++i) out[o++] = E.phi[i];
    }
}

double max_abs_residual(const Eval &E) {
    double m = 0;
    for (double v : E.phi) { if (std::isnan(v)) return NAN; m = std::max(m, std::fabs(v)); }
    for (double v : E.bc) { if (std::isnan(v)) return NAN; m = std::max(m, std::fabs(v)); }
    return m;
}

// ---------------------------------------------------------------- banded LU
// Gaussian elimination with partial pivoting on a band matrix (kl sub-, ku
// super-diagonals), LAPACK-style column storage with kl extra rows for fill.
struct Band {
    int n, kl, ku, ld;
    std::vector<double> ab;
    std::vector<int> piv;
    void init(int n_, int kl_, int ku_) {
        n = n_; kl = kl_; ku = ku_; ld = 2 * kl + ku + 1;
        ab.assign((size_t)ld * n, 0.0);
        piv.assign(n, 0);
    }
    double &at(int i, int j) { return ab[(size_t)j * ld + (kl + ku + i - j)]; }
    // returns 0, or 1 if a zero pivot was met
    int factor() {
        int sing = 0;
        int ju = 0;
        for (int j = 0; j < n; ++j) {
            const int km = std::min(kl, n - 1 - j);
            int p = j;
            double best = std::fabs(at(j, j));
            for (int i = 1; i <= km; ++i) {
                double v = std::fabs(at(j + i, j));
                if (v > best) { best = v; p = j + i; }
            }
            piv[j] = p;
            if (at(p, j) == 0.0) { sing = 1; continue; }
            ju = std::max(ju, std::min(p + ku, n - 1));
            if (p != j)
                for (int c = j; c <= ju; ++c) std::swap(at(j, c), at(p, c));
            const double inv = 1.0 / at(j, j);
            for (int i = 1; i <= km; ++i) at(j + i, j) *= inv;
            for (int c = j + 1; c <= ju; ++c) {
                const double a = at(j, c);
                if (a != 0.0)
                    for (int i = 1; i <= km; ++i) at(j + i, c) -= at(j + i, j) * a;
            }
        }
        return sing;
    }
    void solve(double *b) {
        for (int j = 0; j < n; ++j) {
            const int km = std::min(kl, n - 1 - j);
            if (piv[j] != j) std::swap(b[j], b[piv[j]]);
            for (int i = 1; i <= km; ++i) b[j + i] -= at(j + i, j) * b[j];
        }
        const int kv = kl + ku;
        for (int j = n - 1; j >= 0; --j) {
            b[j] /= at(j, j);
            const int lo = std::max(0, j - kv);
            for (int i = lo; i < j; ++i) b[i] -= at(i, j) * b[j];
        }
    }
};

// Assemble the Newton matrix J (rows ordered node by node so that it is banded:
// at node k, first the BC rows whose Jacobian entry sits on node k, then the
// rows of Phi_k) and solve J delta = -F.  Row order does not change delta.
int newton_solve(const Ctx &C, int N, const double *mesh, const double *y,
                 const std::vector<BcRow> &rows, const Eval &E, double *delta, Band &B) {
    const int M = C.M, MM = M * M, n = M * (N + 1);
    // row descriptors
    struct RowRef { int kind; int idx; };  // kind 0: bc row idx, 1: phi interval idx (M rows)
    std::vector<int> order_bc(rows.size());
    for (size_t k = 0; k < rows.size(); ++k) order_bc[k] = (int)k;
    std::stable_sort(order_bc.begin(), order_bc.end(),
                     [&](int a, int b) { return rows[a].node < rows[b].node; });
    // band widths from the structure
    int kl = 0, ku = 0;
    {
        int r = 0;
        size_t q = 0;
        for (int k = 0; k <= N; ++k) {
            while (q < order_bc.size() && rows[order_bc[q]].node == k) {
                int col = k * M + rows[order_bc[q]].comp;
                kl = std::max(kl, r - col); ku = std::max(ku, col - r);
                ++r; ++q;
            }
            if (k < N)
                for (int c = 0; c < M; ++c) {
                    kl = std::max(kl, r - k * M); ku = std::max(ku, (k + 2) * M - 1 - r);
                    ++r;
                }
        }
    }
    B.init(n, kl, ku);
    std::vector<double> rhs(n);
    std::vector<double> S(MM), R(MM);
    int r = 0;
    size_t q = 0;
    for (int k = 0; k <= N; ++k) {
        while (q < order_bc.size() && rows[order_bc[q]].node == k) {
            const int idx = order_bc[q];
            B.at(r, k * M + rows[idx].comp) = 1.0;
            rhs[r] = -E.bc[idx];
            ++r; ++q;
        }
        if (k < N) {
            const double h = mesh[k + 1] - mesh[k];
            interval_jac(C, mesh[k], h, &E.Y[(size_t)k * C.T.s * M], S.data(), R.data());
            for (int i = 0; i < M; ++i) {
                for (int j = 0; j < M; ++j) {
                    B.at(r, k * M + j) = S[i * M + j];
                    B.at(r, (k + 1) * M + j) = R[i * M + j];
                }
                rhs[r] = -E.phi[(size_t)k * M + i];
                ++r;
            }
        }
    }
    if (B.factor()) return 1;
    B.solve(rhs.data());
    for (int i = 0; i < n; ++i) delta[i] = rhs[i];
    return 0;
}

// ---------------------------------------------------------------- Newton
// The default nonlinear solver is the polyalgorithm
//   NewtonRaphson -> NewtonRaphson(linesearch = BackTracking()) -> TrustRegion()
// (Core/default_internal_solve.jl:31-45), each stage restarting from the same
// initial guess with its own iteration budget, stopping when max|F| <= tol,
// checked before every step (so maxiters = 0 returns the guess untouched,
// MIRKt/Core/mirk_basic_tests.jl:678-720).  `stages` is a bit mask of the
// stages to run (bit0 NR, bit1 NR + BackTracking, bit2 TrustRegion).
//
// The stages are external to the reference (NonlinearSolveFirstOrder "1.2, 2",
// LineSearch "0.1.4"); they are restated here from the published algorithms:
//  * BackTracking: LineSearches.jl's backtracking line search, order 3, on the
//    merit function phi(a) = |F(y + a d)|_2^2 / 2 with phi'(0) = F'J d = -|F|^2
//    for the Newton direction d: Armijo test phi(a) <= phi(0) + c1 a phi'(0),
//    c1 = 1e-4; the first reduction minimises the quadratic through phi(0),
//    phi'(0), phi(a), later ones the cubic through phi(0), phi'(0) and the last
//    two trials; each new step is kept within [rho_lo, rho_hi] = [0.1, 0.5] of
//    the previous one; a non-finite trial is halved first (at most 52 times);
//    more than 1000 reductions fail the stage.
//  * TrustRegion: the dogleg step (Newton step inside the region; else the
//    Cauchy point of the model |F + J d|^2 / 2 along -J'F, cut at the region
//    boundary or joined to the Newton step where the dogleg path leaves the
//    region) with the "Simple" radius update and TrustRegion()'s documented
//    defaults: initial radius max(|F(u0)|_2, max(u0) - min(u0)) / 11, step
//    accepted when rho = actual / predicted reduction > 1e-4, radius times 1/4
//    when rho < 1/4 (failure after more than 32 shrinks in a row), times 2 when
//    rho > 3/4.
// The Jacobian of the boundary rows is the frozen-stage one of the Newton
// matrix (MIRK/mirk.jl:565-572) in J d and J'F as well.
struct NlOut { int ret; int iters; double fnorm; int stage; };

double sumsq_residual(const Eval &E) {
    double s = 0;
    for (double v : E.phi) s += v * v;
    for (double v : E.bc) s += v * v;
    return s;
}

// LineSearches.jl BackTracking (order 3).  phi(a) evaluates the merit function.
// Returns 0 and the accepted step in *alpha, or 1 when the reductions ran out.
template <class Phi>
int backtracking(Phi &&phi, double phi0, double dphi0, double *alpha, double *phi_alpha) {
    const double c1 = 1e-4, rho_hi = 0.5, rho_lo = 0.1;
    const int iterations = 1000, iterfinitemax = 52;  // -log2(eps(Float64))
    auto nanmin = [](double a, double b) { return std::isnan(a) ? b : (std::isnan(b) ? a : std::min(a, b)); };
    auto nanmax = [](double a, double b) { return std::isnan(a) ? b : (std::isnan(b) ? a : std::max(a, b)); };
    double a1 = 1.0, a2 = 1.0;
    double fx0 = phi0, fx1 = phi(a1);
    int iterfinite = 0;
    while (!std::isfinite(fx1) && iterfinite < iterfinitemax) {
        ++iterfinite;
        a1 = a2;
        a2 = a1 / 2;
        fx1 = phi(a2);
    }
    int iteration = 0;
    while (fx1 > phi0 + c1 * a2 * dphi0) {
        if (++iteration > iterations) { *alpha = a2; *phi_alpha = fx1; return 1; }
        double at;
        if (iteration == 1) {
            at = -(dphi0 * a2 * a2) / (2 * (fx1 - phi0 - dphi0 * a2));
        } else {
            const double div = 1.0 / (a1 * a1 * a2 * a2 * (a2 - a1));
            const double e1 = fx1 - phi0 - dphi0 * a2, e0 = fx0 - phi0 - dphi0 * a1;
            const double a = (a1 * a1 * e1 - a2 * a2 * e0) * div;
            const double b = (-a1 * a1 * a1 * e1 + a2 * a2 * a2 * e0) * div;
            if (std::fabs(a) <= 2.220446049250313e-16) {
                at = dphi0 / (2 * b);
            } else {
                const double d = std::max(b * b - 3 * a * dphi0, 0.0);
                at = (-b + std::sqrt(d)) / (3 * a);
            }
        }
        a1 = a2;
        at = nanmin(at, a2 * rho_hi);
        a2 = nanmax(at, a2 * rho_lo);
        fx0 = fx1;
        fx1 = phi(a2);
    }
    *alpha = a2;
    *phi_alpha = fx1;
    return 0;
}

// Products with the Newton matrix at the stages in E.  Residual rows are
// ordered [Phi_1..Phi_N ; bc rows in input order].
void jac_mul(const Ctx &C, int N, const double *mesh, const Eval &E, const std::vector<BcRow> &rows,
             const double *v, double *out) {
    const int M = C.M, MM = M * M;
    std::vector<double> S(MM), R(MM);
    for (int i = 0; i < N; ++i) {
        interval_jac(C, mesh[i], mesh[i + 1] - mesh[i], &E.Y[(size_t)i * C.T.s * M], S.data(), R.data());
        for (int a = 0; a < M; ++a) {
            double acc = 0;
            for (int b = 0; b < M; ++b) acc += S[a * M + b] * v[i * M + b] + R[a * M + b] * v[(i + 1) * M + b];
            out[(size_t)i * M + a] = acc;
        }
    }
    for (size_t k = 0; k < rows.size(); ++k) out[(size_t)N * M + k] = v[rows[k].node * M + rows[k].comp];
}

void jac_tmul(const Ctx &C, int N, const double *mesh, const Eval &E, const std::vector<BcRow> &rows,
              const double *w, double *out) {
    const int M = C.M, MM = M * M;
    std::vector<double> S(MM), R(MM);
    std::fill(out, out + (size_t)M * (N + 1), 0.0);
    for (int i = 0; i < N; ++i) {
        interval_jac(C, mesh[i], mesh[i + 1] - mesh[i], &E.Y[(size_t)i * C.T.s * M], S.data(), R.data());
        for (int b = 0; b < M; ++b) {
            double sa = 0, sr = 0;
            for (int a = 0; a < M; ++a) {
                sa += S[a * M + b] * w[(size_t)i * M + a];
                sr += R[a * M + b] * w[(size_t)i * M + a];
            }
            out[i * M + b] += sa;
            out[(i + 1) * M + b] += sr;
        }
    }
    for (size_t k = 0; k < rows.size(); ++k) out[rows[k].node * M + rows[k].comp] += w[(size_t)N * M + k];
}

// Dogleg step of radius `radius` from the Newton step dn, the steepest-descent
// direction g = J'F with |g|^2 = l2 and |J g|^2 = jg2, and ss = |F|^2.  Writes
// the step into d and returns the predicted reduction |F|^2 - |F + J d|^2 of the
// linear model (F + J dn = 0 for the Newton step):
//   Newton step:            pred = ss
//   cut steepest descent:   d = -(radius/|g|) g, pred = 2 radius |g| - radius^2 jg2 / l2
//   dogleg, d = dc + tau (dn - dc), dc = -(l2/jg2) g:
//                           pred = ss - (1 - tau)^2 (ss - l2^2 / jg2)
// dn_norm2 = |dn|^2 (the caller's test |dn| <= radius selects the Newton step).
double dogleg(int n, const double *dn, const double *g, double l2, double jg2, double ss, double radius,
              double *d) {
    const double lg = std::sqrt(l2);
    const double dcauchy = l2 * lg / jg2;
    if (dcauchy >= radius) {
        const double lam = radius / lg;
        for (int i = 0; i < n; ++i) d[i] = -lam * g[i];
        return 2 * radius * lg - radius * radius * jg2 / l2;
    }
    const double beta = l2 / jg2;
    double a = 0, b = 0;
    for (int i = 0; i < n; ++i) {
        const double dc = -beta * g[i], di = dn[i] - dc;
        a += di * di;
        b += 2 * dc * di;
    }
    const double c = dcauchy * dcauchy - radius * radius;
    const double tau = (-b + std::sqrt(std::max(0.0, b * b - 4 * a * c))) / (2 * a);
    for (int i = 0; i < n; ++i) {
        const double dc = -beta * g[i];
        d[i] = dc + tau * (dn[i] - dc);
    }
    return ss - (1 - tau) * (1 - tau) * (ss - l2 * l2 / jg2);
}

// One NewtonRaphson stage (linesearch: with BackTracking).
int newton_stage(const Ctx &C, int N, const double *mesh, std::vector<double> &y, const std::vector<BcRow> &rows,
                 double tol, int maxiters, bool linesearch, NlOut &out) {
    const int n = C.M * (N + 1);
    std::vector<double> delta(n), yt(n);
    Eval E, Et;
    Band B;
    for (int it = 0;; ++it) {
        evaluate(C, N, mesh, y.data(), rows, E, false);
        out.fnorm = max_abs_residual(E);
        if (!std::isfinite(out.fnorm)) return BVP_RET_UNSTABLE;
        if (out.fnorm <= tol) return BVP_RET_SUCCESS;
        if (it >= maxiters) return BVP_RET_MAXITERS;
        if (newton_solve(C, N, mesh, y.data(), rows, E, delta.data(), B)) return BVP_RET_FAILURE;
        out.iters++;
        double alpha = 1.0;
        if (linesearch) {
            const double ss = sumsq_residual(E);
            auto phi = [&](double a) {
                for (int i = 0; i < n; ++i) yt[i] = y[i] + a * delta[i];
                evaluate(C, N, mesh, yt.data(), rows, Et, false);
                return 0.5 * sumsq_residual(Et);
            };
            double pa;
            if (backtracking(phi, 0.5 * ss, -ss, &alpha, &pa)) return BVP_RET_FAILURE;
        }
        for (int i = 0; i < n; ++i) y[i] += alpha * delta[i];
    }
}

int trust_region_stage(const Ctx &C, int N, const double *mesh, std::vector<double> &y,
                       const std::vector<BcRow> &rows, double tol, int maxiters, NlOut &out) {
    const int M = C.M, n = M * (N + 1), nr = M * N + (int)rows.size();
    std::vector<double> dn(n), g(n), d(n), yt(n), F(nr), Jg(nr);
    Eval E, Et;
    Band B;
    evaluate(C, N, mesh, y.data(), rows, E, false);
    double ss = sumsq_residual(E);
    double ymax = -INFINITY, ymin = INFINITY;
    for (double v : y) { ymax = std::max(ymax, v); ymin = std::min(ymin, v); }
    double radius = std::max(std::sqrt(ss), ymax - ymin) / 11;
    int shrinks = 0;
    bool fresh = true;  // dn, g belong to the current y
    double l2 = 0, jg2 = 0, dn2 = 0;
    for (int it = 0;; ++it) {
        out.fnorm = max_abs_residual(E);
        if (!std::isfinite(out.fnorm)) return BVP_RET_UNSTABLE;
        if (out.fnorm <= tol) return BVP_RET_SUCCESS;
        if (it >= maxiters) return BVP_RET_MAXITERS;
        if (fresh) {
            if (newton_solve(C, N, mesh, y.data(), rows, E, dn.data(), B)) return BVP_RET_FAILURE;
            dn2 = 0;
            for (double v : dn) dn2 += v * v;
            for (int i = 0; i < M * N; ++i) F[i] = E.phi[i];
            for (size_t k = 0; k < rows.size(); ++k) F[(size_t)M * N + k] = E.bc[k];
            jac_tmul(C, N, mesh, E, rows, F.data(), g.data());
            l2 = 0;
            for (double v : g) l2 += v * v;
            jac_mul(C, N, mesh, E, rows, g.data(), Jg.data());
            jg2 = 0;
            for (double v : Jg) jg2 += v * v;
            fresh = false;
        }
        out.iters++;
        double pred;
        if (std::sqrt(dn2) <= radius) {
            d = dn;
            pred = ss;
        } else {
            pred = dogleg(n, dn.data(), g.data(), l2, jg2, ss, radius, d.data());
        }
        for (int i = 0; i < n; ++i) yt[i] = y[i] + d[i];
        evaluate(C, N, mesh, yt.data(), rows, Et, false);
        const double sst = sumsq_residual(Et);
        const double rho = (ss - sst) / pred;
        if (rho > 1e-4) {  // accept
            y = yt;
            std::swap(E, Et);
            ss = sst;
            fresh = true;
        }
        if (rho < 0.25) {
            radius *= 0.25;
            if (++shrinks > 32) { out.fnorm = max_abs_residual(E); return BVP_RET_FAILURE; }
        } else {
            shrinks = 0;
            if (rho > 0.75) radius *= 2;
        }
    }
}

NlOut newton(const Ctx &C, int N, const double *mesh, std::vector<double> &y,
             const std::vector<BcRow> &rows, double tol, int maxiters, int stages) {
    const std::vector<double> y0 = y;
    NlOut out{BVP_RET_MAXITERS, 0, 0, -1};
    bool first = true;
    for (int st = 0; st < 3; ++st) {
        if (!(stages & (1 << st))) continue;
        if (!first) y = y0;
        first = false;
        out.stage = st;
        out.ret = st < 2 ? newton_stage(C, N, mesh, y, rows, tol, maxiters, st == 1, out)
                         : trust_region_stage(C, N, mesh, y, rows, tol, maxiters, out);
        if (out.ret == BVP_RET_SUCCESS) break;
    }
    return out;
}

// ---------------------------------------------------------------- defect
// error_estimate!(::DefectControl)  MIRK/adaptivity.jl:416-461
double defect(const Ctx &C, int N, const double *mesh, const double *y, Eval &E,
              std::vector<double> &errors) {
    const int M = C.M, s = C.T.s, ns = C.T.sstar - C.T.s;
    std::vector<double> w1(9), w1p(9), w2(9), w2p(9), z(M), zp(M), f1(M), f2(M), d1(M), d2(M);
    interp_weights(C.method, C.T.tau, w1.data(), w1p.data());
    interp_weights(C.method, 1.0 - C.T.tau, w2.data(), w2p.data());
    errors.assign((size_t)N * M, 0.0);
    double dn = 0;
    for (int i = 0; i < N; ++i) {
        const double h = mesh[i + 1] - mesh[i];
        interval_interp_stages(C, mesh[i], h, y + i * M, y + (i + 1) * M, &E.K[(size_t)i * s * M],
                               &E.KI[(size_t)i * ns * M]);
    }
    for (int i = 0; i < N; ++i) {
        const double h = mesh[i + 1] - mesh[i];
        const double *K = &E.K[(size_t)i * s * M], *KI = &E.KI[(size_t)i * ns * M];
        sum_stages(C, h, y + i * M, K, KI, w1.data(), w1p.data(), z.data(), zp.data());
        f_eval(C.P, mesh[i] + C.T.tau * h, z.data(), f1.data());
        double e1 = 0;
        for (int c = 0; c < M; ++c) { d1[c] = (zp[c] - f1[c]) / (std::fabs(f1[c]) + 1.0); e1 = std::max(e1, std::fabs(d1[c])); }
        sum_stages(C, h, y + i * M, K, KI, w2.data(), w2p.data(), z.data(), zp.data());
        f_eval(C.P, mesh[i] + (1.0 - C.T.tau) * h, z.data(), f2.data());
        double e2 = 0;
        for (int c = 0; c < M; ++c) { d2[c] = (zp[c] - f2[c]) / (std::fabs(f2[c]) + 1.0); e2 = std::max(e2, std::fabs(d2[c])); }
        const double *d = e1 > e2 ? d1.data() : d2.data();
        double ei = 0;
        for (int c = 0; c < M; ++c) { errors[(size_t)i * M + c] = d[c]; ei = std::max(ei, std::fabs(d[c])); }
        if (!(ei <= dn)) dn = ei;   // NaN propagates
    }
    return dn;
}

// ---------------------------------------------------------------- mesh selection
// half_mesh!  MIRK/adaptivity.jl:287-304
void half_mesh(const std::vector<double> &mesh, std::vector<double> &out) {
    const int n = (int)mesh.size() - 1;
    std::vector<double> dt(n);
    for (int i = 0; i < n; ++i) dt[i] = mesh[i + 1] - mesh[i];
    out.assign(2 * n + 1, 0.0);
    out[2 * n] = mesh[n];
    for (int i = 0; i < n; ++i) out[2 * i] = mesh[i];
    for (int i = 0; i < n; ++i) out[2 * i + 1] = (out[2 * i + 2] + out[2 * i]) / 2.0;
}

// Julia round(Int, x): ties to even
long long round_even(double x) {
    return (long long)std::nearbyint(x);  // default FE_TONEAREST
}

// mesh_selector!(::DefectControl) + redistribute!  MIRK/adaptivity.jl:23-75, 250-278.
// Returns BVP_RET_SUCCESS/FAILURE; fills new_mesh.
// max-abs of each interval's error vector
std::vector<double> interval_max(const std::vector<double> &errors, int n, int M) {
    std::vector<double> m(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < M; ++c) m[i] = std::max(m[i], std::fabs(errors[(size_t)i * M + c]));
    return m;
}

// The common part of the four mesh selectors (MIRK/adaptivity.jl:23-243): shat[i]
// is the controller's per-interval indicator, rho its halving threshold (1 for
// DefectControl, 2 for the global, sequential and hybrid controllers).
int mesh_select_shat(double rho, int max_sub, const std::vector<double> &mesh, std::vector<double> shat,
                     std::vector<double> &new_mesh) {
    const int Nn = (int)mesh.size();  // length(mesh)
    const int n = Nn - 1;
    std::vector<double> dt(n);
    for (int i = 0; i < n; ++i) dt[i] = mesh[i + 1] - mesh[i];
    double r1 = shat[0], r2 = 0;
    for (int i = 0; i < n; ++i) { r1 = std::max(r1, shat[i]); r2 += shat[i]; }
    const double r3 = r2 / (Nn - 1);
    long long n_predict = round_even(1.3 * r2 + 1);
    const double n_ = 0.1 * n;
    if (std::fabs((double)(n_predict - n)) < n_) n_predict = round_even(n + n_);
    if (r1 <= rho * r3) {
        const long long Nsub = 2LL * n;
        if (Nsub > max_sub) return BVP_RET_FAILURE;
        half_mesh(mesh, new_mesh);
        return BVP_RET_SUCCESS;
    }
    const long long lb = Nn / 2, ub = 4LL * (Nn - 1);
    const long long Nsub = std::min(std::max(n_predict, lb), ub);
    if (Nsub > max_sub) return BVP_RET_FAILURE;
    for (int i = 0; i < n; ++i) shat[i] /= dt[i];
    // redistribute!
    double zsum = 0;
    for (int i = 0; i < n; ++i) zsum += shat[i] * dt[i];
    const double zeta = zsum / (double)Nsub;
    new_mesh.assign((size_t)Nsub + 1, 0.0);
    new_mesh[0] = mesh[0];
    double t = mesh[0], integral = 0;
    int k = 0;            // 0-based old interval
    long long i = 0;      // nodes placed
    while (k < n) {
        const double next_piece = shat[k] * (mesh[k + 1] - t);
        const double int_next = integral + next_piece;
        if (int_next > zeta) {
            const double tn = (zeta - integral) / shat[k] + t;
            if (i + 1 > Nsub) return BVP_RET_FAILURE;  // more nodes than planned: the reference would index out of range
            new_mesh[(size_t)(i + 1)] = tn;
            t = tn;
            ++i;
            integral = 0;
        } else {
            integral = int_next;
            t = mesh[k + 1];
            ++k;
        }
    }
    new_mesh[(size_t)Nsub] = mesh[n];
    return BVP_RET_SUCCESS;
}

// mesh_selector!(::DefectControl): shat = (max|d_i| / abstol)^(1/(p+1)), rho = 1
int mesh_select(int order, double abstol, int max_sub, const std::vector<double> &mesh,
                const std::vector<double> &errors, int M, std::vector<double> &new_mesh) {
    const int n = (int)mesh.size() - 1;
    std::vector<double> shat = interval_max(errors, n, M);
    for (double &v : shat) v = std::pow(v / abstol, 1.0 / (order + 1));
    return mesh_select_shat(1.0, max_sub, mesh, shat, new_mesh);
}

// Solution transfer: y_new(node) = continuous extension of the old solve.
// transfer_mode REFERENCE reproduces the in-place update of the reference
// (MIRK/mirk.jl:366-370 with sum_stages! MIRK/adaptivity.jl:606-621): the base
// value y0.u[ii] is read from the array being written, so for ii < j it is the
// already-updated entry of the new mesh.
void transfer(const Ctx &C, int mode, const std::vector<double> &old_mesh, const std::vector<double> &yold,
              const Eval &E, const std::vector<double> &new_mesh, std::vector<double> &ynew) {
    const int M = C.M, s = C.T.s, ns = C.T.sstar - C.T.s;
    const int Nold = (int)old_mesh.size() - 1;
    const int Nnew = (int)new_mesh.size() - 1;
    std::vector<double> work = yold;  // y0.u storage
    work.resize((size_t)M * std::max(Nold + 1, Nnew + 1), 0.0);
    std::vector<double> w(9), wp(9), z(M);
    ynew.assign((size_t)M * (Nnew + 1), 0.0);
    for (int j = 0; j <= Nnew; ++j) {
        const double m = new_mesh[j];
        const int ii = interval_of(old_mesh.data(), Nold, m);
        const double dt = old_mesh[ii + 1] - old_mesh[ii];
        const double tau = (m - old_mesh[ii]) / dt;
        interp_weights(C.method, tau, w.data(), wp.data());
        const double *base = (mode == BVP_TRANSFER_REFERENCE) ? &work[(size_t)ii * M] : &yold[(size_t)ii * M];
        sum_stages(C, dt, base, &E.K[(size_t)ii * s * M], &E.KI[(size_t)ii * ns * M], w.data(), nullptr,
                   z.data(), nullptr);
        for (int c = 0; c < M; ++c) { work[(size_t)j * M + c] = z[c]; ynew[(size_t)j * M + c] = z[c]; }
    }
}

// ---------------------------------------------------------------- full solve
int instance_mesh_and_guess(const bvp_instance &P, std::vector<double> &mesh, std::vector<double> &y) {
    if (P.N < 1) return BVP_EINVAL;
    mesh.resize(P.N + 1);
    if (P.mesh) std::memcpy(mesh.data(), P.mesh, sizeof(double) * (P.N + 1));
    else uniform_mesh(P.t0, P.t1, P.N, mesh.data());
    y.resize((size_t)P.M * (P.N + 1));
    if (P.guess) std::memcpy(y.data(), P.guess, sizeof(double) * y.size());
    else {
        if (!P.u0) return BVP_EINVAL;
        for (int k = 0; k <= P.N; ++k)
            for (int c = 0; c < P.M; ++c) y[(size_t)k * P.M + c] = P.u0[c];
    }
    return 0;
}

int check_instance(const bvp_instance &P) {
    if (!family_ok(P)) return BVP_EINVAL;
    if (P.n_bc != P.M) return BVP_EINVAL;
    if (P.N < 1) return BVP_EINVAL;
    if (!(P.t1 > P.t0)) return BVP_EINVAL;
    return 0;
}

// ---------------------------------------------------------------- global error
// Julia's max: NaN if either argument is NaN
double jmax(double a, double b) { return (a != a || b != b) ? NAN : std::max(a, b); }

// global_error + GE_subinterval! (MIRK/adaptivity.jl:540-556): per node the relative
// difference e_i = (high_i - low_i) / (1 + |low_i|); per interval the one of its two
// end-node vectors with the larger max-abs (the left one on ties, the right one when
// a max-abs is NaN: ifelse(a >= b, ...)); returns the max-abs over all intervals
// (NaN-propagating, as Julia's maximum).
double global_error_of(const std::vector<double> &high, const std::vector<double> &low, int n, int M,
                       std::vector<double> &errors) {
    std::vector<double> e((size_t)(n + 1) * M);
    for (size_t q = 0; q < e.size(); ++q) e[q] = (high[q] - low[q]) / (1.0 + std::fabs(low[q]));
    errors.assign((size_t)n * M, 0.0);
    double norm = 0;
    for (int i = 0; i < n; ++i) {
        double a = 0, b = 0;
        for (int c = 0; c < M; ++c) {
            a = jmax(a, std::fabs(e[(size_t)i * M + c]));
            b = jmax(b, std::fabs(e[(size_t)(i + 1) * M + c]));
        }
        const bool left = a >= b;
        for (int c = 0; c < M; ++c) errors[(size_t)i * M + c] = e[(size_t)(left ? i : i + 1) * M + c];
        norm = jmax(norm, left ? a : b);
    }
    return norm;
}

// error_estimate!(::GlobalErrorControl, ::REErrorControl) (MIRK/adaptivity.jl:511-525):
// the same method, non-adaptive, on the halved mesh from halve_sol (:311-333) of the
// current solution, solved with the same nonlinear solver and tolerance (its return
// code is not consulted, as in the reference); the global error of the current
// solution at the old nodes, its norm scaled by 2^p / (2^p - 1).
// error_estimate!(::GlobalErrorControl, ::HOErrorControl) (:527-538): MIRK(p+2)
// (MIRK6 for MIRK4) on the same mesh from the current solution; norm unscaled.
double global_error_estimate(const bvp_instance &P, const bvp_opts &o, const Ctx &C, const std::vector<double> &mesh,
                             const std::vector<double> &y, double tol, std::vector<double> &errors) {
    const int M = P.M, n = (int)mesh.size() - 1;
    std::vector<BcRow> rows;
    if (o.global_method == BVP_GLOBAL_RE) {
        std::vector<double> mh, yh((size_t)(2 * n + 1) * M);
        half_mesh(mesh, mh);
        for (int i = 0; i <= n; ++i)
            for (int c = 0; c < M; ++c) yh[(size_t)(2 * i) * M + c] = y[(size_t)i * M + c];
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < M; ++c)
                yh[(size_t)(2 * i + 1) * M + c] = (yh[(size_t)(2 * i + 2) * M + c] + yh[(size_t)(2 * i) * M + c]) / 2.0;
        resolve_bc(P, 2 * n, mh.data(), rows);
        newton(C, 2 * n, mh.data(), yh, rows, tol, o.max_newton_iters, o.nlsolve_stages);
        std::vector<double> hi((size_t)(n + 1) * M);
        for (int i = 0; i <= n; ++i)
            for (int c = 0; c < M; ++c) hi[(size_t)i * M + c] = yh[(size_t)(2 * i) * M + c];
        const double p2 = std::pow(2.0, C.T.order);
        return global_error_of(hi, y, n, M, errors) * p2 / (p2 - 1.0);
    }
    Ctx Ch(P, C.method + 2);
    std::vector<double> yh = y;
    resolve_bc(P, n, mesh.data(), rows);
    newton(Ch, n, mesh.data(), yh, rows, tol, o.max_newton_iters, o.nlsolve_stages);
    return global_error_of(yh, y, n, M, errors);
}

int solve_one(const bvp_instance &P, const bvp_opts &o, bvp_result &res) {
    int rc = check_instance(P);
    if (rc) return rc;
    Ctx C(P, o.method);
    std::vector<double> mesh, y;
    rc = instance_mesh_and_guess(P, mesh, y);
    if (rc) return rc;
    const double tol = o.newton_abstol > 0 ? o.newton_abstol : o.abstol;
    if (o.controller < BVP_CONTROL_DEFECT || o.controller > BVP_CONTROL_HYBRID) return BVP_EINVAL;
    if (o.controller != BVP_CONTROL_DEFECT && o.global_method == BVP_GLOBAL_HO && o.method != BVP_MIRK4)
        return BVP_ENOTSUP;  // MIRK(p+2) for MIRK6 would be MIRK8
    const int M = P.M;
    int N = P.N;
    int info = BVP_RET_SUCCESS, nl_ret = BVP_RET_SUCCESS;
    double error_norm = 2 * o.abstol, last_defect = NAN;
    int outer = 0, newton_iters = 0, fb = 0;
    std::vector<BcRow> rows;
    std::vector<double> errors, new_mesh, ynew;
    Eval E;
    const int MAX_OUTER = 64;
    for (;;) {
        ++outer;
        if (resolve_bc(P, N, mesh.data(), rows)) return BVP_EINVAL;
        NlOut nl = newton(C, N, mesh.data(), y, rows, tol, o.max_newton_iters, o.nlsolve_stages);
        newton_iters += nl.iters;
        if (nl.stage == 1) fb |= 1;
        if (nl.stage == 2) fb |= 16;
        nl_ret = nl.ret;
        error_norm = 2 * o.abstol;
        if (!o.adaptive) { info = nl.ret; break; }
        info = nl.ret;
        // error_estimate! of the controller (MIRK/adaptivity.jl:370-525) and the
        // indicator its mesh_selector! equidistributes (:23-243)
        std::vector<double> shat;
        double rho = 1.0;
        if (info == BVP_RET_SUCCESS) {
            const int p = C.T.order;
            evaluate(C, N, mesh.data(), y.data(), rows, E, false);
            if (o.controller == BVP_CONTROL_GLOBAL) {
                error_norm = global_error_estimate(P, o, C, mesh, y, tol, errors);
                shat = interval_max(errors, N, M);
                for (double &v : shat) v = std::pow(v / o.abstol, 1.0 / p);
                rho = 2.0;
                evaluate(C, N, mesh.data(), y.data(), rows, E, true);  // stages for the transfer
            } else if (o.controller == BVP_CONTROL_SEQUENTIAL) {
                error_norm = defect(C, N, mesh.data(), y.data(), E, errors);
                last_defect = error_norm;
                info = (error_norm > o.defect_threshold || std::isnan(error_norm)) ? BVP_RET_FAILURE : BVP_RET_SUCCESS;
                if (info == BVP_RET_SUCCESS && error_norm <= o.abstol)
                    error_norm = global_error_estimate(P, o, C, mesh, y, tol, errors);
                shat = interval_max(errors, N, M);
                for (double &v : shat) v = std::pow(v / o.abstol, 1.0 / (p + 1));
                rho = 2.0;
            } else if (o.controller == BVP_CONTROL_HYBRID) {
                std::vector<double> ge;
                const double dn = defect(C, N, mesh.data(), y.data(), E, errors);
                last_defect = dn;
                const double gn = global_error_estimate(P, o, C, mesh, y, tol, ge);
                error_norm = o.hybrid_de * dn + o.hybrid_ge * gn;
                std::vector<double> s1 = interval_max(errors, N, M), s2 = interval_max(ge, N, M);
                shat.resize(N);
                for (int i = 0; i < N; ++i)
                    shat[i] = std::pow(s1[i] / o.abstol, 1.0 / (p + 1)) + std::pow(s2[i] / o.abstol, 1.0 / (p + 1));
                rho = 2.0;
            } else {
                error_norm = defect(C, N, mesh.data(), y.data(), E, errors);
                last_defect = error_norm;
                info = (error_norm > o.defect_threshold || std::isnan(error_norm)) ? BVP_RET_FAILURE : BVP_RET_SUCCESS;
                shat = interval_max(errors, N, M);
                for (double &v : shat) v = std::pow(v / o.abstol, 1.0 / (p + 1));
            }
        }
        if (info == BVP_RET_SUCCESS) {
            if (error_norm > o.abstol) {
                info = mesh_select_shat(rho, o.max_num_subintervals, mesh, shat, new_mesh);
                if (info == BVP_RET_SUCCESS) {
                    transfer(C, o.transfer_mode, mesh, y, E, new_mesh, ynew);
                    mesh = new_mesh;
                    y = ynew;
                    N = (int)mesh.size() - 1;
                }
            }
        } else {
            if (2 * N > o.max_num_subintervals) {
                info = BVP_RET_FAILURE;
            } else {
                std::vector<double> hm;
                half_mesh(mesh, hm);
                mesh = hm;
                N = (int)mesh.size() - 1;
                y.assign((size_t)M * (N + 1), 0.0);
                info = BVP_RET_SUCCESS;
            }
        }
        if (!(info == BVP_RET_SUCCESS && error_norm > o.abstol)) break;
        if (outer >= MAX_OUTER) { info = BVP_RET_FAILURE; break; }
    }
    // __build_solution: Core/utils.jl:798-801
    res.retcode = (nl_ret == BVP_RET_SUCCESS) ? info : nl_ret;
    res.N_final = N;
    res.newton_iters = newton_iters;
    res.outer_iters = outer;
    res.fallback_used = fb;
    res.defect_norm = last_defect;
    res.err_vs_fp64 = NAN;
    if (resolve_bc(P, N, mesh.data(), rows) == 0) {
        evaluate(C, N, mesh.data(), y.data(), rows, E, false);
        res.resid_norm = max_abs_residual(E);
    }
    if (res.capacity < N + 1) return BVP_EINVAL;
    if (res.mesh) std::memcpy(res.mesh, mesh.data(), sizeof(double) * (N + 1));
    if (res.u) std::memcpy(res.u, y.data(), sizeof(double) * M * (N + 1));
    return 0;
}

// ======================================================================
// FIRK collocation in the expanded form (lib/BoundaryValueDiffEqFIRK/src/,
// FIRK/ below; the reference's default nested_nlsolve = false):
//   tableaus ............ FIRK/radau_tableaus.jl:1-163, FIRK/lobatto_tableaus.jl:1-323
//                         (derived here from the nodes, firk_tableau below)
//   unknowns ............ FIRK/firk.jl:76-92 (extend_y / shrink_y)
//   Phi! (expanded) ..... FIRK/collocation.jl:59-93
//   loss / layout ....... FIRK/firk.jl:1404-1455
//   Jacobian ............ exact (the reference differentiates the loss with sparse
//                         ForwardDiff on the banded prototype, FIRK/firk.jl:806-922)
//   linear solve ........ banded Gaussian elimination with partial pivoting
//   Newton .............. the polyalgorithm of the MIRK path (newton above)
//   outer loop .......... FIRK/firk.jl:481-525, 572-621
//   defect .............. FIRK/adaptivity.jl:349-405, 581-606 (eval_q)
//   mesh_selector! ...... FIRK/adaptivity.jl:247-300 (the MIRK rules, rho = 1)
//   interpolation ....... FIRK/adaptivity.jl:8-48, 183-227; FIRK/interpolation.jl:229-316
// STANDARD boundary rows at t0 or t1 read the node (EvalSol's end-point shortcut,
// FIRK/interpolation.jl:344-345); rows inside the span evaluate EvalSol's quintic
// (firk_evalsol, FIRK/interpolation.jl:338-441).
// The solution transfer after a mesh redistribution restates what interp_eval!
// computes for each new node (the quintic interpolant of the old solution,
// FIRK/adaptivity.jl:8-48) and places the nodes as extend_y does for a new problem
// (stages K_ij = y_{i+1}, FIRK/firk.jl:76-83).  The reference writes the new node
// values into the first N+1 slots of its expanded state and leaves the remaining
// slots of the resized buffer as they were or uninitialised (FIRK/firk.jl:600-606
// with FIRK/utils.jl:1-8), and it scales h by the mesh-growth factor there; that
// state is not reproducible, so the restart differs from the reference's.  The mesh
// sequence and the converged solutions do not depend on the restart as long as Newton
// converges.

struct FTab {
    int code = 0, kind = 0, s = 0, order = 0;
    bool adaptive = true;
    double A[7][7] = {}, b[7] = {}, c[7] = {}, Q[7][7] = {}, tau = 0;
};

using ld = long double;

// P_n(x) and P_n'(x) (|x| < 1) by the three-term recurrence
void legendre(int n, ld x, ld &p, ld &dp) {
    ld p0 = 1, p1 = x;
    if (n == 0) { p = 1; dp = 0; return; }
    for (int k = 2; k <= n; ++k) {
        const ld p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    p = p1;
    dp = n * (x * p1 - p0) / (x * x - 1);
}

// zeros of g on (-1, 1) by a sign scan and bisection
template <class G>
std::vector<ld> zeros_in(G &&g) {
    std::vector<ld> z;
    const int NG = 20000;
    ld xa = -1 + 2.0L / NG, ga = g(xa);
    for (int k = 2; k < NG; ++k) {
        const ld xb = -1 + 2.0L * k / NG, gb = g(xb);
        if (ga == 0) z.push_back(xa);
        else if ((ga < 0) != (gb < 0) && gb != 0) {
            ld lo = xa, hi = xb, glo = ga;
            for (int it = 0; it < 200; ++it) {
                const ld mid = (lo + hi) / 2, gm = g(mid);
                if ((gm < 0) == (glo < 0)) { lo = mid; glo = gm; } else hi = mid;
            }
            z.push_back((lo + hi) / 2);
        }
        xa = xb;
        ga = gb;
    }
    return z;
}

// inverse of an n x n matrix (Gauss-Jordan, partial pivoting)
void inverse(int n, std::vector<ld> a, std::vector<ld> &inv) {
    inv.assign((size_t)n * n, 0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1;
    for (int j = 0; j < n; ++j) {
        int p = j;
        for (int i = j + 1; i < n; ++i)
            if (std::fabs(a[(size_t)i * n + j]) > std::fabs(a[(size_t)p * n + j])) p = i;
        for (int c = 0; c < n; ++c) {
            std::swap(a[(size_t)j * n + c], a[(size_t)p * n + c]);
            std::swap(inv[(size_t)j * n + c], inv[(size_t)p * n + c]);
        }
        const ld d = a[(size_t)j * n + j];
        for (int c = 0; c < n; ++c) { a[(size_t)j * n + c] /= d; inv[(size_t)j * n + c] /= d; }
        for (int i = 0; i < n; ++i) {
            if (i == j) continue;
            const ld m = a[(size_t)i * n + j];
            for (int c = 0; c < n; ++c) {
                a[(size_t)i * n + c] -= m * a[(size_t)j * n + c];
                inv[(size_t)i * n + c] -= m * inv[(size_t)j * n + c];
            }
        }
    }
}

// The tableau of a FIRK method code.  The nodes are those of the methods (Radau IIA:
// the zeros of P_s(2t-1) - P_{s-1}(2t-1); Lobatto: 0, 1 and the zeros of P'_{s-1}(2t-1));
// with l_j the Lagrange basis on the nodes,
//   collocation  A_ij = int_0^{c_i} l_j,  b_j = int_0^1 l_j      (Radau IIA, Lobatto IIIA)
//   Lobatto IIIB A_ij = b_j (1 - A^IIIA_ji / b_i)
//   Lobatto IIIC A_i1 = b_1 and sum_j A_ij c_j^(k-1) = c_i^k / k for k = 1..s-1
// and q_coeff Q_kj = the coefficient of tau^(k+1) in int_0^tau l_j (the stage polynomial
// of the defect estimate and the interpolant, "bvp5c" in the reference's comments).
// tau* is the reference's own sample point (data of the tables).
bool firk_tableau(int code, FTab &T) {
    T = FTab();
    T.code = code;
    code %= 1000;  // BVP_FIRK_NESTED adds 1000
    T.kind = code / 100;
    T.s = code % 100;
    const int s = T.s;
    if (T.kind == 1) {
        if (s != 1 && s != 2 && s != 3 && s != 5 && s != 7) return false;
        T.order = 2 * s - 1;
        T.adaptive = s != 1;
    } else if (T.kind >= 2 && T.kind <= 4) {
        if (s < 2 || s > 5) return false;
        T.order = 2 * s - 2;
        T.adaptive = !(s == 2 && (T.kind == 3 || T.kind == 4));
    } else {
        return false;
    }
    std::vector<ld> c;
    if (T.kind == 1) {
        c = zeros_in([&](ld x) {
            ld p, dp, q, dq;
            legendre(s, x, p, dp);
            legendre(s - 1, x, q, dq);
            return p - q;
        });
        c.push_back(1);
    } else {
        c.push_back(-1);
        if (s > 2) {
            std::vector<ld> z = zeros_in([&](ld x) {
                ld p, dp;
                legendre(s - 1, x, p, dp);
                return dp;
            });
            c.insert(c.end(), z.begin(), z.end());
        }
        c.push_back(1);
    }
    if ((int)c.size() != s) return false;
    for (ld &v : c) v = (v + 1) / 2;
    std::vector<ld> Cp((size_t)s * s), L;
    for (int i = 0; i < s; ++i)
        for (int m = 0; m < s; ++m) Cp[(size_t)i * s + m] = std::pow(c[i], (ld)m);
    inverse(s, Cp, L);  // L[m][j]: coefficient of t^m in l_j
    ld Acol[7][7], bcol[7];
    for (int j = 0; j < s; ++j) {
        bcol[j] = 0;
        for (int m = 0; m < s; ++m) bcol[j] += L[(size_t)m * s + j] / (m + 1);
        for (int i = 0; i < s; ++i) {
            ld a = 0;
            for (int m = 0; m < s; ++m) a += L[(size_t)m * s + j] * std::pow(c[i], (ld)(m + 1)) / (m + 1);
            Acol[i][j] = a;
        }
    }
    ld A[7][7];
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) {
            if (T.kind <= 2) A[i][j] = Acol[i][j];
            else if (T.kind == 3) A[i][j] = bcol[j] * (1 - Acol[j][i] / bcol[i]);
        }
    if (T.kind == 4) {
        const int n = s - 1;
        std::vector<ld> V((size_t)n * n), Vi;
        for (int k = 0; k < n; ++k)
            for (int j = 1; j < s; ++j) V[(size_t)k * n + (j - 1)] = std::pow(c[j], (ld)k);
        inverse(n, V, Vi);
        for (int i = 0; i < s; ++i) {
            A[i][0] = bcol[0];
            for (int j = 1; j < s; ++j) {
                ld a = 0;
                for (int k = 0; k < n; ++k) {
                    const ld rhs = std::pow(c[i], (ld)(k + 1)) / (k + 1) - (k == 0 ? bcol[0] : 0);
                    a += Vi[(size_t)(j - 1) * n + k] * rhs;
                }
                A[i][j] = a;
            }
        }
    }
    for (int i = 0; i < s; ++i) {
        T.c[i] = (double)c[i];
        T.b[i] = (double)bcol[i];
        for (int j = 0; j < s; ++j) {
            T.A[i][j] = (double)A[i][j];
            T.Q[i][j] = (double)(L[(size_t)i * s + j] / (i + 1));
        }
    }
    // tau* (FIRK/radau_tableaus.jl, FIRK/lobatto_tableaus.jl: tau_star of each method)
    if (T.kind == 1) T.tau = 0.0;
    else if (s == 2 || s == 4) T.tau = 0.5;
    else if (s == 3) T.tau = T.kind == 4 ? 0.7886751345948129 : 0.21132486540518713;
    else T.tau = T.kind == 4 ? 0.6699905217924309 : 0.33000947820757126;
    return true;
}

struct FCtx {
    const bvp_instance &P;
    FTab T;
    int M;
    double itol = 3.000213634488528e-13;  // nested form: the inner solves' tolerance, eps(Float64)^(4/5)
    FCtx(const bvp_instance &p, const FTab &t) : P(p), T(t), M(p.M) {}
    int s() const { return T.s; }
    size_t yoff(int i) const { return (size_t)i * (T.s + 1) * M; }
    size_t koff(int i, int j) const { return ((size_t)i * (T.s + 1) + 1 + j) * M; }
    size_t nx(int N) const { return (size_t)M * ((size_t)N * (T.s + 1) + 1); }
};

// extend_y (FIRK/firk.jl:76-83): nodes into the expanded order, every stage of
// interval i set to y_{i+1}; or the given stages
void firk_extend(const FCtx &C, int N, const double *y, const double *stages, std::vector<double> &x) {
    const int M = C.M, s = C.s();
    x.assign(C.nx(N), 0.0);
    for (int i = 0; i <= N; ++i)
        for (int c = 0; c < M; ++c) x[C.yoff(i) + c] = y[(size_t)i * M + c];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < s; ++j)
            for (int c = 0; c < M; ++c)
                x[C.koff(i, j) + c] = stages ? stages[((size_t)i * s + j) * M + c] : y[(size_t)(i + 1) * M + c];
}

struct FEval {
    std::vector<double> col;  // per interval (s+1) M: Phi_i, then G_i1..G_is
    std::vector<double> Y;    // per interval s M: stage points y_i + h sum_j a_rj K_j
    std::vector<double> bc;   // per BC row, input order
    std::vector<double> bcg;  // per BC row (s+2) M: an interior row's derivatives (y_iv, y_iv+1, K_iv)
};

// Phi! of the expanded form (FIRK/collocation.jl:59-93), in the reference's order
// of operations: Y_r = y_i + h (K a_r), G_r = f(Y_r) - K_r, Phi = (y_{i+1} - y_i) - h (K b)
double firk_evalsol(const FCtx &C, const double *mesh, int j, double t, int comp, const double *yi, const double *yj,
                    const double *K, bool expanded, double *g);

void firk_evaluate(const FCtx &C, int N, const double *mesh, const double *x, const std::vector<BcRow> &rows,
                   FEval &E) {
    const int M = C.M, s = C.s();
    E.col.assign((size_t)N * (s + 1) * M, 0.0);
    E.Y.assign((size_t)N * s * M, 0.0);
    std::vector<double> f(M);
    for (int i = 0; i < N; ++i) {
        const double h = mesh[i + 1] - mesh[i];
        const double *yi = x + C.yoff(i), *yj = x + C.yoff(i + 1);
        double *out = &E.col[(size_t)i * (s + 1) * M];
        for (int r = 0; r < s; ++r) {
            double *Yr = &E.Y[((size_t)i * s + r) * M];
            for (int c = 0; c < M; ++c) {
                double a = 0;
                for (int j = 0; j < s; ++j) a += x[C.koff(i, j) + c] * C.T.A[r][j];
                Yr[c] = h * a + yi[c];
            }
            f_eval(C.P, mesh[i] + C.T.c[r] * h, Yr, f.data());
            for (int c = 0; c < M; ++c) out[(size_t)(r + 1) * M + c] = f[c] - x[C.koff(i, r) + c];
        }
        for (int c = 0; c < M; ++c) {
            double a = 0;
            for (int j = 0; j < s; ++j) a += x[C.koff(i, j) + c] * C.T.b[j];
            out[c] = -h * a + (yj[c] - yi[c]);
        }
    }
    E.bc.assign(rows.size(), 0.0);
    E.bcg.assign(rows.size() * (s + 2) * M, 0.0);
    for (size_t k = 0; k < rows.size(); ++k) {
        const BcRow &r = rows[k];
        const double v = r.direct ? x[C.yoff(r.node) + r.comp]
                                  : firk_evalsol(C, mesh, r.iv, C.P.bc_t[k], r.comp, x + C.yoff(r.iv),
                                                 x + C.yoff(r.iv + 1), x + C.koff(r.iv, 0), true,
                                                 &E.bcg[k * (s + 2) * M]);
        E.bc[k] = v - C.P.bc_val[k];
    }
}

double firk_max_abs(const FEval &E) {
    double m = 0;
    for (double v : E.col) { if (std::isnan(v)) return NAN; m = std::max(m, std::fabs(v)); }
    for (double v : E.bc) { if (std::isnan(v)) return NAN; m = std::max(m, std::fabs(v)); }
    return m;
}
double firk_sumsq(const FEval &E) {
    double s = 0;
    for (double v : E.col) s += v * v;
    for (double v : E.bc) s += v * v;
    return s;
}

// reference layout: STANDARD [bc; collocation], TWO_POINT [bca; collocation; bcb]
void firk_flatten(const FCtx &C, const std::vector<BcRow> &rows, const FEval &E, double *out) {
    size_t o = 0;
    const bool tp = C.P.bc_kind == BVP_BC_TWO_POINT;
    for (size_t k = 0; k < rows.size(); ++k) if (!tp || !rows[k].side_b) out[o++] = E.bc[k];
    for (double v : E.col) out[o++] = v;
    if (tp) for (size_t k = 0; k < rows.size(); ++k) if (rows[k].side_b) out[o++] = E.bc[k];
}

// The Newton matrix as (row, col, value) entries with rows ordered so that it is
// banded (BC rows on node 0, then per interval Phi_i and G_i1..G_is, then BC rows on
// node N); rhs receives -F in the same order.  Row order does not change delta.
struct Trip { int r, c; double v; };
void firk_jacobian(const FCtx &C, int N, const double *mesh, const std::vector<BcRow> &rows, const FEval &E,
                   std::vector<Trip> &J, std::vector<double> &rhs) {
    const int M = C.M, s = C.s();
    J.clear();
    rhs.clear();
    int r = 0;
    auto bc_rows = [&](int node) {
        for (size_t k = 0; k < rows.size(); ++k)
            if (rows[k].direct && rows[k].node == node) {
                J.push_back({r, (int)C.yoff(node) + rows[k].comp, 1.0});
                rhs.push_back(-E.bc[k]);
                ++r;
            }
    };
    bc_rows(0);
    std::vector<double> Jr((size_t)M * M);
    for (int i = 0; i < N; ++i) {
        const double h = mesh[i + 1] - mesh[i];
        const double *col = &E.col[(size_t)i * (s + 1) * M];
        // interior STANDARD rows of interval i: derivatives of EvalSol wrt y_i, y_{i+1}, K_i
        for (size_t k = 0; k < rows.size(); ++k) {
            if (rows[k].direct || rows[k].iv != i) continue;
            const double *g = &E.bcg[k * (s + 2) * M];
            for (int c = 0; c < M; ++c) {
                J.push_back({r, (int)C.yoff(i) + c, g[c]});
                J.push_back({r, (int)C.yoff(i + 1) + c, g[M + c]});
            }
            for (int q = 0; q < s * M; ++q) J.push_back({r, (int)C.koff(i, 0) + q, g[2 * M + q]});
            rhs.push_back(-E.bc[k]);
            ++r;
        }
        for (int a = 0; a < M; ++a) {  // Phi_i
            J.push_back({r, (int)C.yoff(i) + a, -1.0});
            for (int j = 0; j < s; ++j) J.push_back({r, (int)C.koff(i, j) + a, -h * C.T.b[j]});
            J.push_back({r, (int)C.yoff(i + 1) + a, 1.0});
            rhs.push_back(-col[a]);
            ++r;
        }
        for (int q = 0; q < s; ++q) {  // G_iq
            J_eval(C.P, mesh[i] + C.T.c[q] * h, &E.Y[((size_t)i * s + q) * M], Jr.data());
            for (int a = 0; a < M; ++a) {
                for (int c = 0; c < M; ++c) J.push_back({r, (int)C.yoff(i) + c, Jr[(size_t)a * M + c]});
                for (int j = 0; j < s; ++j)
                    for (int c = 0; c < M; ++c) {
                        const double v = h * C.T.A[q][j] * Jr[(size_t)a * M + c] - ((j == q && c == a) ? 1.0 : 0.0);
                        J.push_back({r, (int)C.koff(i, j) + c, v});
                    }
                rhs.push_back(-col[(size_t)(q + 1) * M + a]);
                ++r;
            }
        }
    }
    if (N > 0) bc_rows(N);
}

// banded solve of the triplet system; delta in the unknown order
int firk_newton_dir(const FCtx &C, int N, const double *mesh, const std::vector<BcRow> &rows, const FEval &E,
                    double *delta) {
    std::vector<Trip> J;
    std::vector<double> rhs;
    firk_jacobian(C, N, mesh, rows, E, J, rhs);
    const int n = (int)C.nx(N);
    if ((int)rhs.size() != n) return 1;
    int kl = 0, ku = 0;
    for (const Trip &t : J) { kl = std::max(kl, t.r - t.c); ku = std::max(ku, t.c - t.r); }
    Band B;
    B.init(n, kl, ku);
    for (const Trip &t : J) B.at(t.r, t.c) += t.v;
    if (B.factor()) return 1;
    B.solve(rhs.data());
    std::memcpy(delta, rhs.data(), sizeof(double) * n);
    return 0;
}

// The expanded system as the nonlinear solver sees it: x -> F(x), the Newton
// direction and products with J and J' (rows in the banded order).
struct FirkSys {
    const FCtx &C;
    int N;
    const double *mesh;
    const std::vector<BcRow> &rows;
    FEval E;
    std::vector<Trip> J;
    std::vector<double> rhs;
    double eval(const double *x) {
        firk_evaluate(C, N, mesh, x, rows, E);
        return firk_max_abs(E);
    }
    double sumsq() const { return firk_sumsq(E); }
    int dir(double *d) { return firk_newton_dir(C, N, mesh, rows, E, d); }
    // J, -F at the last evaluation (for the trust region)
    void linearise() { firk_jacobian(C, N, mesh, rows, E, J, rhs); }
    void jmul(const double *v, double *out) const {
        std::fill(out, out + rhs.size(), 0.0);
        for (const Trip &t : J) out[t.r] += t.v * v[t.c];
    }
    void jtmul(const double *w, double *out, size_t n) const {
        std::fill(out, out + n, 0.0);
        for (const Trip &t : J) out[t.c] += t.v * w[t.r];
    }
};

int firk_newton_stage(FirkSys &S, std::vector<double> &x, double tol, int maxiters, bool linesearch, NlOut &out) {
    const size_t n = x.size();
    std::vector<double> delta(n), xt(n);
    for (int it = 0;; ++it) {
        out.fnorm = S.eval(x.data());
        if (!std::isfinite(out.fnorm)) return BVP_RET_UNSTABLE;
        if (out.fnorm <= tol) return BVP_RET_SUCCESS;
        if (it >= maxiters) return BVP_RET_MAXITERS;
        if (S.dir(delta.data())) return BVP_RET_FAILURE;
        out.iters++;
        double alpha = 1.0;
        if (linesearch) {
            const double ss = S.sumsq();
            auto phi = [&](double a) {
                for (size_t i = 0; i < n; ++i) xt[i] = x[i] + a * delta[i];
                S.eval(xt.data());
                return 0.5 * S.sumsq();
            };
            double pa;
            if (backtracking(phi, 0.5 * ss, -ss, &alpha, &pa)) return BVP_RET_FAILURE;
        }
        for (size_t i = 0; i < n; ++i) x[i] += alpha * delta[i];
    }
}

// TrustRegion stage on the expanded system (the MIRK restatement's rules above)
int firk_trust_region_stage(FirkSys &S, std::vector<double> &x, double tol, int maxiters, NlOut &out) {
    const size_t n = x.size();
    std::vector<double> dn(n), g(n), d(n), xt(n), F, Jg;
    double fy = S.eval(x.data()), ss = S.sumsq();
    double xmax = -INFINITY, xmin = INFINITY;
    for (double v : x) { xmax = std::max(xmax, v); xmin = std::min(xmin, v); }
    double radius = std::max(std::sqrt(ss), xmax - xmin) / 11;
    int shrinks = 0;
    bool fresh = true;
    double l2 = 0, jg2 = 0, dn2 = 0;
    for (int it = 0;; ++it) {
        out.fnorm = fy;
        if (!std::isfinite(out.fnorm)) return BVP_RET_UNSTABLE;
        if (out.fnorm <= tol) return BVP_RET_SUCCESS;
        if (it >= maxiters) return BVP_RET_MAXITERS;
        if (fresh) {
            S.eval(x.data());
            if (S.dir(dn.data())) return BVP_RET_FAILURE;
            dn2 = 0;
            for (double v : dn) dn2 += v * v;
            S.linearise();
            F.assign(S.rhs.size(), 0.0);
            for (size_t i = 0; i < F.size(); ++i) F[i] = -S.rhs[i];
            S.jtmul(F.data(), g.data(), n);
            l2 = 0;
            for (double v : g) l2 += v * v;
            Jg.assign(S.rhs.size(), 0.0);
            S.jmul(g.data(), Jg.data());
            jg2 = 0;
            for (double v : Jg) jg2 += v * v;
            fresh = false;
        }
        out.iters++;
        double pred;
        if (std::sqrt(dn2) <= radius) {
            d = dn;
            pred = ss;
        } else {
            pred = dogleg((int)n, dn.data(), g.data(), l2, jg2, ss, radius, d.data());
        }
        for (size_t i = 0; i < n; ++i) xt[i] = x[i] + d[i];
        const double ft = S.eval(xt.data());
        const double sst = S.sumsq();
        const double rho = (ss - sst) / pred;
        if (rho > 1e-4) {
            x = xt;
            ss = sst;
            fy = ft;
            fresh = true;
        }
        if (rho < 0.25) {
            radius *= 0.25;
            if (++shrinks > 32) { out.fnorm = fy; return BVP_RET_FAILURE; }
        } else {
            shrinks = 0;
            if (rho > 0.75) radius *= 2;
        }
    }
}

NlOut firk_newton(FirkSys &S, std::vector<double> &x, double tol, int maxiters, int stages) {
    const std::vector<double> x0 = x;
    NlOut out{BVP_RET_MAXITERS, 0, 0, -1};
    bool first = true;
    for (int st = 0; st < 3; ++st) {
        if (!(stages & (1 << st))) continue;
        if (!first) x = x0;
        first = false;
        out.stage = st;
        out.ret = st < 2 ? firk_newton_stage(S, x, tol, maxiters, st == 1, out)
                         : firk_trust_region_stage(S, x, tol, maxiters, out);
        if (out.ret == BVP_RET_SUCCESS) break;
    }
    return out;
}

// eval_q (FIRK/adaptivity.jl:581-606): the stage polynomial q(tau h) = y_i +
// sum_k coeff_k (tau h)^(k+1) and q', coeff_k = (Q K)_k / h^k
void firk_eval_q(const FCtx &C, const double *yi, double tau, double h, const double *x, int i, double *q,
                 double *qp) {
    const int M = C.M, s = C.s();
    for (int c = 0; c < M; ++c) {
        double coef[7];
        for (int k = 0; k < s; ++k) {
            double a = 0;
            for (int j = 0; j < s; ++j) a += C.T.Q[k][j] * x[C.koff(i, j) + c];
            coef[k] = a / std::pow(h, (double)k);
        }
        const double th = tau * h;
        double v = 0, dv = 0, p = 1;
        for (int k = 0; k < s; ++k) {
            dv += (k + 1) * coef[k] * p;
            p *= th;
            v += coef[k] * p;
        }
        q[c] = yi[c] + v;
        qp[c] = dv;
    }
}

// defect_estimate! (FIRK/adaptivity.jl:349-405)
double firk_defect(const FCtx &C, int N, const double *mesh, const double *x, std::vector<double> &errors) {
    const int M = C.M;
    std::vector<double> z(M), zp(M), f(M), d1(M), d2(M);
    errors.assign((size_t)N * M, 0.0);
    double dn = 0;
    const double tau = C.T.tau;
    for (int i = 0; i < N; ++i) {
        const double h = mesh[i + 1] - mesh[i];
        const double *yi = x + C.yoff(i);
        firk_eval_q(C, yi, tau, h, x, i, z.data(), zp.data());
        f_eval(C.P, mesh[i] + tau * h, z.data(), f.data());
        double e1 = 0;
        for (int c = 0; c < M; ++c) { d1[c] = (zp[c] - f[c]) / (std::fabs(f[c]) + 1.0); e1 = std::max(e1, std::fabs(d1[c])); }
        firk_eval_q(C, yi, 1.0 - tau, h, x, i, z.data(), zp.data());
        f_eval(C.P, mesh[i] + (1.0 - tau) * h, z.data(), f.data());
        double e2 = 0;
        for (int c = 0; c < M; ++c) { d2[c] = (zp[c] - f[c]) / (std::fabs(f[c]) + 1.0); e2 = std::max(e2, std::fabs(d2[c])); }
        const double *d = e1 > e2 ? d1.data() : d2.data();
        double ei = 0;
        for (int c = 0; c < M; ++c) { errors[(size_t)i * M + c] = d[c]; ei = std::max(ei, std::fabs(d[c])); }
        if (!(ei <= dn)) dn = ei;
    }
    return dn;
}

// 6 x 6 solve of the quintic through value and slope at 0, h, h/2
// (get_S_coeffs / s_constraints_block, FIRK/adaptivity.jl:183-227)
void quintic(double h, const double *bk, double *coef) {
    const double t[3] = {0.0, h, h / 2};
    double A[6][6], b[6];
    for (int r = 0; r < 3; ++r)
        for (int m = 0; m < 6; ++m) {
            A[r][m] = std::pow(t[r], (double)m);
            A[3 + r][m] = m == 0 ? 0.0 : m * std::pow(t[r], (double)(m - 1));
        }
    // rows: value 0, value h, value h/2, slope 0, slope h, slope h/2
    for (int r = 0; r < 6; ++r) b[r] = bk[r];
    for (int j = 0; j < 6; ++j) {
        int p = j;
        for (int i = j + 1; i < 6; ++i) if (std::fabs(A[i][j]) > std::fabs(A[p][j])) p = i;
        for (int c = 0; c < 6; ++c) std::swap(A[j][c], A[p][c]);
        std::swap(b[j], b[p]);
        for (int i = j + 1; i < 6; ++i) {
            const double m = A[i][j] / A[j][j];
            for (int c = j; c < 6; ++c) A[i][c] -= m * A[j][c];
            b[i] -= m * b[j];
        }
    }
    for (int i = 5; i >= 0; --i) {
        double v = b[i];
        for (int c = i + 1; c < 6; ++c) v -= A[i][c] * coef[c];
        coef[i] = v / A[i][i];
    }
}

// EvalSol of a FIRK cache at a point t inside interval j (FIRK/interpolation.jl:338-385
// expanded, 398-441 nested), component comp: the quintic through y_j, y_{j+1}, their
// slopes f and a third value/slope pair from the stage polynomial, evaluated at
// tau = t - t_j.  The expanded form takes that pair at (tau h) in the stage polynomial's
// variable (tau the absolute offset, as the reference writes it) from the state's stages;
// the nested form takes it at the midpoint, from stages the caller solved from zero (the
// inner problem's u0 there).  g (optional): the derivatives of the value, [M] wrt y_j,
// [M] wrt y_{j+1}, [s M] wrt K_j (zero in the nested form, whose inner solve sees y_j
// without its dual part, nodual_value).
double firk_evalsol(const FCtx &C, const double *mesh, int j, double t, int comp, const double *yi, const double *yj,
                    const double *K, bool expanded, double *g) {
    const int M = C.M, s = C.s();
    const double h = mesh[j + 1] - mesh[j], tau = t - mesh[j];
    const double th = (expanded ? tau : 0.5) * h;
    std::vector<double> fi(M), fj(M);
    f_eval(C.P, mesh[j], yi, fi.data());
    f_eval(C.P, mesh[j + 1], yj, fj.data());
    // stage polynomial of component comp (get_q_coeffs_interp / apply_q) and its
    // derivatives wrt the stages K_j.comp
    double v = 0, zp = 0, p = 1, dz[7] = {0}, dzp[7] = {0};
    for (int k = 0; k < s; ++k) {
        const double hk = std::pow(h, (double)k);
        double a = 0;
        for (int q = 0; q < s; ++q) a += C.T.Q[k][q] * K[q * M + comp];
        const double coef = a / hk;
        zp += (k + 1) * coef * p;
        for (int q = 0; q < s; ++q) dzp[q] += (k + 1) * (C.T.Q[k][q] / hk) * p;
        p *= th;
        v += coef * p;
        for (int q = 0; q < s; ++q) dz[q] += (C.T.Q[k][q] / hk) * p;
    }
    const double z = yi[comp] + v;
    const double bk[6] = {yi[comp], yj[comp], z, fi[comp], fj[comp], zp};
    double cf[6], val = 0;
    quintic(h, bk, cf);
    for (int m = 0; m < 6; ++m) val += cf[m] * std::pow(tau, (double)m);
    if (g) {
        // weights of the six conditions in the value (the quintic is linear in them)
        double w[6];
        for (int r = 0; r < 6; ++r) {
            double e[6] = {0, 0, 0, 0, 0, 0}, ce[6];
            e[r] = 1.0;
            quintic(h, e, ce);
            w[r] = 0;
            for (int m = 0; m < 6; ++m) w[r] += ce[m] * std::pow(tau, (double)m);
        }
        std::vector<double> Ji((size_t)M * M), Jj((size_t)M * M);
        J_eval(C.P, mesh[j], yi, Ji.data());
        J_eval(C.P, mesh[j + 1], yj, Jj.data());
        std::fill(g, g + (size_t)(s + 2) * M, 0.0);
        for (int c = 0; c < M; ++c) {
            g[c] = w[3] * Ji[(size_t)comp * M + c] + (c == comp ? w[0] + w[2] : 0.0);
            g[M + c] = w[4] * Jj[(size_t)comp * M + c] + (c == comp ? w[1] : 0.0);
        }
        if (expanded)
            for (int q = 0; q < s; ++q) g[2 * M + q * M + comp] = w[2] * dz[q] + w[5] * dzp[q];
    }
    return val;
}

// interpolant! of FIRKExpandInterpolation (FIRK/interpolation.jl:229-316), also
// interp_eval! of the transfer (FIRK/adaptivity.jl:8-48) without its h scaling
void firk_interp(const FCtx &C, int N, const double *mesh, const double *x, double t, double *u, double *du) {
    const int M = C.M;
    const int j = interval_of(mesh, N, t);
    const double h = mesh[j + 1] - mesh[j], tau = t - mesh[j];
    const double *yi = x + C.yoff(j), *yj = x + C.yoff(j + 1);
    std::vector<double> fi(M), fj(M), z(M), zp(M);
    f_eval(C.P, mesh[j], yi, fi.data());
    f_eval(C.P, mesh[j + 1], yj, fj.data());
    firk_eval_q(C, yi, 0.5, h, x, j, z.data(), zp.data());
    for (int c = 0; c < M; ++c) {
        const double bk[6] = {yi[c], yj[c], z[c], fi[c], fj[c], zp[c]};
        double coef[6];
        quintic(h, bk, coef);
        double v = 0, dv = 0;
        for (int m = 0; m < 6; ++m) {
            v += coef[m] * std::pow(tau, (double)m);
            if (m > 0) dv += m * coef[m] * std::pow(tau, (double)(m - 1));
        }
        u[c] = v;
        if (du) du[c] = dv;
    }
}

int firk_check(const bvp_instance &P, int code, FTab &T) {
    if (!firk_tableau(code, T)) return BVP_EINVAL;
    if (check_instance(P)) return BVP_EINVAL;
    return 0;
}

// state of an instance: mesh, nodes (guess/u0) and stages in the expanded order
int firk_state(const FCtx &C, const double *stages, std::vector<double> &mesh, std::vector<double> &x,
               std::vector<BcRow> &rows) {
    std::vector<double> y;
    int rc = instance_mesh_and_guess(C.P, mesh, y);
    if (rc) return rc;
    if ((rc = resolve_bc(C.P, C.P.N, mesh.data(), rows))) return rc;
    firk_extend(C, C.P.N, y.data(), stages, x);
    return 0;
}

int firk_solve_one(const bvp_instance &P, const bvp_opts &o, bvp_result &res) {
    FTab T;
    int rc = firk_check(P, o.method, T);
    if (rc) return rc;
    if (o.adaptive && !T.adaptive) return BVP_ENOTSUP;
    if (o.controller != BVP_CONTROL_DEFECT) return BVP_ENOTSUP;
    FCtx C(P, T);
    std::vector<double> mesh, x;
    std::vector<BcRow> rows;
    if ((rc = firk_state(C, nullptr, mesh, x, rows))) return rc;
    const double tol = o.newton_abstol > 0 ? o.newton_abstol : o.abstol;
    const int M = P.M, s = T.s;
    int N = P.N;
    int info = BVP_RET_SUCCESS, nl_ret = BVP_RET_SUCCESS;
    double defect_norm = 2 * o.abstol, last_defect = NAN;
    int outer = 0, newton_iters = 0, fb = 0;
    std::vector<double> errors, new_mesh;
    const int MAX_OUTER = 64;
    for (;;) {
        ++outer;
        if (resolve_bc(P, N, mesh.data(), rows)) return BVP_EINVAL;
        FirkSys S{C, N, mesh.data(), rows, {}, {}, {}};
        NlOut nl = firk_newton(S, x, tol, o.max_newton_iters, o.nlsolve_stages);
        newton_iters += nl.iters;
        if (nl.stage == 1) fb |= 1;
        if (nl.stage == 2) fb |= 16;
        nl_ret = nl.ret;
        defect_norm = 2 * o.abstol;
        info = nl.ret;
        if (!o.adaptive) break;
        if (info == BVP_RET_SUCCESS) {
            defect_norm = firk_defect(C, N, mesh.data(), x.data(), errors);
            last_defect = defect_norm;
            if (defect_norm > o.defect_threshold || std::isnan(defect_norm)) info = BVP_RET_FAILURE;
        }
        if (info == BVP_RET_SUCCESS) {
            if (defect_norm > o.abstol) {
                info = mesh_select(T.order, o.abstol, o.max_num_subintervals, mesh, errors, M, new_mesh);
                if (info == BVP_RET_SUCCESS) {
                    const int Nn = (int)new_mesh.size() - 1;
                    std::vector<double> yn((size_t)(Nn + 1) * M);
                    for (int k = 0; k <= Nn; ++k)
                        firk_interp(C, N, mesh.data(), x.data(), new_mesh[k], &yn[(size_t)k * M], nullptr);
                    mesh = new_mesh;
                    N = Nn;
                    firk_extend(C, N, yn.data(), nullptr, x);
                }
            }
        } else {
            if (2 * N > o.max_num_subintervals) {
                info = BVP_RET_FAILURE;
            } else {
                std::vector<double> hm;
                half_mesh(mesh, hm);
                mesh = hm;
                N = (int)mesh.size() - 1;
                x.assign(C.nx(N), 0.0);
                info = BVP_RET_SUCCESS;
            }
        }
        if (!(info == BVP_RET_SUCCESS && defect_norm > o.abstol)) break;
        if (outer >= MAX_OUTER) { info = BVP_RET_FAILURE; break; }
    }
    res.retcode = (nl_ret == BVP_RET_SUCCESS) ? info : nl_ret;
    res.N_final = N;
    res.newton_iters = newton_iters;
    res.outer_iters = outer;
    res.fallback_used = fb;
    res.defect_norm = last_defect;
    res.err_vs_fp64 = NAN;
    if (resolve_bc(P, N, mesh.data(), rows) == 0) {
        FEval E;
        firk_evaluate(C, N, mesh.data(), x.data(), rows, E);
        res.resid_norm = firk_max_abs(E);
    }
    if (res.capacity < N + 1) return BVP_EINVAL;
    if (res.mesh) std::memcpy(res.mesh, mesh.data(), sizeof(double) * (N + 1));
    if (res.u)
        for (int i = 0; i <= N; ++i)
            for (int c = 0; c < M; ++c) res.u[(size_t)i * M + c] = x[C.yoff(i) + c];
    if (res.stages)
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < s; ++j)
                for (int c = 0; c < M; ++c) res.stages[((size_t)i * s + j) * M + c] = x[C.koff(i, j) + c];
    return 0;
}

// ---------------------------------------------------------------- nested form
// nested_nlsolve = true (FIRK/firk.jl:114-288, FIRK/collocation.jl:140-428): the
// unknowns are the node values only; every evaluation solves each interval's stage
// equations G(K) = 0 for K from the fixed start K0 (__K0_on_u0, FIRK/utils.jl:41-83:
// u0 repeated per stage, or the mean of a vector guess) and the residual is
// Phi_i = y_{i+1} - y_i - h sum_j b_j K_j(y_i).  The inner solve restates
// NonlinearSolve's NewtonRaphson at its default tolerance for Float64,
// max|G| <= eps^(4/5) (or nested_nlsolve_kwargs.abstol, bvp_opts.nest_abstol), at most
// 1000 steps; its Jacobian enters the outer one by the
// implicit function theorem, dK/dy_i = W^-1 [J_1; ...; J_s] (what ForwardDiff through
// the nested solve computes).
bool firk_is_nested(int code) { return code >= 1000; }

// K0 of an instance (M values)
void firk_k0(const bvp_instance &P, std::vector<double> &k0) {
    const int M = P.M;
    k0.assign(M, 0.0);
    if (P.u0) {
        for (int c = 0; c < M; ++c) k0[c] = P.u0[c];
    } else if (P.guess) {
        for (int k = 0; k <= P.N; ++k)
            for (int c = 0; c < M; ++c) k0[c] += P.guess[(size_t)k * M + c];
        for (int c = 0; c < M; ++c) k0[c] /= (P.N + 1);
    }
}

// W = I - h (A (x) J) blockwise at the stage points of K (s M x s M, row-major), the
// stage residuals G and the stacked stage Jacobians Jh (s M x M)
void firk_stage_system(const FCtx &C, double ti, double h, const double *yi, const double *K, std::vector<double> &W,
                       std::vector<double> &G, std::vector<double> &Jh) {
    const int M = C.M, s = C.s(), n = s * M;
    W.assign((size_t)n * n, 0.0);
    G.assign(n, 0.0);
    Jh.assign((size_t)n * M, 0.0);
    std::vector<double> Y(M), f(M), J((size_t)M * M);
    for (int r = 0; r < s; ++r) {
        for (int c = 0; c < M; ++c) {
            double a = 0;
            for (int j = 0; j < s; ++j) a += K[j * M + c] * C.T.A[r][j];
            Y[c] = h * a + yi[c];
        }
        const double t = ti + C.T.c[r] * h;
        f_eval(C.P, t, Y.data(), f.data());
        J_eval(C.P, t, Y.data(), J.data());
        for (int a = 0; a < M; ++a) {
            G[r * M + a] = f[a] - K[r * M + a];
            for (int c = 0; c < M; ++c) {
                Jh[(size_t)(r * M + a) * M + c] = J[(size_t)a * M + c];
                for (int j = 0; j < s; ++j)
                    W[(size_t)(r * M + a) * n + j * M + c] =
                        ((r == j && a == c) ? 1.0 : 0.0) - h * C.T.A[r][j] * J[(size_t)a * M + c];
            }
        }
    }
}

// dense solve W X = B (n x n, nr right-hand sides, row-major), partial pivoting
bool dense_lu_solve(int n, int nr, std::vector<double> W, std::vector<double> &B) {
    for (int j = 0; j < n; ++j) {
        int p = j;
        for (int i = j + 1; i < n; ++i) if (std::fabs(W[(size_t)i * n + j]) > std::fabs(W[(size_t)p * n + j])) p = i;
        if (W[(size_t)p * n + j] == 0.0) return false;
        if (p != j) {
            for (int c = 0; c < n; ++c) std::swap(W[(size_t)j * n + c], W[(size_t)p * n + c]);
            for (int c = 0; c < nr; ++c) std::swap(B[(size_t)j * nr + c], B[(size_t)p * nr + c]);
        }
        for (int i = j + 1; i < n; ++i) {
            const double m = W[(size_t)i * n + j] / W[(size_t)j * n + j];
            for (int c = j + 1; c < n; ++c) W[(size_t)i * n + c] -= m * W[(size_t)j * n + c];
            for (int c = 0; c < nr; ++c) B[(size_t)i * nr + c] -= m * B[(size_t)j * nr + c];
        }
    }
    for (int i = n - 1; i >= 0; --i)
        for (int c = 0; c < nr; ++c) {
            double v = B[(size_t)i * nr + c];
            for (int k = i + 1; k < n; ++k) v -= W[(size_t)i * n + k] * B[(size_t)k * nr + c];
            B[(size_t)i * nr + c] = v / W[(size_t)i * n + i];
        }
    return true;
}

// The interval's stage values K (s M) from K0: NewtonRaphson, K <- K + W^-1 G until
// max|G| <= C.itol (checked before every step; NaN stops), at most 1000 steps or a
// singular W.
void firk_inner_solve(const FCtx &C, double ti, double h, const double *yi, const std::vector<double> &k0,
                      double *K) {
    const int M = C.M, s = C.s(), n = s * M;
    for (int r = 0; r < s; ++r)
        for (int c = 0; c < M; ++c) K[r * M + c] = k0[c];
    std::vector<double> W, G, Jh;
    for (int it = 0; it < 1000; ++it) {
        firk_stage_system(C, ti, h, yi, K, W, G, Jh);
        double mx = 0;
        bool bad = false;
        for (double v : G) { bad |= std::isnan(v); mx = std::max(mx, std::fabs(v)); }
        if (bad || !(mx > C.itol)) return;
        if (!dense_lu_solve(n, 1, W, G)) return;
        for (int q = 0; q < n; ++q) K[q] += G[q];
    }
}

struct FNEval {
    std::vector<double> phi;  // [N][M]
    std::vector<double> K;    // [N][s][M] converged stages
    std::vector<double> bc;
    std::vector<double> bcg;  // per BC row (s+2) M: an interior row's derivatives (y_iv, y_iv+1, -)
};

void firk_nested_evaluate(const FCtx &C, int N, const double *mesh, const double *y, const std::vector<double> &k0,
                          const std::vector<BcRow> &rows, FNEval &E) {
    const int M = C.M, s = C.s();
    E.phi.assign((size_t)N * M, 0.0);
    E.K.assign((size_t)N * s * M, 0.0);
    for (int i = 0; i < N; ++i) {
        const double h = mesh[i + 1] - mesh[i];
        double *K = &E.K[(size_t)i * s * M];
        firk_inner_solve(C, mesh[i], h, y + (size_t)i * M, k0, K);
        for (int c = 0; c < M; ++c) {
            double a = 0;
            for (int j = 0; j < s; ++j) a += K[j * M + c] * C.T.b[j];
            E.phi[(size_t)i * M + c] = -h * a + (y[(size_t)(i + 1) * M + c] - y[(size_t)i * M + c]);
        }
    }
    E.bc.assign(rows.size(), 0.0);
    E.bcg.assign(rows.size() * (s + 2) * M, 0.0);
    std::vector<double> zero(M, 0.0), Kz((size_t)s * M);
    for (size_t k = 0; k < rows.size(); ++k) {
        const BcRow &r = rows[k];
        double v;
        if (r.direct) {
            v = y[(size_t)r.node * M + r.comp];
        } else {
            // EvalSol's own inner solve, from zero (FIRK/interpolation.jl:432-434)
            const int i = r.iv;
            firk_inner_solve(C, mesh[i], mesh[i + 1] - mesh[i], y + (size_t)i * M, zero, Kz.data());
            v = firk_evalsol(C, mesh, i, C.P.bc_t[k], r.comp, y + (size_t)i * M, y + (size_t)(i + 1) * M, Kz.data(),
                             false, &E.bcg[k * (s + 2) * M]);
        }
        E.bc[k] = v - C.P.bc_val[k];
    }
}

double firk_nested_max_abs(const FNEval &E) {
    double m = 0;
    for (double v : E.phi) { if (std::isnan(v)) return NAN; m = std::max(m, std::fabs(v)); }
    for (double v : E.bc) { if (std::isnan(v)) return NAN; m = std::max(m, std::fabs(v)); }
    return m;
}
double firk_nested_sumsq(const FNEval &E) {
    double s = 0;
    for (double v : E.phi) s += v * v;
    for (double v : E.bc) s += v * v;
    return s;
}

// the Newton matrix of the nested residual (rows: BC rows on node 0, Phi_i, BC rows on
// node N) with S_i = -(I + h sum_j b_j KJ_j), KJ = W^-1 [J_r] at the converged stages
void firk_nested_jacobian(const FCtx &C, int N, const double *mesh, const double *y, const std::vector<BcRow> &rows,
                          const FNEval &E, std::vector<Trip> &J, std::vector<double> &rhs) {
    const int M = C.M, s = C.s(), n = s * M;
    J.clear();
    rhs.clear();
    int r = 0;
    auto bc_rows = [&](int node) {
        for (size_t k = 0; k < rows.size(); ++k)
            if (rows[k].direct && rows[k].node == node) {
                J.push_back({r, node * M + rows[k].comp, 1.0});
                rhs.push_back(-E.bc[k]);
                ++r;
            }
    };
    bc_rows(0);
    std::vector<double> W, G, Jh;
    for (int i = 0; i < N; ++i) {
        const double h = mesh[i + 1] - mesh[i];
        for (size_t k = 0; k < rows.size(); ++k) {  // interior STANDARD rows of interval i
            if (rows[k].direct || rows[k].iv != i) continue;
            const double *g = &E.bcg[k * (s + 2) * M];
            for (int c = 0; c < M; ++c) {
                J.push_back({r, i * M + c, g[c]});
                J.push_back({r, (i + 1) * M + c, g[M + c]});
            }
            rhs.push_back(-E.bc[k]);
            ++r;
        }
        firk_stage_system(C, mesh[i], h, y + (size_t)i * M, &E.K[(size_t)i * s * M], W, G, Jh);
        dense_lu_solve(n, M, W, Jh);  // Jh <- KJ
        for (int a = 0; a < M; ++a) {
            for (int c = 0; c < M; ++c) {
                double v = 0;
                for (int j = 0; j < s; ++j) v += C.T.b[j] * Jh[(size_t)(j * M + a) * M + c];
                J.push_back({r, i * M + c, -((a == c) ? 1.0 : 0.0) - h * v});
            }
            J.push_back({r, (i + 1) * M + a, 1.0});
            rhs.push_back(-E.phi[(size_t)i * M + a]);
            ++r;
        }
    }
    if (N > 0) bc_rows(N);
}

int firk_nested_dir(const FCtx &C, int N, const double *mesh, const double *y, const std::vector<BcRow> &rows,
                    const FNEval &E, double *delta) {
    std::vector<Trip> J;
    std::vector<double> rhs;
    firk_nested_jacobian(C, N, mesh, y, rows, E, J, rhs);
    const int n = C.M * (N + 1);
    if ((int)rhs.size() != n) return 1;
    int kl = 0, ku = 0;
    for (const Trip &t : J) { kl = std::max(kl, t.r - t.c); ku = std::max(ku, t.c - t.r); }
    Band B;
    B.init(n, kl, ku);
    for (const Trip &t : J) B.at(t.r, t.c) += t.v;
    if (B.factor()) return 1;
    B.solve(rhs.data());
    std::memcpy(delta, rhs.data(), sizeof(double) * n);
    return 0;
}

// the NewtonRaphson stages on the nested residual (the outer solve)
int firk_nested_newton_stage(const FCtx &C, int N, const double *mesh, std::vector<double> &y,
                             const std::vector<double> &k0, const std::vector<BcRow> &rows, double tol, int maxiters,
                             bool linesearch, NlOut &out, FNEval &E) {
    const size_t n = y.size();
    std::vector<double> delta(n), yt(n);
    FNEval Et;
    for (int it = 0;; ++it) {
        firk_nested_evaluate(C, N, mesh, y.data(), k0, rows, E);
        out.fnorm = firk_nested_max_abs(E);
        if (!std::isfinite(out.fnorm)) return BVP_RET_UNSTABLE;
        if (out.fnorm <= tol) return BVP_RET_SUCCESS;
        if (it >= maxiters) return BVP_RET_MAXITERS;
        if (firk_nested_dir(C, N, mesh, y.data(), rows, E, delta.data())) return BVP_RET_FAILURE;
        out.iters++;
        double alpha = 1.0;
        if (linesearch) {
            const double ss = firk_nested_sumsq(E);
            auto phi = [&](double a) {
                for (size_t i = 0; i < n; ++i) yt[i] = y[i] + a * delta[i];
                firk_nested_evaluate(C, N, mesh, yt.data(), k0, rows, Et);
                return 0.5 * firk_nested_sumsq(Et);
            };
            double pa;
            if (backtracking(phi, 0.5 * ss, -ss, &alpha, &pa)) return BVP_RET_FAILURE;
        }
        for (size_t i = 0; i < n; ++i) y[i] += alpha * delta[i];
    }
}

// the quintic interpolant of interval j at t with stages K (s M) and step hq (the
// transfer passes the reference's scaled step, FIRK/adaptivity.jl:98-104)
void firk_quintic_at(const FCtx &C, const double *mesh, const double *y, int j, const double *K, double hq, double t,
                     double *u, double *du) {
    const int M = C.M, s = C.s();
    const double tau = t - mesh[j];
    const double *yi = y + (size_t)j * M, *yj = y + (size_t)(j + 1) * M;
    std::vector<double> fi(M), fj(M), z(M), zp(M);
    f_eval(C.P, mesh[j], yi, fi.data());
    f_eval(C.P, mesh[j + 1], yj, fj.data());
    for (int c = 0; c < M; ++c) {
        double coef[7];
        for (int k = 0; k < s; ++k) {
            double a = 0;
            for (int jj = 0; jj < s; ++jj) a += C.T.Q[k][jj] * K[jj * M + c];
            coef[k] = a / std::pow(hq, (double)k);
        }
        const double th = 0.5 * hq;
        double v = 0, dv = 0, p = 1;
        for (int k = 0; k < s; ++k) {
            dv += (k + 1) * coef[k] * p;
            p *= th;
            v += coef[k] * p;
        }
        z[c] = yi[c] + v;
        zp[c] = dv;
    }
    for (int c = 0; c < M; ++c) {
        const double bk[6] = {yi[c], yj[c], z[c], fi[c], fj[c], zp[c]};
        double cf[6];
        quintic(hq, bk, cf);
        double v = 0, dv = 0;
        for (int m = 0; m < 6; ++m) {
            v += cf[m] * std::pow(tau, (double)m);
            if (m > 0) dv += m * cf[m] * std::pow(tau, (double)(m - 1));
        }
        u[c] = v;
        if (du) du[c] = dv;
    }
}

int firk_nested_solve_one(const bvp_instance &P, const bvp_opts &o, bvp_result &res) {
    FTab T;
    int rc = firk_check(P, o.method, T);
    if (rc) return rc;
    if (o.adaptive && !T.adaptive) return BVP_ENOTSUP;
    if (o.controller != BVP_CONTROL_DEFECT) return BVP_ENOTSUP;
    FCtx C(P, T);
    if (o.nest_abstol > 0) C.itol = o.nest_abstol;
    std::vector<double> mesh, y, k0;
    if ((rc = instance_mesh_and_guess(P, mesh, y))) return rc;
    firk_k0(P, k0);
    const double tol = o.newton_abstol > 0 ? o.newton_abstol : o.abstol;
    const int M = P.M, s = T.s;
    int N = P.N;
    int info = BVP_RET_SUCCESS, nl_ret = BVP_RET_SUCCESS;
    double defect_norm = 2 * o.abstol, last_defect = NAN;
    int outer = 0, newton_iters = 0, fb = 0;
    std::vector<double> errors, new_mesh, x;
    std::vector<BcRow> rows;
    FNEval E;
    const int MAX_OUTER = 64;
    bool eval_valid = false;
    for (;;) {
        ++outer;
        if (resolve_bc(P, N, mesh.data(), rows)) return BVP_EINVAL;
        const std::vector<double> y0 = y;
        NlOut nl{BVP_RET_MAXITERS, 0, 0, -1};
        bool first = true;
        for (int st = 0; st < 2; ++st) {
            if (!(o.nlsolve_stages & (1 << st))) continue;
            if (!first) y = y0;
            first = false;
            nl.stage = st;
            nl.ret = firk_nested_newton_stage(C, N, mesh.data(), y, k0, rows, tol, o.max_newton_iters, st == 1, nl, E);
            if (nl.ret == BVP_RET_SUCCESS) break;
        }
        eval_valid = !first;  // E holds the last stage's state (none ran: no NewtonRaphson stage asked for)
        if (first) firk_nested_evaluate(C, N, mesh.data(), y.data(), k0, rows, E);
        newton_iters += nl.iters;
        if (nl.stage == 1) fb |= 1;
        nl_ret = nl.ret;
        defect_norm = 2 * o.abstol;
        info = nl.ret;
        if (!o.adaptive) break;
        firk_extend(C, N, y.data(), E.K.data(), x);  // nodes and the converged stages
        if (info == BVP_RET_SUCCESS) {
            defect_norm = firk_defect(C, N, mesh.data(), x.data(), errors);
            last_defect = defect_norm;
            if (defect_norm > o.defect_threshold || std::isnan(defect_norm)) info = BVP_RET_FAILURE;
        }
        if (info == BVP_RET_SUCCESS) {
            if (defect_norm > o.abstol) {
                info = mesh_select(T.order, o.abstol, o.max_num_subintervals, mesh, errors, M, new_mesh);
                if (info == BVP_RET_SUCCESS) {
                    // interp_eval! (FIRK/adaptivity.jl:90-135): the old interpolant with h scaled
                    // by the growth factor of the node count when the mesh grows
                    const int Nn = (int)new_mesh.size() - 1;
                    const double lf = (double)Nn / (double)N;
                    std::vector<double> yn((size_t)(Nn + 1) * M);
                    for (int k = 0; k <= Nn; ++k) {
                        const int j = interval_of(mesh.data(), N, new_mesh[k]);
                        const double hq = (mesh[j + 1] - mesh[j]) * (lf > 1 ? lf : 1.0);
                        firk_quintic_at(C, mesh.data(), y.data(), j, &E.K[(size_t)j * s * M], hq, new_mesh[k],
                                        &yn[(size_t)k * M], nullptr);
                    }
                    mesh = new_mesh;
                    N = Nn;
                    y = yn;
                    eval_valid = false;
                }
            }
        } else {
            if (2 * N > o.max_num_subintervals) {
                info = BVP_RET_FAILURE;
            } else {
                std::vector<double> hm;
                half_mesh(mesh, hm);
                mesh = hm;
                N = (int)mesh.size() - 1;
                y.assign((size_t)(N + 1) * M, 0.0);
                eval_valid = false;
                info = BVP_RET_SUCCESS;
            }
        }
        if (!(info == BVP_RET_SUCCESS && defect_norm > o.abstol)) break;
        if (outer >= MAX_OUTER) { info = BVP_RET_FAILURE; break; }
    }
    res.retcode = (nl_ret == BVP_RET_SUCCESS) ? info : nl_ret;
    res.N_final = N;
    res.newton_iters = newton_iters;
    res.outer_iters = outer;
    res.fallback_used = fb;
    res.defect_norm = last_defect;
    res.err_vs_fp64 = NAN;
    if (resolve_bc(P, N, mesh.data(), rows) == 0) {
        if (!eval_valid) firk_nested_evaluate(C, N, mesh.data(), y.data(), k0, rows, E);
        res.resid_norm = firk_nested_max_abs(E);
    }
    if (res.capacity < N + 1) return BVP_EINVAL;
    if (res.mesh) std::memcpy(res.mesh, mesh.data(), sizeof(double) * (N + 1));
    if (res.u) std::memcpy(res.u, y.data(), sizeof(double) * (size_t)M * (N + 1));
    if (res.stages) std::memcpy(res.stages, E.K.data(), sizeof(double) * (size_t)N * s * M);
    return 0;
}

// [bc; Phi] (STANDARD) or [bca; Phi; bcb] (TWO_POINT) of node-layout residuals
void flatten_residual_phi(const bvp_instance &P, const std::vector<BcRow> &rows, const std::vector<double> &phi,
                          const std::vector<double> &bc, double *out) {
    size_t o = 0;
    const bool tp = P.bc_kind == BVP_BC_TWO_POINT;
    for (size_t k = 0; k < rows.size(); ++k) if (!tp || !rows[k].side_b) out[o++] = bc[k];
    for (double v : phi) out[o++] = v;
    if (tp) for (size_t k = 0; k < rows.size(); ++k) if (rows[k].side_b) out[o++] = bc[k];
}

// nested codes (BVP_FIRK_NESTED): the state is (mesh, guess); stages are ignored (the
// inner solves compute them) and the vectors have the node layout M (N + 1)
int firk_nested_state(const FCtx &C, std::vector<double> &mesh, std::vector<double> &y, std::vector<double> &k0,
                      std::vector<BcRow> &rows, FNEval &E) {
    int rc = instance_mesh_and_guess(C.P, mesh, y);
    if (rc) return rc;
    if ((rc = resolve_bc(C.P, C.P.N, mesh.data(), rows))) return rc;
    firk_k0(C.P, k0);
    firk_nested_evaluate(C, C.P.N, mesh.data(), y.data(), k0, rows, E);
    return 0;
}

int solve_any(const bvp_instance &P, const bvp_opts &o, bvp_result &res) {
    if (o.method >= 1000) return firk_nested_solve_one(P, o, res);
    return o.method >= 100 ? firk_solve_one(P, o, res) : solve_one(P, o, res);
}

// banded row order of firk_jacobian / firk_nested_jacobian -> the reference's row order
// ([bc; collocation] or [bca; collocation; bcb]); per_iv collocation rows per interval
std::vector<int> firk_row_map(const std::vector<BcRow> &rows, int N, int per_iv, bool tp) {
    std::vector<int> bc_ref(rows.size());
    int r = 0;
    for (size_t k = 0; k < rows.size(); ++k) if (!tp || !rows[k].side_b) bc_ref[k] = r++;
    const int top = r;
    r += N * per_iv;
    if (tp) for (size_t k = 0; k < rows.size(); ++k) if (rows[k].side_b) bc_ref[k] = r++;
    std::vector<int> map;
    for (size_t k = 0; k < rows.size(); ++k) if (rows[k].direct && rows[k].node == 0) map.push_back(bc_ref[k]);
    for (int i = 0; i < N; ++i) {
        for (size_t k = 0; k < rows.size(); ++k) if (!rows[k].direct && rows[k].iv == i) map.push_back(bc_ref[k]);
        for (int q = 0; q < per_iv; ++q) map.push_back(top + i * per_iv + q);
    }
    for (size_t k = 0; k < rows.size(); ++k) if (rows[k].direct && rows[k].node == N) map.push_back(bc_ref[k]);
    return map;
}

}  // namespace

// ======================================================================
// C ABI of the oracle (ctypes from tests/ and the benchmark's CPU leg)
// ======================================================================
extern "C" {

void ref_opts_default(bvp_opts *o) {
    o->method = BVP_MIRK4;
    o->abstol = 1e-6;
    o->newton_abstol = 0;
    o->adaptive = 1;
    o->defect_threshold = 0.1;
    o->max_num_subintervals = 3000;
    o->max_newton_iters = 1000;
    o->nlsolve_stages = 7;
    o->transfer_mode = BVP_TRANSFER_REFERENCE;
    o->precision = BVP_PREC_FP64;
    o->refine_iters = 0;
    o->warm_start = BVP_WARM_NONE;
    o->report_err = 0;
    o->controller = BVP_CONTROL_DEFECT;
    o->global_method = BVP_GLOBAL_HO;
    o->hybrid_de = 1.0;
    o->hybrid_ge = 1.0;
    o->nest_abstol = 0.0;
}

// The global error estimate of opts.global_method at (mesh, guess): per-interval
// error vectors (N*M) and the norm the controller compares with abstol.
int ref_global_error(const bvp_opts *o, const bvp_instance *P, double *errors, double *norm) {
    int rc = check_instance(*P);
    if (rc) return rc;
    if (o->global_method == BVP_GLOBAL_HO && o->method != BVP_MIRK4) return BVP_ENOTSUP;
    Ctx C(*P, o->method);
    std::vector<double> mesh, y, e;
    if ((rc = instance_mesh_and_guess(*P, mesh, y))) return rc;
    const double tol = o->newton_abstol > 0 ? o->newton_abstol : o->abstol;
    *norm = global_error_estimate(*P, *o, C, mesh, y, tol, e);
    std::memcpy(errors, e.data(), sizeof(double) * e.size());
    return 0;
}

// BackTracking on the polynomial merit function phi(a) = sum_k coef[k] a^k
// (phi(0) = coef[0], phi'(0) = coef[1]).  trace (room for max_trace) receives
// the trial steps in order, *n_trace their number.  Returns 0, or 1 when the
// reductions ran out.
int ref_backtracking_poly(const double *coef, int ncoef, double *alpha, double *trace, int max_trace,
                          int *n_trace) {
    int nt = 0;
    auto phi = [&](double a) {
        double v = 0, p = 1;
        for (int k = 0; k < ncoef; ++k) { v += coef[k] * p; p *= a; }
        if (nt < max_trace) trace[nt] = a;
        ++nt;
        return v;
    };
    double pa;
    const int rc = backtracking(phi, coef[0], coef[1], alpha, &pa);
    *n_trace = nt;
    return rc;
}

// Dogleg step for a dense system (J n x n row-major, F) of radius `radius`:
// d and the predicted reduction |F|^2 - |F + J d|^2 of the linear model.
int ref_dogleg(int n, const double *J, const double *F, double radius, double *d, double *pred) {
    std::vector<double> A(J, J + (size_t)n * n), dn(n), g(n, 0.0), Jg(n, 0.0);
    // Newton step by Gaussian elimination with partial pivoting
    for (int i = 0; i < n; ++i) dn[i] = -F[i];
    for (int j = 0; j < n; ++j) {
        int p = j;
        for (int i = j + 1; i < n; ++i) if (std::fabs(A[(size_t)i * n + j]) > std::fabs(A[(size_t)p * n + j])) p = i;
        if (A[(size_t)p * n + j] == 0.0) return 1;
        for (int c = 0; c < n; ++c) std::swap(A[(size_t)j * n + c], A[(size_t)p * n + c]);
        std::swap(dn[j], dn[p]);
        for (int i = j + 1; i < n; ++i) {
            const double m = A[(size_t)i * n + j] / A[(size_t)j * n + j];
            for (int c = j; c < n; ++c) A[(size_t)i * n + c] -= m * A[(size_t)j * n + c];
            dn[i] -= m * dn[j];
        }
    }
    for (int j = n - 1; j >= 0; --j) {
        for (int c = j + 1; c < n; ++c) dn[j] -= A[(size_t)j * n + c] * dn[c];
        dn[j] /= A[(size_t)j * n + j];
    }
    double ss = 0, dn2 = 0, l2 = 0, jg2 = 0;
    for (int i = 0; i < n; ++i) { ss += F[i] * F[i]; dn2 += dn[i] * dn[i]; }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) g[i] += J[(size_t)k * n + i] * F[k];
    for (int i = 0; i < n; ++i) l2 += g[i] * g[i];
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < n; ++k) Jg[i] += J[(size_t)i * n + k] * g[k];
        jg2 += Jg[i] * Jg[i];
    }
    if (std::sqrt(dn2) <= radius) {
        for (int i = 0; i < n; ++i) d[i] = dn[i];
        *pred = ss;
        return 0;
    }
    *pred = dogleg(n, dn.data(), g.data(), l2, jg2, ss, radius, d);
    return 0;
}

void ref_uniform_mesh(double t0, double t1, int N, double *mesh) { uniform_mesh(t0, t1, N, mesh); }

int ref_interval(const double *mesh, int N, double t) { return interval_of(mesh, N, t); }

void ref_interp_weights(int method, double tau, double *w, double *wp) { interp_weights(method, tau, w, wp); }

// tableau export: s, sstar, order, c[5], v[5], b[5], x[25], cs[4], vs[4], xs[36], tau
int ref_tableau(int method, int *ints, double *dbl) {
    const Tab &T = tableau(method);
    ints[0] = T.s; ints[1] = T.sstar; ints[2] = T.order;
    int o = 0;
    for (int i = 0; i < 5; ++i) dbl[o++] = T.c[i];
    for (int i = 0; i < 5; ++i) dbl[o++] = T.v[i];
    for (int i = 0; i < 5; ++i) dbl[o++] = T.b[i];
    for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) dbl[o++] = T.x[i][j];
    for (int i = 0; i < 4; ++i) dbl[o++] = T.cs[i];
    for (int i = 0; i < 4; ++i) dbl[o++] = T.vs[i];
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 9; ++j) dbl[o++] = T.xs[i][j];
    dbl[o++] = T.tau;
    return o;
}

void ref_f(const bvp_instance *P, double t, const double *y, double *out) { f_eval(*P, t, y, out); }

int ref_residual(int method, const bvp_instance *P, double *resid) {
    int rc = check_instance(*P);
    if (rc) return rc;
    Ctx C(*P, method);
    std::vector<double> mesh, y;
    if ((rc = instance_mesh_and_guess(*P, mesh, y))) return rc;
    std::vector<BcRow> rows;
    if ((rc = resolve_bc(*P, P->N, mesh.data(), rows))) return rc;
    Eval E;
    evaluate(C, P->N, mesh.data(), y.data(), rows, E, false);
    flatten_residual(C, P->N, rows, E, resid);
    return 0;
}

int ref_jacobian_blocks(int method, const bvp_instance *P, double *S, double *R, int32_t *bc_node) {
    int rc = check_instance(*P);
    if (rc) return rc;
    Ctx C(*P, method);
    std::vector<double> mesh, y;
    if ((rc = instance_mesh_and_guess(*P, mesh, y))) return rc;
    std::vector<BcRow> rows;
    if ((rc = resolve_bc(*P, P->N, mesh.data(), rows))) return rc;
    Eval E;
    evaluate(C, P->N, mesh.data(), y.data(), rows, E, false);
    const int M = P->M, MM = M * M;
    for (int i = 0; i < P->N; ++i)
        interval_jac(C, mesh[i], mesh[i + 1] - mesh[i], &E.Y[(size_t)i * C.T.s * M], S + (size_t)i * MM,
                     R + (size_t)i * MM);
    for (size_t k = 0; k < rows.size(); ++k) bc_node[k] = rows[k].node;
    return 0;
}

// One Newton correction delta solving J delta = -F (point-major, M*(N+1)).
int ref_newton_step(int method, const bvp_instance *P, double *delta) {
    int rc = check_instance(*P);
    if (rc) return rc;
    Ctx C(*P, method);
    std::vector<double> mesh, y;
    if ((rc = instance_mesh_and_guess(*P, mesh, y))) return rc;
    std::vector<BcRow> rows;
    if ((rc = resolve_bc(*P, P->N, mesh.data(), rows))) return rc;
    Eval E;
    evaluate(C, P->N, mesh.data(), y.data(), rows, E, false);
    Band B;
    if (newton_solve(C, P->N, mesh.data(), y.data(), rows, E, delta, B)) return 1;
    return 0;
}

// Dense check path: assemble the full Newton matrix (reference row layout) as
// a dense n*n row-major array plus the residual.  For small N only.
int ref_dense_system(int method, const bvp_instance *P, double *J, double *F) {
    int rc = check_instance(*P);
    if (rc) return rc;
    Ctx C(*P, method);
    std::vector<double> mesh, y;
    if ((rc = instance_mesh_and_guess(*P, mesh, y))) return rc;
    std::vector<BcRow> rows;
    if ((rc = resolve_bc(*P, P->N, mesh.data(), rows))) return rc;
    Eval E;
    evaluate(C, P->N, mesh.data(), y.data(), rows, E, false);
    flatten_residual(C, P->N, rows, E, F);
    const int M = P->M, MM = M * M, N = P->N, n = M * (N + 1);
    std::memset(J, 0, sizeof(double) * (size_t)n * n);
    std::vector<double> S(MM), R(MM);
    int r = 0;
    auto put_bc = [&](size_t k) { J[(size_t)r * n + rows[k].node * M + rows[k].comp] = 1.0; ++r; };
    if (P->bc_kind == BVP_BC_TWO_POINT) { for (size_t k = 0; k < rows.size(); ++k) if (!rows[k].side_b) put_bc(k); }
    else for (size_t k = 0; k < rows.size(); ++k) put_bc(k);
    for (int i = 0; i < N; ++i) {
        interval_jac(C, mesh[i], mesh[i + 1] - mesh[i], &E.Y[(size_t)i * C.T.s * M], S.data(), R.data());
        for (int a = 0; a < M; ++a) {
            for (int b = 0; b < M; ++b) {
                J[(size_t)r * n + i * M + b] = S[a * M + b];
                J[(size_t)r * n + (i + 1) * M + b] = R[a * M + b];
            }
            ++r;
        }
    }
    if (P->bc_kind == BVP_BC_TWO_POINT) for (size_t k = 0; k < rows.size(); ++k) if (rows[k].side_b) put_bc(k);
    return 0;
}

int ref_defect(int method, const bvp_instance *P, double *errors, double *defect_norm) {
    int rc = check_instance(*P);
    if (rc) return rc;
    Ctx C(*P, method);
    std::vector<double> mesh, y;
    if ((rc = instance_mesh_and_guess(*P, mesh, y))) return rc;
    std::vector<BcRow> rows;
    if ((rc = resolve_bc(*P, P->N, mesh.data(), rows))) return rc;
    Eval E;
    evaluate(C, P->N, mesh.data(), y.data(), rows, E, false);
    std::vector<double> err;
    *defect_norm = defect(C, P->N, mesh.data(), y.data(), E, err);
    std::memcpy(errors, err.data(), sizeof(double) * err.size());
    return 0;
}

// Dense output at (mesh, guess): sol(t) and sol(t, Val{1}) of MIRKInterpolation
// (MIRK/interpolation.jl:77-196): interval(mesh, t), tau = (t - t_i)/h,
// u = y_i + h (K w + KI w*), u' = K w' + KI w'*, stages at (y_i, y_{i+1}).
int ref_interpolate(int method, const bvp_instance *P, int nt, const double *t, double *u, double *du) {
    int rc = check_instance(*P);
    if (rc) return rc;
    Ctx C(*P, method);
    std::vector<double> mesh, y;
    if ((rc = instance_mesh_and_guess(*P, mesh, y))) return rc;
    const int M = P->M, s = C.T.s, ns = C.T.sstar - C.T.s, N = P->N;
    std::vector<double> K(s * M), Y(s * M), KI(std::max(1, ns) * M), w(9), wp(9);
    for (int q = 0; q < nt; ++q) {
        const int i = interval_of(mesh.data(), N, t[q]);
        const double h = mesh[i + 1] - mesh[i], tau = (t[q] - mesh[i]) / h;
        interval_stages(C, mesh[i], h, &y[(size_t)i * M], &y[(size_t)(i + 1) * M], K.data(), Y.data());
        interval_interp_stages(C, mesh[i], h, &y[(size_t)i * M], &y[(size_t)(i + 1) * M], K.data(), KI.data());
        interp_weights(method, tau, w.data(), wp.data());
        sum_stages(C, h, &y[(size_t)i * M], K.data(), KI.data(), w.data(), wp.data(), u + (size_t)q * M,
                   du ? du + (size_t)q * M : nullptr);
    }
    return 0;
}

int ref_mesh_select(const bvp_opts *o, const bvp_instance *P, int capacity, int32_t *new_N, int32_t *info,
                    double *new_mesh, double *new_y) {
    int rc = check_instance(*P);
    if (rc) return rc;
    Ctx C(*P, o->method);
    std::vector<double> mesh, y;
    if ((rc = instance_mesh_and_guess(*P, mesh, y))) return rc;
    std::vector<BcRow> rows;
    if ((rc = resolve_bc(*P, P->N, mesh.data(), rows))) return rc;
    Eval E;
    evaluate(C, P->N, mesh.data(), y.data(), rows, E, false);
    std::vector<double> err, nm, ny;
    defect(C, P->N, mesh.data(), y.data(), E, err);
    *info = mesh_select(C.T.order, o->abstol, o->max_num_subintervals, mesh, err, P->M, nm);
    if (*info != BVP_RET_SUCCESS) { *new_N = P->N; return 0; }
    transfer(C, o->transfer_mode, mesh, y, E, nm, ny);
    *new_N = (int)nm.size() - 1;
    if ((int)nm.size() > capacity) return BVP_EINVAL;
    std::memcpy(new_mesh, nm.data(), sizeof(double) * nm.size());
    std::memcpy(new_y, ny.data(), sizeof(double) * ny.size());
    return 0;
}

// mesh_selector! alone on given per-interval errors (for hand-computed cases)
int ref_mesh_select_errors(int order, double abstol, int max_sub, int N, const double *mesh, int M,
                           const double *errors, int capacity, double *new_mesh) {
    std::vector<double> m(mesh, mesh + N + 1), e(errors, errors + (size_t)N * M), nm;
    int info = mesh_select(order, abstol, max_sub, m, e, M, nm);
    if (info != BVP_RET_SUCCESS) return -(int)info;  // negative: failure
    if ((int)nm.size() > capacity) return BVP_EINVAL - 100;
    std::memcpy(new_mesh, nm.data(), sizeof(double) * nm.size());
    return (int)nm.size() - 1;
}

// ---- FIRK (expanded form).  `stages` is [N*s*M] (interval, stage, component) or
// NULL for extend_y's stages.
// tableau export: ints = [s, order, adaptive], dbl = A[49] (row-major, Butcher
// order), b[7], c[7], Q[49], tau
int ref_firk_tableau(int code, int *ints, double *dbl) {
    FTab T;
    if (!firk_tableau(code, T)) return BVP_EINVAL;
    ints[0] = T.s; ints[1] = T.order; ints[2] = T.adaptive ? 1 : 0;
    int o = 0;
    for (int i = 0; i < 7; ++i) for (int j = 0; j < 7; ++j) dbl[o++] = T.A[i][j];
    for (int i = 0; i < 7; ++i) dbl[o++] = T.b[i];
    for (int i = 0; i < 7; ++i) dbl[o++] = T.c[i];
    for (int i = 0; i < 7; ++i) for (int j = 0; j < 7; ++j) dbl[o++] = T.Q[i][j];
    dbl[o++] = T.tau;
    return 0;
}

int ref_firk_residual(int code, const bvp_instance *P, const double *stages, double *resid) {
    FTab T;
    int rc = firk_check(*P, code, T);
    if (rc) return rc;
    FCtx C(*P, T);
    if (firk_is_nested(code)) {
        std::vector<double> mesh, y, k0;
        std::vector<BcRow> rows;
        FNEval E;
        if ((rc = firk_nested_state(C, mesh, y, k0, rows, E))) return rc;
        flatten_residual_phi(*P, rows, E.phi, E.bc, resid);
        return 0;
    }
    std::vector<double> mesh, x;
    std::vector<BcRow> rows;
    if ((rc = firk_state(C, stages, mesh, x, rows))) return rc;
    FEval E;
    firk_evaluate(C, P->N, mesh.data(), x.data(), rows, E);
    firk_flatten(C, rows, E, resid);
    return 0;
}

// Dense expanded Jacobian (reference row layout, n_x x n_x row-major) and residual.
int ref_firk_dense_system(int code, const bvp_instance *P, const double *stages, double *J, double *F) {
    FTab T;
    int rc = firk_check(*P, code, T);
    if (rc) return rc;
    FCtx C(*P, T);
    std::vector<Trip> Jt;
    std::vector<double> rhs;
    std::vector<BcRow> rows;
    int n, per_iv;
    if (firk_is_nested(code)) {
        // node layout: J is M (N+1) square, rows in the reference's order
        std::vector<double> mesh, y, k0;
        FNEval E;
        if ((rc = firk_nested_state(C, mesh, y, k0, rows, E))) return rc;
        flatten_residual_phi(*P, rows, E.phi, E.bc, F);
        firk_nested_jacobian(C, P->N, mesh.data(), y.data(), rows, E, Jt, rhs);
        n = C.M * (P->N + 1);
        per_iv = C.M;
    } else {
        std::vector<double> mesh, x;
        if ((rc = firk_state(C, stages, mesh, x, rows))) return rc;
        FEval E;
        firk_evaluate(C, P->N, mesh.data(), x.data(), rows, E);
        firk_flatten(C, rows, E, F);
        firk_jacobian(C, P->N, mesh.data(), rows, E, Jt, rhs);
        n = (int)C.nx(P->N);
        per_iv = (T.s + 1) * C.M;
    }
    const std::vector<int> map = firk_row_map(rows, P->N, per_iv, P->bc_kind == BVP_BC_TWO_POINT);
    if ((int)map.size() != n) return 1;
    std::memset(J, 0, sizeof(double) * (size_t)n * n);
    for (const Trip &t : Jt) J[(size_t)map[t.r] * n + t.c] += t.v;
    return 0;
}

// EvalSol of interval [tj, tj1] at t (firk_evalsol), for the checks of its derivatives:
// value and g ((s+2) M: y_j, y_{j+1}, K_j)
int ref_firk_evalsol(int code, const bvp_instance *P, double tj, double tj1, double t, int comp, const double *yi,
                     const double *yj, const double *K, double *value, double *g) {
    FTab T;
    if (!firk_tableau(code, T) || comp < 0 || comp >= P->M) return BVP_EINVAL;
    FCtx C(*P, T);
    const double mesh[2] = {tj, tj1};
    *value = firk_evalsol(C, mesh, 0, t, comp, yi, yj, K, !firk_is_nested(code), g);
    return 0;
}

int ref_firk_newton_step(int code, const bvp_instance *P, const double *stages, double *delta) {
    FTab T;
    int rc = firk_check(*P, code, T);
    if (rc) return rc;
    FCtx C(*P, T);
    if (firk_is_nested(code)) {
        std::vector<double> mesh, y, k0;
        std::vector<BcRow> rows;
        FNEval E;
        if ((rc = firk_nested_state(C, mesh, y, k0, rows, E))) return rc;
        return firk_nested_dir(C, P->N, mesh.data(), y.data(), rows, E, delta) ? 1 : 0;
    }
    std::vector<double> mesh, x;
    std::vector<BcRow> rows;
    if ((rc = firk_state(C, stages, mesh, x, rows))) return rc;
    FEval E;
    firk_evaluate(C, P->N, mesh.data(), x.data(), rows, E);
    return firk_newton_dir(C, P->N, mesh.data(), rows, E, delta) ? 1 : 0;
}

int ref_firk_defect(int code, const bvp_instance *P, const double *stages, double *errors, double *defect_norm) {
    FTab T;
    int rc = firk_check(*P, code, T);
    if (rc) return rc;
    FCtx C(*P, T);
    std::vector<double> mesh, x, err;
    std::vector<BcRow> rows;
    if (firk_is_nested(code)) {
        std::vector<double> y, k0;
        FNEval E;
        if ((rc = firk_nested_state(C, mesh, y, k0, rows, E))) return rc;
        firk_extend(C, P->N, y.data(), E.K.data(), x);
    } else if ((rc = firk_state(C, stages, mesh, x, rows))) {
        return rc;
    }
    *defect_norm = firk_defect(C, P->N, mesh.data(), x.data(), err);
    std::memcpy(errors, err.data(), sizeof(double) * err.size());
    return 0;
}

int ref_firk_interpolate(int code, const bvp_instance *P, const double *stages, int nt, const double *t, double *u,
                         double *du) {
    FTab T;
    int rc = firk_check(*P, code, T);
    if (rc) return rc;
    FCtx C(*P, T);
    std::vector<double> mesh, x;
    std::vector<BcRow> rows;
    if (firk_is_nested(code)) {
        // interpolant! of FIRKNestedInterpolation (FIRK/interpolation.jl:127-181): the stages
        // of the point's interval from the inner solve
        std::vector<double> y, k0;
        FNEval E;
        if ((rc = firk_nested_state(C, mesh, y, k0, rows, E))) return rc;
        firk_extend(C, P->N, y.data(), E.K.data(), x);
    } else if ((rc = firk_state(C, stages, mesh, x, rows))) {
        return rc;
    }
    for (int q = 0; q < nt; ++q)
        firk_interp(C, P->N, mesh.data(), x.data(), t[q], u + (size_t)q * C.M, du ? du + (size_t)q * C.M : nullptr);
    return 0;
}

int ref_solve(const bvp_opts *o, const bvp_instance *P, bvp_result *res) { return solve_any(*P, *o, *res); }

// EnsembleThreads analogue: one instance per thread, dynamic schedule.
int ref_solve_batch(const bvp_opts *o, int n, const bvp_instance *inst, bvp_result *res, int nthreads) {
    int bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
    for (int i = 0; i < n; ++i) {
        int rc = solve_any(inst[i], *o, res[i]);
        if (rc) bad |= 1;
    }
    return bad ? BVP_EINVAL : 0;
}

int ref_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
