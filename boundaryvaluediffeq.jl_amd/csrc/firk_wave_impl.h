// firk_wave_impl.h -- wave-per-instance FIRK collocation kernels (gfx950): Radau IIA
// and Lobatto IIIA/B/C in the expanded form of lib/BoundaryValueDiffEqFIRK/src
// ("FIRK/" below; FIRK/firk.jl:290-453, the reference's default nested_nlsolve = false).
//
// The unknowns of an instance are the node values y_i and the stage values K_i1..K_is
// of every interval.  The reference hands the whole expanded Jacobian (banded,
// FIRK/firk.jl:806-922) to a sparse LU.  Here the stage unknowns of each interval are
// eliminated inside the interval first (static condensation): with the stage
// Jacobians J_r = df/dy(Y_r) and W = I - h (A (x) J) (row block r: delta_rj I - h a_rj J_r),
//   dK = kg + KJ dy_i,   kg = W^-1 G,   KJ = W^-1 [J_1; ...; J_s],
// and the continuity rows Phi_i = y_{i+1} - y_i - h sum_j b_j K_j become the two-node
// relation
//   -(I + h sum_j b_j KJ_j) dy_i + dy_{i+1} = -Phi_i + h sum_j b_j kg_j,
// which is the block structure of the MIRK path: the same chunk elimination across
// lanes, LDS tree and root solve (mirk_wave_impl.h) give dy, then dK per interval
// from the stored kg, KJ.  This is exact linear algebra, so the Newton iterates are
// those of the expanded system.
//
// One wave (64 lanes) per instance, lanes own contiguous chunks of intervals.
// Template parameters: M state dimension (1..4), S stages (1, 2, 3, 4, 5, 7).
#pragma once

#include "mirk_wave_impl.h"

namespace bvpdev {

// stage values of instance b: [cap][S][M] (interval, stage, component)
template <int M, int S>
__device__ inline double *firk_stages_of(const BatchDev &B, int b) {
    return as_global(B.stK + (size_t)b * B.cap * S * M);
}

// ------------------------------------------------------------------ dense solve
// W X = R in place (R becomes X), Gaussian elimination with partial pivoting (first
// maximum).  Small systems stay in registers (unrolled, row swaps by selects); larger
// ones (S M > 8) run as loops over private arrays.  Returns false on a zero pivot.
template <int NN, int NR>
__device__ inline bool lu_solve_small(double (&W)[NN][NN], double (&R)[NN][NR]) {
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        int p = j;
        double best = fabs(W[j][j]);
#pragma unroll
        for (int i = j + 1; i < NN; ++i) {
            const double v = fabs(W[i][j]);
            if (v > best) { best = v; p = i; }
        }
        if (!(best > 0.0)) return false;
#pragma unroll
        for (int i = j + 1; i < NN; ++i) {
            const bool sw = (i == p);
#pragma unroll
            for (int c = j; c < NN; ++c) {
                const double t = W[j][c];
                W[j][c] = sw ? W[i][c] : W[j][c];
                W[i][c] = sw ? t : W[i][c];
            }
#pragma unroll
            for (int c = 0; c < NR; ++c) {
                const double t = R[j][c];
                R[j][c] = sw ? R[i][c] : R[j][c];
                R[i][c] = sw ? t : R[i][c];
            }
        }
        const double inv = 1.0 / W[j][j];
#pragma unroll
        for (int i = j + 1; i < NN; ++i) {
            const double m = W[i][j] * inv;
#pragma unroll
            for (int c = j + 1; c < NN; ++c) W[i][c] -= m * W[j][c];
#pragma unroll
            for (int c = 0; c < NR; ++c) R[i][c] -= m * R[j][c];
        }
    }
#pragma unroll
    for (int i = NN - 1; i >= 0; --i) {
        const double inv = 1.0 / W[i][i];
#pragma unroll
        for (int c = 0; c < NR; ++c) {
            double v = R[i][c];
#pragma unroll
            for (int k = i + 1; k < NN; ++k) v -= W[i][k] * R[k][c];
            R[i][c] = v * inv;
        }
    }
    return true;
}

template <int NN, int NR>
__device__ inline bool lu_solve_loop(double (&W)[NN][NN], double (&R)[NN][NR]) {
#pragma unroll 1
    for (int j = 0; j < NN; ++j) {
        int p = j;
        double best = fabs(W[j][j]);
#pragma unroll 1
        for (int i = j + 1; i < NN; ++i) {
            const double v = fabs(W[i][j]);
            if (v > best) { best = v; p = i; }
        }
        if (!(best > 0.0)) return false;
        if (p != j) {
#pragma unroll 1
            for (int c = j; c < NN; ++c) {
                const double t = W[j][c];
                W[j][c] = W[p][c];
                W[p][c] = t;
            }
#pragma unroll
            for (int c = 0; c < NR; ++c) {
                const double t = R[j][c];
                R[j][c] = R[p][c];
                R[p][c] = t;
            }
        }
        const double inv = 1.0 / W[j][j];
#pragma unroll 1
        for (int i = j + 1; i < NN; ++i) {
            const double m = W[i][j] * inv;
#pragma unroll 4
            for (int c = j + 1; c < NN; ++c) W[i][c] -= m * W[j][c];
#pragma unroll
            for (int c = 0; c < NR; ++c) R[i][c] -= m * R[j][c];
        }
    }
#pragma unroll 1
    for (int i = NN - 1; i >= 0; --i) {
        const double inv = 1.0 / W[i][i];
#pragma unroll
        for (int c = 0; c < NR; ++c) {
            double v = R[i][c];
#pragma unroll 1
            for (int k = i + 1; k < NN; ++k) v -= W[i][k] * R[k][c];
            R[i][c] = v * inv;
        }
    }
    return true;
}

// Local systems up to 12 unknowns run fully unrolled (the arrays at fixed offsets: in
// registers, the overflow spilled at known slots); above, the rolled loops over private
// arrays spill less.  Measured at M = 2, s = 5 (10 unknowns) on 16,384 Bratu BVPs:
// 121 ms rolled, 88.5 ms unrolled.
template <int NN, int NR>
__device__ inline bool lu_solve(double (&W)[NN][NN], double (&R)[NN][NR]) {
    if constexpr (NN <= 12) return lu_solve_small<NN, NR>(W, R);
    else return lu_solve_loop<NN, NR>(W, R);
}

// ------------------------------------------------------------------ one interval
// Stage points Y_r = h (sum_j a_rj K_j) + y_i, stage residuals G_r = f(t_i + c_r h, Y_r)
// - K_r and the continuity residual Phi = (y_{i+1} - y_i) - h (sum_j b_j K_j), in the
// order of operations of Phi! (FIRK/collocation.jl:59-93); with JAC the stage
// Jacobians J_r by forward-mode dual numbers.
template <int M, int S, bool JAC>
__device__ inline void firk_interval(const Inst<M> &I, const FirkTab &T, double ti, double h, const double *yi,
                                     const double *yj, const double (&K)[S][M], double (&G)[S][M], double (&Phi)[M],
                                     double (&J)[S][M][M]) {
#pragma unroll
    for (int r = 0; r < S; ++r) {
        double Y[M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < S; ++j) a += K[j][c] * T.A[r][j];
            Y[c] = h * a + yi[c];
        }
        const double t = ti + T.c[r] * h;
        if constexpr (JAC) {
            Dual<M> Yd[M], Fd[M];
#pragma unroll
            for (int c = 0; c < M; ++c) {
                Yd[c] = dconst<M>(Y[c]);
                Yd[c].d[c] = 1.0;
            }
            f_eval<M, Dual<M>>(I.fam, I.P, t, Yd, Fd);
#pragma unroll
            for (int a = 0; a < M; ++a) {
                G[r][a] = Fd[a].v - K[r][a];
#pragma unroll
                for (int c = 0; c < M; ++c) J[r][a][c] = Fd[a].d[c];
            }
        } else {
            double F[M];
            f_eval<M, double>(I.fam, I.P, t, Y, F);
#pragma unroll
            for (int a = 0; a < M; ++a) G[r][a] = F[a] - K[r][a];
        }
    }
#pragma unroll
    for (int c = 0; c < M; ++c) {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < S; ++j) a += K[j][c] * T.b[j];
        Phi[c] = -h * a + (yj[c] - yi[c]);
    }
}

// local record of an interval: kg [S M] then KJ [S M][M] (row (r, a), column c)
template <int M, int S> constexpr int kFirkLoc = S * M * (M + 1);

// Static condensation of interval k: solves W [kg | KJ] = [G | J], stores the record
// at loc, and forms the two-node relation S_blk dy_i + dy_{i+1} = rhs.
// nested: the stages are solved (G ~ 0 to the inner tolerance) and the relation is
// the Newton step of the nested residual, rhs = -Phi.
template <int M, int S>
__device__ inline bool firk_condense(const FirkTab &T, double h, const double (&G)[S][M], const double (&Phi)[M],
                                     const double (&J)[S][M][M], double *loc, double (&Sb)[M][M], double (&rhs)[M],
                                     bool nested = false) {
    constexpr int NN = S * M;
    double W[NN][NN], X[NN][M + 1];
#pragma unroll
    for (int r = 0; r < S; ++r)
#pragma unroll
        for (int a = 0; a < M; ++a) {
#pragma unroll
            for (int j = 0; j < S; ++j)
#pragma unroll
                for (int c = 0; c < M; ++c)
                    W[r * M + a][j * M + c] = ((r == j && a == c) ? 1.0 : 0.0) - h * T.A[r][j] * J[r][a][c];
            X[r * M + a][0] = G[r][a];
#pragma unroll
            for (int c = 0; c < M; ++c) X[r * M + a][1 + c] = J[r][a][c];
        }
    const bool ok = lu_solve<NN, M + 1>(W, X);
#pragma unroll
    for (int q = 0; q < NN; ++q) {
        loc[q] = X[q][0];
#pragma unroll
        for (int c = 0; c < M; ++c) loc[NN + q * M + c] = X[q][1 + c];
    }
#pragma unroll
    for (int a = 0; a < M; ++a) {
        double g = 0.0;
#pragma unroll
        for (int j = 0; j < S; ++j) g += T.b[j] * X[j * M + a][0];
        rhs[a] = nested ? -Phi[a] : -Phi[a] + h * g;
#pragma unroll
        for (int c = 0; c < M; ++c) {
            double v = 0.0;
#pragma unroll
            for (int j = 0; j < S; ++j) v += T.b[j] * X[j * M + a][1 + c];
            Sb[a][c] = -((a == c) ? 1.0 : 0.0) - h * v;
        }
    }
    return ok;
}

template <int M, int S>
__device__ inline void load_stages(const double *K, int k, double (&Kk)[S][M]) {
#pragma unroll
    for (int r = 0; r < S; ++r)
#pragma unroll
        for (int c = 0; c < M; ++c) Kk[r][c] = K[((size_t)k * S + r) * M + c];
}

// The nested form (nested_nlsolve = true, FIRK/firk.jl:114-288): the unknowns are the
// node values; each evaluation solves an interval's stage equations G(K) = 0 for K
// from the fixed start k0 by NewtonRaphson (NonlinearSolve's, restated: K <- K + W^-1 G
// until max|G| <= tol, checked before every step, NaN stops, at most 1000 steps) and
// the residual is Phi_i.  Its Jacobian enters the outer one by the implicit function
// theorem, dK/dy_i = W^-1 [J_1; ...; J_s] -- the KJ of the condensation.
template <int M> struct FirkNest {
    bool on;
    bool irows;  // boundary rows inside the span may occur (EvalSol rows)
    double tol;
    double k0[M];
};

template <int M>
__device__ inline FirkNest<M> firk_nest_of(const BatchDev &B, int b) {
    FirkNest<M> n;
    n.on = B.firk_nested != 0;
    n.tol = B.nest_tol;
    n.irows = true;
#pragma unroll
    for (int c = 0; c < M; ++c) n.k0[c] = n.on ? B.u0c[(size_t)b * M + c] : 0.0;
    return n;
}

template <int M, int S>
__device__ inline void firk_inner(const Inst<M> &I, const FirkTab &T, const FirkNest<M> &nest, double ti, double h,
                                  const double *yi, double (&K)[S][M]) {
    constexpr int NN = S * M;
#pragma unroll
    for (int r = 0; r < S; ++r)
#pragma unroll
        for (int c = 0; c < M; ++c) K[r][c] = nest.k0[c];
    for (int it = 0; it < 1000; ++it) {
        double G[S][M], Phi[M], J[S][M][M];
        firk_interval<M, S, true>(I, T, ti, h, yi, yi, K, G, Phi, J);
        double mx = 0.0;
        bool bad = false;
#pragma unroll
        for (int r = 0; r < S; ++r)
#pragma unroll
            for (int c = 0; c < M; ++c) {
                bad |= isnan(G[r][c]);
                mx = fmax(mx, fabs(G[r][c]));
            }
        if (bad || !(mx > nest.tol)) break;
        double W[NN][NN], X[NN][1];
#pragma unroll
        for (int r = 0; r < S; ++r)
#pragma unroll
            for (int a = 0; a < M; ++a) {
#pragma unroll
                for (int j = 0; j < S; ++j)
#pragma unroll
                    for (int c = 0; c < M; ++c)
                        W[r * M + a][j * M + c] = ((r == j && a == c) ? 1.0 : 0.0) - h * T.A[r][j] * J[r][a][c];
                X[r * M + a][0] = G[r][a];
            }
        if (!lu_solve<NN, 1>(W, X)) break;
#pragma unroll
        for (int r = 0; r < S; ++r)
#pragma unroll
            for (int c = 0; c < M; ++c) K[r][c] += X[r * M + c][0];
    }
}

template <int M, int S>
__device__ inline void store_stages(double *K, int k, const double (&Kk)[S][M]) {
#pragma unroll
    for (int r = 0; r < S; ++r)
#pragma unroll
        for (int c = 0; c < M; ++c) K[((size_t)k * S + r) * M + c] = Kk[r][c];
}

// The quintic through value and slope at 0, h, h/2 (get_S_coeffs / s_constraints_block,
// FIRK/adaptivity.jl:183-227), solved with partial pivoting as the 6 x 6 system of the
// reference; bk = (y(0), y(h), y(h/2), y'(0), y'(h), y'(h/2)).
__device__ inline void firk_quintic(double h, const double (&bk)[6], double (&coef)[6]) {
    double A[6][6], b[6];
    const double t[3] = {0.0, h, 0.5 * h};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double p = 1.0, pd = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            A[r][m] = p;
            A[3 + r][m] = m * pd;
            pd = p;
            p *= t[r];
        }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) b[r] = bk[r];
    double X[6][1];
#pragma unroll
    for (int r = 0; r < 6; ++r) X[r][0] = b[r];
    lu_solve_small<6, 1>(A, X);
#pragma unroll
    for (int r = 0; r < 6; ++r) coef[r] = X[r][0];
}

// LDS of the FIRK kernels: the wave kernels' area plus, per boundary row, the
// linearisation of an interior row for the root solve ([2M + 2]: coefficients of
// dy_iv and dy_iv+1, constant, flag; the flag is zero for a row read from a node).
template <int M> struct FirkShared : Shared<M> {
    double irow[M * (2 * M + 2)];
};

// ------------------------------------------------------------------ interior boundary rows
// A STANDARD row inside the span reads EvalSol of the FIRK cache (FIRK/interpolation.jl:
// 338-441) on interval j (interval(), Core/utils.jl:119-121: the left interval at an
// interior node): the quintic through y_j, y_{j+1}, their slopes and a value/slope pair
// of the stage polynomial, at tau = t - t_j.  Expanded: the pair at (tau h) in the stage
// polynomial's variable (tau the absolute offset, as the reference writes it), stages of
// the state; nested: the midpoint pair, stages the caller solved from zero (EvalSol's own
// inner problem).  With g: the derivatives of the value, [M] wrt y_j, [M] wrt y_{j+1},
// [S M] wrt K_j (expanded only; the nested inner solve sees y_j without its dual part).
template <int M, int S, bool WANTG>
__device__ inline double firk_evalsol(const Inst<M> &I, const FirkTab &T, bool expanded, double tj, double tj1,
                                      double t, int comp, const double *yi, const double *yj,
                                      const double (&Kk)[S][M], double (&g)[(S + 2) * M]) {
    const double h = tj1 - tj, tau = t - tj;
    const double th = (expanded ? tau : 0.5) * h;
    double fi = 0.0, fj = 0.0, Ji[M], Jj[M], yic = 0.0, yjc = 0.0, kc[S];
    {
        Dual<M> Yd[M], Fd[M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            Yd[c] = dconst<M>(yi[c]);
            Yd[c].d[c] = 1.0;
        }
        f_eval<M, Dual<M>>(I.fam, I.P, tj, Yd, Fd);
#pragma unroll
        for (int a = 0; a < M; ++a)
            if (a == comp) {
                fi = Fd[a].v;
#pragma unroll
                for (int c = 0; c < M; ++c) Ji[c] = Fd[a].d[c];
            }
#pragma unroll
        for (int c = 0; c < M; ++c) {
            Yd[c] = dconst<M>(yj[c]);
            Yd[c].d[c] = 1.0;
        }
        f_eval<M, Dual<M>>(I.fam, I.P, tj1, Yd, Fd);
#pragma unroll
        for (int a = 0; a < M; ++a)
            if (a == comp) {
                fj = Fd[a].v;
#pragma unroll
                for (int c = 0; c < M; ++c) Jj[c] = Fd[a].d[c];
            }
    }
#pragma unroll
    for (int c = 0; c < M; ++c)
        if (c == comp) {
            yic = yi[c];
            yjc = yj[c];
#pragma unroll
            for (int q = 0; q < S; ++q) kc[q] = Kk[q][c];
        }
    // stage polynomial of component comp and its derivatives wrt K_j.comp
    double v = 0.0, zp = 0.0, p = 1.0, hk = 1.0, dz[S], dzp[S];
#pragma unroll
    for (int q = 0; q < S; ++q) { dz[q] = 0.0; dzp[q] = 0.0; }
#pragma unroll
    for (int k = 0; k < S; ++k) {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < S; ++q) a += T.Q[k][q] * kc[q];
        const double coef = a / hk;
        zp += (k + 1) * coef * p;
#pragma unroll
        for (int q = 0; q < S; ++q) dzp[q] += (k + 1) * (T.Q[k][q] / hk) * p;
        p *= th;
        v += coef * p;
#pragma unroll
        for (int q = 0; q < S; ++q) dz[q] += (T.Q[k][q] / hk) * p;
        hk *= h;
    }
    const double bk[6] = {yic, yjc, yic + v, fi, fj, zp};
    double cf[6];
    firk_quintic(h, bk, cf);
    double val = 0.0, tp = 1.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) { val += cf[m] * tp; tp *= tau; }
    if constexpr (WANTG) {
        // the quintic is linear in its six conditions: their weights w in the value solve
        // A' w = (1, tau, ..., tau^5), A the constraint matrix of firk_quintic
        double At[6][6], X[6][1], w[6];
        const double tn[3] = {0.0, h, 0.5 * h};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double pw = 1.0, pd = 0.0;
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                At[m][r] = pw;
                At[m][3 + r] = m * pd;
                pd = pw;
                pw *= tn[r];
            }
        }
        double pp = 1.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) { X[m][0] = pp; pp *= tau; }
        lu_solve_small<6, 1>(At, X);
#pragma unroll
        for (int r = 0; r < 6; ++r) w[r] = X[r][0];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            g[c] = w[3] * Ji[c] + (c == comp ? w[0] + w[2] : 0.0);
            g[M + c] = w[4] * Jj[c] + (c == comp ? w[1] : 0.0);
        }
#pragma unroll
        for (int q = 0; q < S; ++q)
#pragma unroll
            for (int c = 0; c < M; ++c)
                g[2 * M + q * M + c] = (expanded && c == comp) ? w[2] * dz[q] + w[5] * dzp[q] : 0.0;
    }
    return val;
}

// Residual of the boundary row k inside the span (interval j) at (y + alpha dy,
// K + alpha dK) (dy, dK may be null: alpha 0; K unused by the nested form).  WANTG: also
// its linearisation for the root solve into ir: [M] coefficients of dy_j, [M] of dy_{j+1},
// the constant, with the stage corrections eliminated through the interval's record lk
// (dK = kg + KJ dy_j).
template <int M, int S, bool WANTG>
__device__ inline double firk_bc_interior(const Inst<M> &I, const FirkTab &T, const FirkNest<M> &nest,
                                          const double *mesh, const double *y, const double *K, const double *dy,
                                          const double *dK, double alpha, int k, int j, const double *lk,
                                          double *ir, const double *dy2 = nullptr, const double *dK2 = nullptr,
                                          double beta = 0.0, double *graw = nullptr) {
    const int comp = sel<M>(I.bcc, k);
    double yi[M], yj[M], Kk[S][M];
#pragma unroll
    for (int c = 0; c < M; ++c) {
        yi[c] = y[(size_t)j * M + c];
        yj[c] = y[(size_t)(j + 1) * M + c];
        if (dy) {
            yi[c] += alpha * dy[(size_t)j * M + c];
            yj[c] += alpha * dy[(size_t)(j + 1) * M + c];
        }
        if (dy2) {
            yi[c] += beta * dy2[(size_t)j * M + c];
            yj[c] += beta * dy2[(size_t)(j + 1) * M + c];
        }
    }
    if (nest.on) {
        FirkNest<M> z = nest;
#pragma unroll
        for (int c = 0; c < M; ++c) z.k0[c] = 0.0;
        firk_inner<M, S>(I, T, z, mesh[j], mesh[j + 1] - mesh[j], yi, Kk);
    } else {
#pragma unroll
        for (int r = 0; r < S; ++r)
#pragma unroll
            for (int c = 0; c < M; ++c) {
                const size_t q = ((size_t)j * S + r) * M + c;
                Kk[r][c] = (dK ? K[q] + alpha * dK[q] : K[q]) + (dK2 ? beta * dK2[q] : 0.0);
            }
    }
    double g[(S + 2) * M];
    const double v =
        firk_evalsol<M, S, WANTG>(I, T, !nest.on, mesh[j], mesh[j + 1], sel<M>(I.bct, k), comp, yi, yj, Kk, g);
    if constexpr (WANTG) {
        if (graw)
#pragma unroll
            for (int q = 0; q < (S + 2) * M; ++q) graw[q] = g[q];
    }
    if (WANTG && ir) {
        constexpr int NN = S * M;
        double cst = 0.0, ca[M];
#pragma unroll
        for (int c = 0; c < M; ++c) ca[c] = g[c];
        if (!nest.on) {
#pragma unroll
            for (int q = 0; q < NN; ++q) {
                const double gq = g[2 * M + q];
                cst += gq * lk[q];
#pragma unroll
                for (int c = 0; c < M; ++c) ca[c] += gq * lk[NN + q * M + c];
            }
        }
#pragma unroll
        for (int c = 0; c < M; ++c) { ir[c] = ca[c]; ir[M + c] = g[M + c]; }
        ir[2 * M] = cst;
    }
    return v - sel<M>(I.bcv, k);
}

// Residual of boundary row k (L: where it sits) at (y + alpha dy, K + alpha dK).
template <int M, int S>
__device__ inline double firk_bc_row(const Inst<M> &I, const FirkTab &T, const FirkNest<M> &nest, const double *mesh,
                                     const double *y, const double *K, const double *dy, const double *dK,
                                     double alpha, int k, const BcLoc &L, const double *dy2 = nullptr,
                                     const double *dK2 = nullptr, double beta = 0.0) {
    if (L.iv < 0 || !nest.irows) {
        const size_t q = (size_t)L.node * M + sel<M>(I.bcc, k);
        return (dy ? y[q] + alpha * dy[q] : y[q]) + (dy2 ? beta * dy2[q] : 0.0) - sel<M>(I.bcv, k);
    }
    return firk_bc_interior<M, S, false>(I, T, nest, mesh, y, K, dy, dK, alpha, k, L.iv, nullptr, nullptr, dy2, dK2,
                                         beta);
}

// Boundary rows of the FIRK path at (y, K): residuals and nodes in sh.
template <int M, int S>
__device__ inline void firk_bc_rows(const Inst<M> &I, const FirkTab &T, const FirkNest<M> &nest, Shared<M> &sh,
                                    const double *mesh, const double *y, const double *K, int N, int lane) {
    if (lane < M) {
        const BcLoc L = bc_locate<M>(I, lane, mesh, N);
        sh.bcnode[lane] = L.node;
        sh.bcres[lane] = firk_bc_row<M, S>(I, T, nest, mesh, y, K, nullptr, nullptr, 0.0, lane, L);
    }
    __syncthreads();
}

// ------------------------------------------------------------------ Newton pieces
// Forward pass at (y, K): residual max (NaN-aware), per-interval records (loc) and the
// chunk elimination (Back records in scr, the chunk relation in rel).  Returns
// max|F| over Phi, G and the boundary rows in all lanes (NaN if not finite); *sing is
// set when a local system was singular.  Nested form: K is written, not read (each
// interval's stages solved from k0), and max|F| is over Phi and the boundary rows.
template <int M, int S>
__device__ inline double firk_forward(const Inst<M> &I, const FirkTab &T, const FirkNest<M> &nest, Shared<M> &sh,
                                      const double *mesh, const double *y, double *K, int N, int lane, double *scr,
                                      double *loc, Rel<M> &rel, bool &sing) {
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    double mx = 0.0;
    bool bad = false, sg = false;
    for (int k = start; k < start + len; ++k) {
        const double ti = mesh[k], h = mesh[k + 1] - ti;
        double yi[M], yj[M], Kk[S][M], G[S][M], Phi[M], J[S][M][M];
#pragma unroll
        for (int c = 0; c < M; ++c) { yi[c] = y[(size_t)k * M + c]; yj[c] = y[(size_t)(k + 1) * M + c]; }
        if (nest.on) {
            firk_inner<M, S>(I, T, nest, ti, h, yi, Kk);
            store_stages<M, S>(K, k, Kk);
        } else {
            load_stages<M, S>(K, k, Kk);
        }
        firk_interval<M, S, true>(I, T, ti, h, yi, yj, Kk, G, Phi, J);
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const double a = fabs(Phi[c]);
            bad |= !(a <= 1.7976931348623157e308);
            mx = fmax(mx, a);
        }
        if (!nest.on) {
#pragma unroll
            for (int r = 0; r < S; ++r)
#pragma unroll
                for (int c = 0; c < M; ++c) {
                    const double a = fabs(G[r][c]);
                    bad |= !(a <= 1.7976931348623157e308);
                    mx = fmax(mx, a);
                }
        }
        double Sb[M][M], rhs[M], Rb[M][M];
        sg |= !firk_condense<M, S>(T, h, G, Phi, J, loc + (size_t)k * kFirkLoc<M, S>, Sb, rhs, nest.on);
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
            for (int c = 0; c < M; ++c) Rb[a][c] = (a == c) ? 1.0 : 0.0;
        if (k == start) {
#pragma unroll
            for (int a = 0; a < M; ++a) {
#pragma unroll
                for (int c = 0; c < M; ++c) { rel.E[a][c] = Sb[a][c]; rel.F[a][c] = Rb[a][c]; }
                rel.g[a] = rhs[a];
            }
        } else {
            Rel<M> nr;
            Back<M> bk;
            eliminate<M>(rel, Sb, Rb, rhs, nr, bk);
            put_rec<M, false>(scr, k - start - 1, lane, bk);
            rel = nr;
        }
    }
    firk_bc_rows<M, S>(I, T, nest, sh, mesh, y, K, N, lane);
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double a = fabs(sh.bcres[k]);
        bad |= !(a <= 1.7976931348623157e308);
        mx = fmax(mx, a);
    }
    const bool anybad = wave_any(bad);
    sing = wave_any(sg);
    mx = wave_max(mx);
    __syncthreads();
    return anybad ? NAN : mx;
}

// Tree solve and back-substitution of the condensed system: dy at the nodes, then
// dK = kg + KJ dy_i per interval.  dy: [N+1][M], dK: [N][S][M].  Returns false when
// the root system is singular.
template <int M, int S, bool IR = true>
__device__ inline bool firk_correction(const Inst<M> &I, const FirkTab &T, const FirkNest<M> &nest, Shared<M> &sh,
                                       const double *mesh, const double *y, const double *K, int N, int lane,
                                       double *scr, const double *loc, Rel<M> &rel, double *dy, double *dK) {
    const bool nested = nest.on;
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    // the linearisation of rows inside the span (rows read from a node carry a zero flag)
    double *irow = static_cast<FirkShared<M> &>(sh).irow;
    if constexpr (IR) {
        __syncthreads();  // the interval records are read across lanes
        if (lane < M) {
            const BcLoc L = bc_locate<M>(I, lane, mesh, N);
            double *ir = irow + lane * (2 * M + 2);
            ir[2 * M + 1] = L.iv >= 0 ? 1.0 : 0.0;
            if (L.iv >= 0)
                firk_bc_interior<M, S, true>(I, T, nest, mesh, y, K, nullptr, nullptr, 0.0, lane, L.iv,
                                             loc + (size_t)L.iv * kFirkLoc<M, S>, ir);
        }
        if (!tree_solve<M, false, true>(I, sh, N, lane, P, scr, rel, irow)) return false;
    } else if (!tree_solve<M, false>(I, sh, N, lane, P, scr, rel)) {
        return false;
    }
    if (len > 0) {
        double ya[M], yn[M];
#pragma unroll
        for (int c = 0; c < M; ++c) { ya[c] = sh.lo[lane][c]; yn[c] = sh.hi[lane][c]; }
        if (lane == P - 1)
#pragma unroll
            for (int c = 0; c < M; ++c) dy[(size_t)N * M + c] = yn[c];
        for (int k = start + len - 1; k > start; --k) {
            Back<M> bk;
            get_rec<M, false>(scr, k - start - 1, lane, bk);
            double yk[M];
            back_apply<M>(bk, ya, yn, yk);
#pragma unroll
            for (int c = 0; c < M; ++c) { dy[(size_t)k * M + c] = yk[c]; yn[c] = yk[c]; }
        }
#pragma unroll
        for (int c = 0; c < M; ++c) dy[(size_t)start * M + c] = ya[c];
        // stage corrections of the chunk's intervals (dy of their left nodes: just written by
        // this lane); the nested form has no stage unknowns
        constexpr int NN = S * M;
        for (int k = start; k < start + (nested ? 0 : len); ++k) {
            const double *lk = loc + (size_t)k * kFirkLoc<M, S>;
            double d[M];
#pragma unroll
            for (int c = 0; c < M; ++c) d[c] = dy[(size_t)k * M + c];
#pragma unroll
            for (int q = 0; q < NN; ++q) {
                double v = lk[q];
#pragma unroll
                for (int c = 0; c < M; ++c) v += lk[NN + q * M + c] * d[c];
                dK[(size_t)k * NN + q] = v;
            }
        }
    }
    __syncthreads();
    return true;
}

// Sum of squares of the residual at (y + alpha dy, K + alpha dK) (the line search's
// merit, times 2) in all lanes; NaN if not finite.
template <int M, int S>
__device__ inline double firk_trial_sumsq(const Inst<M> &I, const FirkTab &T, const FirkNest<M> &nest,
                                          const double *mesh, const double *y, const double *K, const double *dy,
                                          const double *dK, double alpha, int N, int lane) {
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    double ss = 0.0;
    bool bad = false;
    for (int k = start; k < start + len; ++k) {
        const double ti = mesh[k], h = mesh[k + 1] - ti;
        double yi[M], yj[M], Kk[S][M], G[S][M], Phi[M], J[S][M][M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            yi[c] = y[(size_t)k * M + c] + alpha * dy[(size_t)k * M + c];
            yj[c] = y[(size_t)(k + 1) * M + c] + alpha * dy[(size_t)(k + 1) * M + c];
        }
        if (nest.on) {
            firk_inner<M, S>(I, T, nest, ti, h, yi, Kk);
        } else {
#pragma unroll
            for (int r = 0; r < S; ++r)
#pragma unroll
                for (int c = 0; c < M; ++c) {
                    const size_t q = ((size_t)k * S + r) * M + c;
                    Kk[r][c] = K[q] + alpha * dK[q];
                }
        }
        firk_interval<M, S, false>(I, T, ti, h, yi, yj, Kk, G, Phi, J);
#pragma unroll
        for (int c = 0; c < M; ++c) ss += Phi[c] * Phi[c];
        if (!nest.on) {
#pragma unroll
            for (int r = 0; r < S; ++r)
#pragma unroll
                for (int c = 0; c < M; ++c) ss += G[r][c] * G[r][c];
        }
    }
    if (lane < M) {
        const BcLoc L = bc_locate<M>(I, lane, mesh, N);
        const double v = firk_bc_row<M, S>(I, T, nest, mesh, y, K, dy, dK, alpha, lane, L);
        ss += v * v;
    }
    bad = !(fabs(ss) <= 1.7976931348623157e308);
    const bool anybad = wave_any(bad);
    ss = wave_sum(ss);
    return anybad ? NAN : uniform(ss);
}

// x += alpha d over the nodes and stages of the lane's chunk (node N by the last lane)
template <int M, int S>
__device__ inline void firk_apply(double *y, double *K, const double *dy, const double *dK, double alpha, int N,
                                  int lane, bool nested) {
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    for (int k = start; k < start + len; ++k) {
#pragma unroll
        for (int c = 0; c < M; ++c) y[(size_t)k * M + c] += alpha * dy[(size_t)k * M + c];
        if (!nested) {
#pragma unroll
            for (int q = 0; q < S * M; ++q) K[(size_t)k * S * M + q] += alpha * dK[(size_t)k * S * M + q];
        }
    }
    if (lane == P - 1)
#pragma unroll
        for (int c = 0; c < M; ++c) y[(size_t)N * M + c] += alpha * dy[(size_t)N * M + c];
    __syncthreads();
}

template <int M, int S>
__device__ inline void firk_copy_state(double *yd, double *Kd, const double *ys, const double *Ks, int N, int lane) {
    for (int q = lane; q < (N + 1) * M; q += WAVE) yd[q] = ys[q];
    for (int q = lane; q < N * S * M; q += WAVE) Kd[q] = Ks[q];
    __syncthreads();
}

struct FirkNewtonOut {
    int ret, iters, stage;
    double fn;
};

// ------------------------------------------------------------------ trust region (expanded form)
// the expanded form with boundary rows anywhere in the span
template <int M> __device__ inline FirkNest<M> firk_expanded_rows() {
    FirkNest<M> n{};
    n.on = false;
    n.irows = true;
    return n;
}

// Products with the expanded Jacobian for the TrustRegion stage, rows
//   Phi_i = y_{i+1} - y_i - h sum_j b_j K_ij,  G_ir = f(t_i + c_r h, Y_ir) - K_ir,
//   Y_ir = y_i + h sum_j a_rj K_ij,  boundary rows y_node.comp - value
// (the Jacobian of FIRK/collocation.jl:59-93 that the oracle assembles), stage
// Jacobians J_r evaluated at (y, K).  firk_jt_f: g = J'F into (gy [N+1][M], gK [N][S][M]).
template <int M, int S>
__device__ inline void firk_jt_f(const Inst<M> &I, const FirkTab &T, const double *mesh, const double *y,
                                 const double *K, int N, int lane, double *gy, double *gK) {
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    // pass 1: each interval's terms on y_i and K_i (node N: zero)
    for (int k = start; k < start + len; ++k) {
        const double ti = mesh[k], h = mesh[k + 1] - ti;
        double yi[M], yj[M], Kk[S][M], G[S][M], Phi[M], J[S][M][M];
#pragma unroll
        for (int c = 0; c < M; ++c) { yi[c] = y[(size_t)k * M + c]; yj[c] = y[(size_t)(k + 1) * M + c]; }
        load_stages<M, S>(K, k, Kk);
        firk_interval<M, S, true>(I, T, ti, h, yi, yj, Kk, G, Phi, J);
#pragma unroll
        for (int c = 0; c < M; ++c) {
            double v = -Phi[c];
#pragma unroll
            for (int r = 0; r < S; ++r)
#pragma unroll
                for (int a = 0; a < M; ++a) v += J[r][a][c] * G[r][a];
            gy[(size_t)k * M + c] = v;
        }
#pragma unroll
        for (int j = 0; j < S; ++j)
#pragma unroll
            for (int c = 0; c < M; ++c) {
                double v = -h * T.b[j] * Phi[c] - G[j][c];
#pragma unroll
                for (int r = 0; r < S; ++r)
#pragma unroll
                    for (int a = 0; a < M; ++a) v += h * T.A[r][j] * J[r][a][c] * G[r][a];
                gK[((size_t)k * S + j) * M + c] = v;
            }
    }
    if (lane == P - 1)
#pragma unroll
        for (int c = 0; c < M; ++c) gy[(size_t)N * M + c] = 0.0;
    __syncthreads();
    // pass 2: Phi_i on y_{i+1}
    for (int k = start; k < start + len; ++k) {
        const double h = mesh[k + 1] - mesh[k];
        double Kk[S][M];
        load_stages<M, S>(K, k, Kk);
#pragma unroll
        for (int c = 0; c < M; ++c) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < S; ++j) a += Kk[j][c] * T.b[j];
            gy[(size_t)(k + 1) * M + c] += -h * a + (y[(size_t)(k + 1) * M + c] - y[(size_t)k * M + c]);
        }
    }
    __syncthreads();
    if (lane == 0)
        for (int r = 0; r < M; ++r) {
            const BcLoc L = bc_locate<M>(I, r, mesh, N);
            if (L.iv < 0) {
                const size_t q = (size_t)L.node * M + sel<M>(I.bcc, r);
                gy[q] += y[q] - sel<M>(I.bcv, r);
            } else {
                // a row inside the span: its EvalSol derivatives on y_j, y_{j+1}, K_j
                const int j = L.iv;
                double g[(S + 2) * M];
                const double v = firk_bc_interior<M, S, true>(I, T, firk_expanded_rows<M>(), mesh, y, K, nullptr,
                                                              nullptr, 0.0, r, j, nullptr, nullptr, nullptr, nullptr,
                                                              0.0, g);
                for (int c = 0; c < M; ++c) {
                    gy[(size_t)j * M + c] += v * g[c];
                    gy[(size_t)(j + 1) * M + c] += v * g[M + c];
                }
                for (int q = 0; q < S * M; ++q) gK[(size_t)j * S * M + q] += v * g[2 * M + q];
            }
        }
    __syncthreads();
}

// |J g|^2 for g = (gy, gK), in all lanes
template <int M, int S>
__device__ inline double firk_jac_norm2(const Inst<M> &I, const FirkTab &T, const double *mesh, const double *y,
                                        const double *K, const double *gy, const double *gK, int N, int lane) {
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    double acc = 0.0;
    for (int k = start; k < start + len; ++k) {
        const double ti = mesh[k], h = mesh[k + 1] - ti;
        double yi[M], yj[M], Kk[S][M], G[S][M], Phi[M], J[S][M][M], gk[S][M];
#pragma unroll
        for (int c = 0; c < M; ++c) { yi[c] = y[(size_t)k * M + c]; yj[c] = y[(size_t)(k + 1) * M + c]; }
        load_stages<M, S>(K, k, Kk);
        load_stages<M, S>(gK, k, gk);
        firk_interval<M, S, true>(I, T, ti, h, yi, yj, Kk, G, Phi, J);
#pragma unroll
        for (int a = 0; a < M; ++a) {
            double v = gy[(size_t)(k + 1) * M + a] - gy[(size_t)k * M + a];
#pragma unroll
            for (int j = 0; j < S; ++j) v -= h * T.b[j] * gk[j][a];
            acc += v * v;
        }
#pragma unroll
        for (int r = 0; r < S; ++r)
#pragma unroll
            for (int a = 0; a < M; ++a) {
                double v = -gk[r][a];
#pragma unroll
                for (int c = 0; c < M; ++c) {
                    double w = gy[(size_t)k * M + c];
#pragma unroll
                    for (int j = 0; j < S; ++j) w += h * T.A[r][j] * gk[j][c];
                    v += J[r][a][c] * w;
                }
                acc += v * v;
            }
    }
    if (lane == 0)
        for (int r = 0; r < M; ++r) {
            const BcLoc L = bc_locate<M>(I, r, mesh, N);
            double v;
            if (L.iv < 0) {
                v = gy[(size_t)L.node * M + sel<M>(I.bcc, r)];
            } else {
                const int j = L.iv;
                double g[(S + 2) * M];
                firk_bc_interior<M, S, true>(I, T, firk_expanded_rows<M>(), mesh, y, K, nullptr, nullptr, 0.0, r, j,
                                             nullptr, nullptr, nullptr, nullptr, 0.0, g);
                v = 0.0;
                for (int c = 0; c < M; ++c) v += g[c] * gy[(size_t)j * M + c] + g[M + c] * gy[(size_t)(j + 1) * M + c];
                for (int q = 0; q < S * M; ++q) v += g[2 * M + q] * gK[(size_t)j * S * M + q];
            }
            acc += v * v;
        }
    const double s = uniform(wave_sum(acc));
    __syncthreads();
    return s;
}

// |F|^2 and max|F| at (y + a dy + b gy, K + a dK + b gK), in all lanes (NaN if not finite)
template <int M, int S>
__device__ inline double firk_trial2(const Inst<M> &I, const FirkTab &T, const double *mesh, const double *y,
                                     const double *K, const double *dy, const double *dK, double a, const double *gy,
                                     const double *gK, double b, int N, int lane, double &mx) {
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    double ss = 0.0, m = 0.0;
    for (int k = start; k < start + len; ++k) {
        const double ti = mesh[k], h = mesh[k + 1] - ti;
        double yi[M], yj[M], Kk[S][M], G[S][M], Phi[M], J[S][M][M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const size_t p = (size_t)k * M + c, q = (size_t)(k + 1) * M + c;
            yi[c] = y[p] + a * dy[p] + b * gy[p];
            yj[c] = y[q] + a * dy[q] + b * gy[q];
        }
#pragma unroll
        for (int r = 0; r < S; ++r)
#pragma unroll
            for (int c = 0; c < M; ++c) {
                const size_t q = ((size_t)k * S + r) * M + c;
                Kk[r][c] = K[q] + a * dK[q] + b * gK[q];
            }
        firk_interval<M, S, false>(I, T, ti, h, yi, yj, Kk, G, Phi, J);
#pragma unroll
        for (int c = 0; c < M; ++c) { ss += Phi[c] * Phi[c]; m = fmax(m, fabs(Phi[c])); }
#pragma unroll
        for (int r = 0; r < S; ++r)
#pragma unroll
            for (int c = 0; c < M; ++c) { ss += G[r][c] * G[r][c]; m = fmax(m, fabs(G[r][c])); }
    }
    if (lane < M) {
        const BcLoc L = bc_locate<M>(I, lane, mesh, N);
        const double v =
            firk_bc_row<M, S>(I, T, firk_expanded_rows<M>(), mesh, y, K, dy, dK, a, lane, L, gy, gK, b);
        ss += v * v;
        m = fmax(m, fabs(v));
    }
    const bool bad = !(fabs(ss) <= 1.7976931348623157e308);
    const bool anybad = wave_any(bad);
    ss = uniform(wave_sum(ss));
    mx = anybad ? NAN : uniform(wave_max(m));
    __syncthreads();
    return anybad ? NAN : ss;
}

// The TrustRegion stage on the expanded system (dogleg between the Newton step and the
// Cauchy point along J'F, initial radius max(|F|, max x - min x) / 11, the "Simple"
// radius update; the restatement the oracle follows, Core/default_internal_solve.jl:
// 37-43).  Boundary rows anywhere in the span.
// gy, gK: work areas for J'F.
template <int M, int S>
__device__ inline int firk_trust_region(const BatchDev &B, const Inst<M> &I, const FirkTab &T,
                                        const FirkNest<M> &nest, Shared<M> &sh, const double *mesh, double *y,
                                        double *K, int N, int lane, double *scr, double *loc, double *dy, double *dK,
                                        double *gy, double *gK, FirkNewtonOut &out) {
    double mx;
    double ss = firk_trial2<M, S>(I, T, mesh, y, K, y, K, 0.0, y, K, 0.0, N, lane, mx);
    double xmax = -INFINITY, xmin = INFINITY;
    for (int q = lane; q < (N + 1) * M; q += WAVE) { xmax = fmax(xmax, y[q]); xmin = fmin(xmin, y[q]); }
    for (int q = lane; q < N * S * M; q += WAVE) { xmax = fmax(xmax, K[q]); xmin = fmin(xmin, K[q]); }
    xmax = uniform(wave_max(xmax));
    xmin = uniform(-wave_max(-xmin));
    __syncthreads();
    double radius = fmax(sqrt(ss), xmax - xmin) / 11.0;
    int shrinks = 0;
    bool fresh = true;
    double dn2 = 0.0, l2 = 0.0, gdn = 0.0, jg2 = 0.0;
    for (int it = 0;; ++it) {
        out.fn = mx;
        if (!(mx <= 1.7976931348623157e308)) return BVP_RET_UNSTABLE;
        if (mx <= B.newton_tol) return BVP_RET_SUCCESS;
        if (it >= B.max_newton) return BVP_RET_MAXITERS;
        if (fresh) {
            Rel<M> rel;
            bool sing = false;
            firk_forward<M, S>(I, T, nest, sh, mesh, y, K, N, lane, scr, loc, rel, sing);
            if (sing || !firk_correction<M, S, false>(I, T, nest, sh, mesh, y, K, N, lane, scr, loc, rel, dy, dK))
                return BVP_RET_FAILURE;
            firk_jt_f<M, S>(I, T, mesh, y, K, N, lane, gy, gK);
            double a = 0.0, b = 0.0, c = 0.0;
            for (int q = lane; q < (N + 1) * M; q += WAVE) {
                a += dy[q] * dy[q];
                b += gy[q] * gy[q];
                c += gy[q] * dy[q];
            }
            for (int q = lane; q < N * S * M; q += WAVE) {
                a += dK[q] * dK[q];
                b += gK[q] * gK[q];
                c += gK[q] * dK[q];
            }
            dn2 = uniform(wave_sum(a));
            l2 = uniform(wave_sum(b));
            gdn = uniform(wave_sum(c));
            __syncthreads();
            jg2 = firk_jac_norm2<M, S>(I, T, mesh, y, K, gy, gK, N, lane);
            fresh = false;
        }
        out.iters++;
        // step d = cn dn + cg g and the predicted reduction of the linear model
        double cn, cg, pred;
        if (sqrt(dn2) <= radius) {
            cn = 1.0; cg = 0.0; pred = ss;
        } else {
            const double lg = sqrt(l2), dcauchy = l2 * lg / jg2;
            if (dcauchy >= radius) {
                cn = 0.0; cg = -radius / lg;
                pred = 2.0 * radius * lg - radius * radius * jg2 / l2;
            } else {
                const double beta = l2 / jg2;
                const double A2 = dn2 + 2.0 * beta * gdn + beta * beta * l2;
                const double B2 = -2.0 * (beta * gdn + beta * beta * l2);
                const double C2 = dcauchy * dcauchy - radius * radius;
                const double tau = (-B2 + sqrt(fmax(0.0, B2 * B2 - 4.0 * A2 * C2))) / (2.0 * A2);
                cn = tau; cg = -(1.0 - tau) * beta;
                pred = ss - (1.0 - tau) * (1.0 - tau) * (ss - l2 * l2 / jg2);
            }
        }
        double mxt;
        const double sst = firk_trial2<M, S>(I, T, mesh, y, K, dy, dK, cn, gy, gK, cg, N, lane, mxt);
        const double rho = (ss - sst) / pred;
        if (rho > 1e-4) {
            for (int q = lane; q < (N + 1) * M; q += WAVE) y[q] += cn * dy[q] + cg * gy[q];
            for (int q = lane; q < N * S * M; q += WAVE) K[q] += cn * dK[q] + cg * gK[q];
            __syncthreads();
            ss = sst;
            mx = mxt;
            fresh = true;
        }
        if (rho < 0.25) {
            radius *= 0.25;
            if (++shrinks > 32) { out.fn = mx; return BVP_RET_FAILURE; }
        } else {
            shrinks = 0;
            if (rho > 0.75) radius *= 2.0;
        }
    }
}


// The NewtonRaphson stages of the polyalgorithm (Core/default_internal_solve.jl:31-45):
// plain, then with the BackTracking line search, each from the same start, stopping
// when max|F| <= tol (checked before every step).  Work areas: dy [cap][M], dK, the
// start copy y0 [cap][M], K0, loc.
template <int M, int S, bool IR = true>
__device__ inline FirkNewtonOut firk_newton(const BatchDev &B, const Inst<M> &I, const FirkTab &T,
                                            const FirkNest<M> &nest, Shared<M> &sh, const double *mesh, double *y,
                                            double *K, int N, int lane, double *scr, double *loc, double *dy,
                                            double *dK, double *y0, double *K0) {
    FirkNewtonOut out{BVP_RET_MAXITERS, 0, -1, NAN};
    // the TrustRegion stage: compiled into the general kernels (IR) only
    const int stages = B.stages & (IR ? 7 : 3);
    bool first = true;
    if (__popc(stages) > 1) firk_copy_state<M, S>(y0, K0, y, K, N, lane);
    for (int st = 0; st < 3; ++st) {
        if (!(stages & (1 << st))) continue;
        if (!first) firk_copy_state<M, S>(y, K, y0, K0, N, lane);
        first = false;
        out.stage = st;
        if constexpr (IR) {
            if (st == 2 && !nest.on) {
                // the last stage: the start copy is free for J'F
                out.ret = firk_trust_region<M, S>(B, I, T, nest, sh, mesh, y, K, N, lane, scr, loc, dy, dK, y0, K0,
                                                  out);
                break;
            }
        }
        for (int it = 0;; ++it) {
            Rel<M> rel;
            bool sing = false;
            const double fn = firk_forward<M, S>(I, T, nest, sh, mesh, y, K, N, lane, scr, loc, rel, sing);
            out.fn = fn;
            if (!(fabs(fn) <= 1.7976931348623157e308)) { out.ret = BVP_RET_UNSTABLE; break; }
            if (fn <= B.newton_tol) { out.ret = BVP_RET_SUCCESS; break; }
            if (it >= B.max_newton) { out.ret = BVP_RET_MAXITERS; break; }
            if (sing || !firk_correction<M, S, IR>(I, T, nest, sh, mesh, y, K, N, lane, scr, loc, rel, dy, dK)) {
                out.ret = BVP_RET_FAILURE;
                break;
            }
            out.iters++;
            double alpha = 1.0;
            if (st == 1) {
                const double ss = firk_trial_sumsq<M, S>(I, T, nest, mesh, y, K, dy, dK, 0.0, N, lane);
                auto phi = [&](double a) {
                    return 0.5 * firk_trial_sumsq<M, S>(I, T, nest, mesh, y, K, dy, dK, a, N, lane);
                };
                if (!backtracking(phi, 0.5 * ss, -ss, alpha)) { out.ret = BVP_RET_FAILURE; break; }
            }
            firk_apply<M, S>(y, K, dy, dK, alpha, N, lane, nest.on);
        }
        if (out.ret == BVP_RET_SUCCESS) break;
    }
    return out;
}

// ------------------------------------------------------------------ defect, interpolation
// eval_q (FIRK/adaptivity.jl:581-606): q(tau h) = y_i + sum_k coeff_k (tau h)^(k+1),
// q' = sum_k (k+1) coeff_k (tau h)^k, coeff_k = (Q K)_k / h^k
template <int M, int S>
__device__ inline void firk_eval_q(const FirkTab &T, const double *yi, double tau, double h, const double (&Kk)[S][M],
                                   double *q, double *qp) {
    const double th = tau * h;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        double v = 0.0, dv = 0.0, p = 1.0, hk = 1.0;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < S; ++j) a += T.Q[k][j] * Kk[j][c];
            const double coef = a / hk;
            dv += (k + 1) * coef * p;
            p *= th;
            v += coef * p;
            hk *= h;
        }
        q[c] = yi[c] + v;
        qp[c] = dv;
    }
}

// Defect pass (defect_estimate!, FIRK/adaptivity.jl:349-405): per interval the relative
// defect of the stage polynomial at tau* and 1 - tau* (the sample with the larger
// max-abs); writes shat_k = (max|d_k| / abstol)^(1/(p+1)) (mesh_selector!,
// FIRK/adaptivity.jl:247-300) and, when errs is set, the vectors; r1, r2 the max and
// sum of shat.  Returns the defect norm (NaN-aware) in all lanes.
template <int M, int S>
__device__ inline double firk_defect_pass(const Inst<M> &I, const FirkTab &T, const double *mesh, const double *y,
                                          const double *K, int N, int lane, double abstol, double *shat,
                                          double *errs, double &r1, double &r2) {
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    double dmx = 0.0, s1 = 0.0, s2 = 0.0;
    bool bad = false;
    const double expo = 1.0 / (T.order + 1);
    for (int k = start; k < start + len; ++k) {
        const double ti = mesh[k], h = mesh[k + 1] - ti;
        double yi[M], Kk[S][M], z[M], zp[M], f[M], d1[M], d2[M];
#pragma unroll
        for (int c = 0; c < M; ++c) yi[c] = y[(size_t)k * M + c];
        load_stages<M, S>(K, k, Kk);
        firk_eval_q<M, S>(T, yi, T.tau, h, Kk, z, zp);
        f_eval<M, double>(I.fam, I.P, ti + T.tau * h, z, f);
        double e1 = 0.0, e2 = 0.0;
#pragma unroll
        for (int c = 0; c < M; ++c) { d1[c] = (zp[c] - f[c]) / (fabs(f[c]) + 1.0); e1 = fmax(e1, fabs(d1[c])); }
        firk_eval_q<M, S>(T, yi, 1.0 - T.tau, h, Kk, z, zp);
        f_eval<M, double>(I.fam, I.P, ti + (1.0 - T.tau) * h, z, f);
#pragma unroll
        for (int c = 0; c < M; ++c) { d2[c] = (zp[c] - f[c]) / (fabs(f[c]) + 1.0); e2 = fmax(e2, fabs(d2[c])); }
        const bool first = e1 > e2;
        double ei = 0.0;
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const double d = first ? d1[c] : d2[c];
            bad |= !(fabs(d) <= 1.7976931348623157e308);
            ei = fmax(ei, fabs(d));
            if (errs) errs[(size_t)k * M + c] = d;
        }
        dmx = fmax(dmx, ei);
        if (shat) {
            const double sh = pow(ei / abstol, expo);
            shat[k] = sh;
            s1 = fmax(s1, sh);
            s2 += sh;
        }
    }
    const bool anybad = wave_any(bad);
    dmx = uniform(wave_max(dmx));
    r1 = uniform(wave_max(s1));
    r2 = uniform(wave_sum(s2));
    __syncthreads();
    return anybad ? NAN : dmx;
}


// interpolant! of FIRKExpandInterpolation / FIRKNestedInterpolation
// (FIRK/interpolation.jl:127-316), and the node values of interp_eval!
// (FIRK/adaptivity.jl:8-135): the quintic of interval(mesh, t) through y_i, y_{i+1},
// f(y_i), f(y_{i+1}) and the stage polynomial at the midpoint.  hscale multiplies the
// interval's step in the stage polynomial and the quintic (the nested form's transfer
// passes the reference's growth factor, FIRK/adaptivity.jl:98-104); with nest.on the
// stages come from the inner solve (at the unscaled step), else from K.
template <int M, int S>
__device__ inline void firk_interp_at(const Inst<M> &I, const FirkTab &T, const FirkNest<M> &nest, const double *mesh,
                                      const double *y, const double *K, int N, double t, double *u, double *du,
                                      double hscale = 1.0) {
    const int j = interval_of(mesh, N, t);
    const double tj = mesh[j], h0 = mesh[j + 1] - tj, tau = t - tj, h = h0 * hscale;
    double yi[M], yj[M], fi[M], fj[M], z[M], zp[M], Kk[S][M];
#pragma unroll
    for (int c = 0; c < M; ++c) { yi[c] = y[(size_t)j * M + c]; yj[c] = y[(size_t)(j + 1) * M + c]; }
    if (nest.on && !K) firk_inner<M, S>(I, T, nest, tj, h0, yi, Kk);
    else load_stages<M, S>(K, j, Kk);
    f_eval<M, double>(I.fam, I.P, tj, yi, fi);
    f_eval<M, double>(I.fam, I.P, mesh[j + 1], yj, fj);
    firk_eval_q<M, S>(T, yi, 0.5, h, Kk, z, zp);
#pragma unroll
    for (int c = 0; c < M; ++c) {
        const double bk[6] = {yi[c], yj[c], z[c], fi[c], fj[c], zp[c]};
        double coef[6];
        firk_quintic(h, bk, coef);
        double v = 0.0, dv = 0.0, p = 1.0, pd = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            v += coef[m] * p;
            dv += m * coef[m] * pd;
            pd = p;
            p *= tau;
        }
        u[c] = v;
        if (du) du[c] = dv;
    }
}

// extend_y (FIRK/firk.jl:76-83): every stage of interval i set to y_{i+1}
template <int M, int S>
__device__ inline void firk_extend(const double *y, double *K, int N, int lane) {
    for (int q = lane; q < N * S * M; q += WAVE) {
        const int i = q / (S * M), c = q % M;
        K[q] = y[(size_t)(i + 1) * M + c];
    }
    __syncthreads();
}

// ------------------------------------------------------------------ full solve of one instance
// The outer loop of FIRK/firk.jl:481-525, 572-621: Newton, defect, mesh_selector! and
// the transfer of the solution (the old quintic interpolant at the new nodes, stages as
// extend_y sets them for a new problem), or halving with a zero restart when Newton or
// the defect threshold failed.
template <int M, int S, bool NEST, bool IR>
__device__ inline void firk_solve_instance(const BatchDev &B, const FirkTab &T, int b, Shared<M> &sh, int lane,
                                           double *scr, double *aux) {
    Inst<M> I;
    load_inst<M>(B, b, I);
    I.minc = 1;
    Bufs<M> bf = bufs_of<M>(B, b);
    double *K = firk_stages_of<M, S>(B, b);
    FirkNest<M> nest = firk_nest_of<M>(B, b);
    nest.on = NEST;  // compile-time forms: the expanded kernels carry no inner-solve code,
    nest.irows = IR;  // and kernels for batches without interior rows no EvalSol code
    const size_t cap = (size_t)B.cap;
    double *dy = aux;                          // [cap][M]
    double *dK = dy + cap * M;                 // [cap][S][M]
    double *y0 = dK + cap * S * M;             // [cap][M]
    double *K0 = y0 + cap * M;                 // [cap][S][M]
    double *loc = K0 + cap * S * M;            // [cap][kFirkLoc]
    double *shat = as_global(B.shat + (size_t)b * B.cap);
    int cur = 0;
    int N = B.N0[b];
    {
        const long long off = B.init_off[b];
        for (int k = lane; k <= N; k += WAVE) {
            bf.mesh[0][k] = B.init_mesh[off + k];
#pragma unroll
            for (int c = 0; c < M; ++c) bf.y[0][(size_t)k * M + c] = B.init_y[(off + k) * M + c];
        }
        __syncthreads();
        if (!nest.on) firk_extend<M, S>(bf.y[0], K, N, lane);
    }
    int info = BVP_RET_SUCCESS, nl_ret = BVP_RET_SUCCESS;
    int outer = 0, iters = 0, fb = 0;
    double last_defect = NAN, rn = NAN;
    bool rn_valid = false;
    long long cnt_it = 0, cnt_def = 0, cnt_tr = 0;
    for (;;) {
        ++outer;
        const double *mesh = bf.mesh[cur];
        double *y = bf.y_of(cur);
        double *mesh_o = bf.mesh[1 - cur], *y_o = bf.y_of(1 - cur);
        FirkNewtonOut nl = firk_newton<M, S, IR>(B, I, T, nest, sh, mesh, y, K, N, lane, scr, loc, dy, dK, y0, K0);
        iters += nl.iters;
        rn = nl.fn;
        rn_valid = true;
        cnt_it += (long long)N * nl.iters;
        if (nl.stage == 1) fb |= 1;
        if (nl.stage == 2) fb |= 16;  // the TrustRegion stage ran
        nl_ret = nl.ret;
        double error_norm = 2.0 * B.abstol;
        info = nl.ret;
        if (!B.adaptive) break;
        double r1 = 0.0, r2 = 0.0;
        if (info == BVP_RET_SUCCESS) {
            error_norm = firk_defect_pass<M, S>(I, T, mesh, y, K, N, lane, B.abstol, shat, nullptr, r1, r2);
            cnt_def += N;
            last_defect = error_norm;
            info = (error_norm > B.defect_threshold || isnan(error_norm)) ? BVP_RET_FAILURE : BVP_RET_SUCCESS;
        }
        if (info == BVP_RET_SUCCESS) {
            if (error_norm > B.abstol) {
                const int Nn = mesh_select(mesh, N, shat, r1, r2, B.max_sub, mesh_o, lane, sh);
                if (Nn < 0) {
                    info = BVP_RET_FAILURE;
                } else {
                    // the nested form restates interp_eval! with its step scaled by the growth
                    // of the node count (FIRK/adaptivity.jl:98-104); the expanded form's
                    // restart is the interpolated nodes with extend_y's stages
                    const double hs = (nest.on && Nn > N) ? (double)Nn / (double)N : 1.0;
                    for (int j = lane; j <= Nn; j += WAVE)
                        firk_interp_at<M, S>(I, T, nest, mesh, y, K, N, mesh_o[j], y_o + (size_t)j * M, nullptr, hs);
                    __syncthreads();
                    if (!nest.on) firk_extend<M, S>(y_o, K, Nn, lane);
                    cnt_tr += N + Nn;
                    N = Nn;
                    cur = 1 - cur;
                    rn_valid = false;
                }
            }
        } else {
            if (2 * N > B.max_sub) {
                info = BVP_RET_FAILURE;
            } else {
                half_mesh_into(mesh, N, mesh_o, lane);
                for (int q = lane; q <= 2 * N * M + M - 1; q += WAVE) y_o[q] = 0.0;
                for (int q = lane; q < 2 * N * S * M; q += WAVE) K[q] = 0.0;
                __syncthreads();
                N = 2 * N;
                cur = 1 - cur;
                rn_valid = false;
                info = BVP_RET_SUCCESS;
            }
        }
        if (!(info == BVP_RET_SUCCESS && error_norm > B.abstol)) break;
        if (outer >= 64) { info = BVP_RET_FAILURE; break; }
    }
    if (!rn_valid) {
        Rel<M> rel;
        bool sing = false;
        rn = firk_forward<M, S>(I, T, nest, sh, bf.mesh[cur], bf.y_of(cur), K, N, lane, scr, loc, rel, sing);
    }
    if (lane == 0) {
        B.err[b] = NAN;
        B.retcode[b] = (nl_ret == BVP_RET_SUCCESS) ? info : nl_ret;
        B.N_final[b] = N;
        B.newton_iters[b] = iters;
        B.outer_iters[b] = outer;
        B.fallback[b] = fb;
        B.cur_buf[b] = cur;
        B.defect_norm[b] = last_defect;
        B.resid_norm[b] = rn;
        B.cnt[(size_t)b * 4 + 0] = cnt_it;
        B.cnt[(size_t)b * 4 + 1] = cnt_def;
        B.cnt[(size_t)b * 4 + 2] = cnt_tr;
        B.cnt[(size_t)b * 4 + 3] = 0;
    }
    __syncthreads();
}

// ------------------------------------------------------------------ kernels
// Persistent FIRK solve: each wave pulls instances from the queue.  NEST: the nested form;
// IR: the general kernels (STANDARD rows inside the span, the TrustRegion stage).
template <int M, int S, bool NEST, bool IR>
__device__ __forceinline__ void firk_solve_queue(const BatchDev &B) {
    __shared__ FirkShared<M> sh;
    const int lane = threadIdx.x;
    const FirkTab &T = *reinterpret_cast<const FirkTab *>(B.firk_tab);
    double *scr = as_global(B.scratch + (size_t)blockIdx.x * B.scratch_per_slot);
    double *aux = as_global(B.aux + (size_t)blockIdx.x * B.aux_per_slot);
    for (;;) {
        const int item = next_item<M>(B.counter, sh, lane);
        if (item >= B.n) break;
        firk_solve_instance<M, S, NEST, IR>(B, T, B.order[item], sh, lane, scr, aux);
    }
}

// Two register-allocation variants the autotuner chooses between: one wave per SIMD
// (few spills) and two waves per SIMD (twice the waves to hide latency, more spills).
template <int M, int S, bool NEST, bool IR = false>
__global__ __launch_bounds__(64, 1) void k_firk_solve(BatchDev B) {
    firk_solve_queue<M, S, NEST, IR>(B);
}
template <int M, int S, bool NEST>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_firk_solve_w2(BatchDev B) {
    firk_solve_queue<M, S, NEST, false>(B);
}

// Component kernels at the packed state (mesh, guess and the packed stages init_K; the
// nested form solves the stages from k0 instead).  Slots of n_x per instance, n_x =
// M (N (S+1) + 1) (expanded) or M (N + 1) (nested):
//   OP_RESIDUAL    out0: the residual in the reference's layout
//   OP_NEWTON_STEP out0: the correction in the unknown order ([y_0, K_0., y_1, ...] or the
//                  node values)
//   OP_DEFECT      out0: [N][M] defect vectors (interval offset), out1: norm
template <int M, int S>
__global__ __launch_bounds__(64, 1) void k_firk_component(BatchDev B, int op) {
    __shared__ FirkShared<M> sh;
    const int lane = threadIdx.x;
    const FirkTab &T = *reinterpret_cast<const FirkTab *>(B.firk_tab);
    double *scr = as_global(B.scratch + (size_t)blockIdx.x * B.scratch_per_slot);
    double *aux = as_global(B.aux + (size_t)blockIdx.x * B.aux_per_slot);
    const size_t cap = (size_t)B.cap;
    for (;;) {
        const int b = next_item<M>(B.counter, sh, lane);
        if (b >= B.n) break;
        Inst<M> I;
        load_inst<M>(B, b, I);
        I.minc = 1;
        const int N = B.N0[b];
        const long long off = B.init_off[b];
        const long long ioff = off - b;  // interval offset
        const double *mesh = as_global(B.init_mesh + off);
        const double *y = as_global(B.init_y + off * M);
        double *dy = aux, *dK = dy + cap * M, *loc = dK + cap * S * M;
        const FirkNest<M> nest = firk_nest_of<M>(B, b);
        // the stage state: the packed input, or (nested) the inner solves' stages in the
        // slot's dK area, which the nested correction does not use
        double *K = nest.on ? dK : const_cast<double *>(as_global(B.init_K + ioff * S * M));
        const size_t xo = nest.on ? (size_t)M * off : (size_t)M * ((size_t)(S + 1) * ioff + b);  // slot offset
        const int rowlen = nest.on ? M : (S + 1) * M;  // residual entries per interval
        if (op == OP_RESIDUAL) {
            double *out = B.out0 + xo;
            int nbca = 0;
            if (I.bc_kind == BVP_BC_TWO_POINT)
                for (int k = 0; k < M; ++k) nbca += (sel<M>(I.bct, k) == I.t0) ? 1 : 0;
            else
                nbca = M;
            int P, start, len;
            chunk_of(N, lane, P, start, len);
            for (int k = start; k < start + len; ++k) {
                const double ti = mesh[k], h = mesh[k + 1] - ti;
                double yi[M], yj[M], Kk[S][M], G[S][M], Phi[M], J[S][M][M];
#pragma unroll
                for (int c = 0; c < M; ++c) { yi[c] = y[(size_t)k * M + c]; yj[c] = y[(size_t)(k + 1) * M + c]; }
                if (nest.on) firk_inner<M, S>(I, T, nest, ti, h, yi, Kk);
                else load_stages<M, S>(K, k, Kk);
                firk_interval<M, S, false>(I, T, ti, h, yi, yj, Kk, G, Phi, J);
                double *o = out + nbca + (size_t)k * rowlen;
#pragma unroll
                for (int c = 0; c < M; ++c) o[c] = Phi[c];
                if (!nest.on) {
#pragma unroll
                    for (int r = 0; r < S; ++r)
#pragma unroll
                        for (int c = 0; c < M; ++c) o[(size_t)(r + 1) * M + c] = G[r][c];
                }
            }
            if (lane == 0) {
                // BC rows: input order; TWO_POINT splits them into bca (before) and bcb (after)
                int ia = 0, ib = 0;
                for (int k = 0; k < M; ++k) {
                    const BcLoc L = bc_locate<M>(I, k, mesh, N);
                    const double v = firk_bc_row<M, S>(I, T, nest, mesh, y, K, nullptr, nullptr, 0.0, k, L);
                    if (I.bc_kind == BVP_BC_TWO_POINT && sel<M>(I.bct, k) != I.t0)
                        out[nbca + (size_t)N * rowlen + (ib++)] = v;
                    else
                        out[ia++] = v;
                }
            }
        } else if (op == OP_NEWTON_STEP) {
            Rel<M> rel;
            bool sing = false;
            firk_forward<M, S>(I, T, nest, sh, mesh, y, K, N, lane, scr, loc, rel, sing);
            const bool ok = !sing && firk_correction<M, S>(I, T, nest, sh, mesh, y, K, N, lane, scr, loc, rel, dy, dK);
            double *out = B.out0 + xo;
            for (int q = lane; q < (N + 1) * M; q += WAVE) {
                const int i = q / M, c = q % M;
                out[(size_t)i * rowlen + c] = ok ? dy[q] : NAN;
            }
            for (int q = lane; q < (nest.on ? 0 : N * S * M); q += WAVE) {
                const int i = q / (S * M), rc = q % (S * M);
                out[(size_t)i * (S + 1) * M + M + rc] = ok ? dK[q] : NAN;
            }
        } else if (op == OP_DEFECT) {
            if (nest.on) {
                // the stages from the inner solves, as defect_estimate! solves them
                // (FIRK/adaptivity.jl:465-517)
                for (int k = lane; k < N; k += WAVE) {
                    double yi[M], Kk[S][M];
#pragma unroll
                    for (int c = 0; c < M; ++c) yi[c] = y[(size_t)k * M + c];
                    firk_inner<M, S>(I, T, nest, mesh[k], mesh[k + 1] - mesh[k], yi, Kk);
                    store_stages<M, S>(K, k, Kk);
                }
                __syncthreads();
            }
            double r1, r2;
            const double dn = firk_defect_pass<M, S>(I, T, mesh, y, K, N, lane, B.abstol, nullptr,
                                                     B.out0 + ioff * M, r1, r2);
            if (lane == 0) B.out1[b] = dn;
        }
        __syncthreads();
    }
}

// Dense output: one thread per query point (qinst[q] its instance, tq[q] its t).
template <int M, int S>
__global__ __launch_bounds__(256) void k_firk_interp(BatchDev B, const int *qinst, const double *tq, double *uo,
                                                     double *duo, long long nq) {
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    const FirkTab &T = *reinterpret_cast<const FirkTab *>(B.firk_tab);
    const int b = qinst[q];
    Inst<M> I;
    load_inst<M>(B, b, I);
    const long long off = B.init_off[b];
    const long long ioff = off - b;
    const double *mesh = as_global(B.init_mesh + off);
    const double *y = as_global(B.init_y + off * M);
    const int N = B.N0[b];
    const double t = tq[q];
    const FirkNest<M> nest = firk_nest_of<M>(B, b);
    const double *K = nest.on ? nullptr : as_global(B.init_K + ioff * S * M);
    double u[M], du[M];
    firk_interp_at<M, S>(I, T, nest, mesh, y, K, N, t, u, du);
#pragma unroll
    for (int c = 0; c < M; ++c) {
        uo[q * M + c] = u[c];
        if (duo) duo[q * M + c] = du[c];
    }
}

// Final stage values of each instance to out at interval offsets out_off (one block
// per instance).
template <int M, int S>
__global__ __launch_bounds__(256) void k_firk_gather(BatchDev B, const long long *out_off, double *out) {
    const int b = blockIdx.x;
    if (b >= B.n) return;
    const int N = B.N_final[b];
    const double *K = firk_stages_of<M, S>(B, b);
    const long long o = out_off[b] * S * M;
    for (int q = threadIdx.x; q < N * S * M; q += blockDim.x) out[o + q] = K[q];
}

// The two-waves-per-SIMD register variant where the compiler reaches that occupancy
// (M = 1; M = 2 with s <= 3; M = 3 with s <= 2); elsewhere both variants launch the
// one-wave kernels.
template <int M, int S> constexpr bool kFirkW2 = M == 1 || (M <= 3 && M * S <= 6);

template <int M, int S, int WPE>
hipError_t firk_launch_solve_t(const BatchDev &B, int blocks, hipStream_t s) {
    if (B.firk_irows) {
        // batches with STANDARD rows inside the span: the one-wave-per-SIMD kernels only
        if (B.firk_nested) k_firk_solve<M, S, true, true><<<blocks, 64, 0, s>>>(B);
        else k_firk_solve<M, S, false, true><<<blocks, 64, 0, s>>>(B);
    } else if (B.firk_nested) {
        if constexpr (WPE == 2 && kFirkW2<M, S>) k_firk_solve_w2<M, S, true><<<blocks, 64, 0, s>>>(B);
        else k_firk_solve<M, S, true><<<blocks, 64, 0, s>>>(B);
    } else {
        if constexpr (WPE == 2 && kFirkW2<M, S>) k_firk_solve_w2<M, S, false><<<blocks, 64, 0, s>>>(B);
        else k_firk_solve<M, S, false><<<blocks, 64, 0, s>>>(B);
    }
    return hipGetLastError();
}
template <int M, int S>
hipError_t firk_launch_component_t(const BatchDev &B, int op, int blocks, hipStream_t s) {
    k_firk_component<M, S><<<blocks, 64, 0, s>>>(B, op);
    return hipGetLastError();
}
template <int M, int S>
hipError_t firk_launch_interp_t(const BatchDev &B, const int *qinst, const double *tq, double *u, double *du,
                                long long nq, hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((nq + 255) / 256);
    k_firk_interp<M, S><<<nb, 256, 0, s>>>(B, qinst, tq, u, du, nq);
    return hipGetLastError();
}
// resident FIRK solve waves per CU (registers and LDS) of variant WPE
template <int M, int S, int WPE>
int firk_occupancy_t(int nested) {
    int nb = 0;
    hipError_t e;
    if (nested) {
        if constexpr (WPE == 2 && kFirkW2<M, S>)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_firk_solve_w2<M, S, true>, 64, 0);
        else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_firk_solve<M, S, true>, 64, 0);
    } else {
        if constexpr (WPE == 2 && kFirkW2<M, S>)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_firk_solve_w2<M, S, false>, 64, 0);
        else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_firk_solve<M, S, false>, 64, 0);
    }
    if (e != hipSuccess) return 4 * WPE;
    return nb;
}
template <int M, int S>
hipError_t firk_launch_gather_t(const BatchDev &B, const long long *out_off, double *out, hipStream_t s) {
    k_firk_gather<M, S><<<B.n, 256, 0, s>>>(B, out_off, out);
    return hipGetLastError();
}

}  // namespace bvpdev
