// mirk_wave_impl.h -- MIRK collocation on gfx950, one 64-lane wavefront per BVP
// instance, for small state dimensions (M <= 4).
//
// Layout of one instance's work in the wave:
//   * the N mesh intervals are split into P = min(N, 64) contiguous chunks, one
//     per lane;
//   * forward pass: each lane walks its chunk once, computing for interval k the
//     residual Phi_k and the Jacobian blocks S_k = dPhi_k/dy_k, R_k = dPhi_k/dy_{k+1}
//     together with forward-mode dual numbers (2M directions), and eliminates
//     the interior chunk nodes as it goes (stable block elimination with row
//     pivoting on the 2M x M pivot column block).  What back-substitution needs
//     is written to per-wave scratch in a lane-interleaved layout
//     ([step][field][lane]: every store is one 512-byte coalesced line);
//   * the P chunk relations "E dy_a + F dy_b = g" are merged pairwise up a
//     binary tree across lanes (same elimination step), leaving one relation
//     between the two end nodes; with the M boundary rows this is a 2M x 2M
//     dense system solved with partial pivoting on lane 0;
//   * corrections flow back down the tree and then backwards through each
//     chunk, updating y in place.
// The reference's counterparts (paths relative to the reference root,
// MIRK/ = lib/BoundaryValueDiffEqMIRK/src/, Core/ = lib/BoundaryValueDiffEqCore/src/):
//   residual Phi           MIRK/collocation.jl:44-102
//   Jacobian blocks        MIRK/mirk.jl:697-862, 927-1002 (sparse ForwardDiff)
//   linear solve           LinearSolve on the almost-banded matrix (Core/internal_problems.jl:107-169)
//   Newton                 Core/default_internal_solve.jl:31-45 (NewtonRaphson, + backtracking)
//   defect                 MIRK/adaptivity.jl:370-461, MIRK/interpolation.jl:300-372
//   mesh selection         MIRK/adaptivity.jl:23-75, 250-304
//   transfer, outer loop   MIRK/mirk.jl:286-388, MIRK/adaptivity.jl:6-13
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "anchors.h"
#include "mirk_device.h"
#include "mirk_wave.h"
#include "bvp_internal.h"

namespace bvpdev {

constexpr int WAVE = 64;
constexpr int MAX_OUTER = 64;

// ------------------------------------------------------------------ helpers
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline bool wave_any(bool p) { return __any(p ? 1 : 0) != 0; }
// v as a wave-uniform value (lane 0's; after a reduction every lane holds the same
// bits), so that the compiler can keep it in scalar registers
__device__ inline double uniform(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffLL));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// p as a pointer into global memory.  Pointers that reach the compiler as
// generic (flat) pointers -- read back from a runtime-indexed array, or derived
// in ways it cannot follow -- compile to flat_* accesses, and the use of any
// flat load's result waits for every outstanding memory operation
// (vmcnt(0) lgkmcnt(0)), which serialises loads that were issued ahead.
template <typename T>
__device__ __forceinline__ T *as_global(T *p) {
    return (T *)(__attribute__((address_space(1))) T *)p;
}

// a[k] for a runtime k without dynamic register indexing (which would move the
// array to scratch memory)
template <int N, typename T>
__device__ inline T sel(const T (&a)[N], int k) {
    T r = a[0];
#pragma unroll
    for (int i = 1; i < N; ++i) r = (k == i) ? a[i] : r;
    return r;
}

// minc: at least this many intervals per chunk (the interval-per-lane mapping the
// autotuner chooses; 1 spreads the intervals over up to 64 lanes)
__device__ inline void chunk_of(int N, int lane, int &P, int &start, int &len, int minc = 1) {
    P = N / minc;
    P = P < 1 ? 1 : (P > WAVE ? WAVE : P);
    if (lane >= P) { start = N; len = 0; return; }
    const int q = N / P, r = N % P;
    len = q + (lane < r ? 1 : 0);
    start = lane * q + (lane < r ? lane : r);
}

// Walks intervals start .. start+len-1 of a lane's chunk in order and calls
// body(k, t_k, t_{k+1}, y_k, y_{k+1}) for each.  The nodes are read PF intervals
// ahead into a ring of registers: the body is unrolled PF times so that every
// ring slot has a fixed register, and the load of node k+1+PF is issued before
// interval k is evaluated, so that PF intervals of arithmetic cover its latency.
template <int M, int PF, typename Body>
__device__ __forceinline__ void walk_chunk(const double *mesh, const double *y, int start, int len, Body &&body) {
    if (len <= 0) return;
    double ti = mesh[start], yi[M];
#pragma unroll
    for (int c = 0; c < M; ++c) yi[c] = y[(size_t)start * M + c];
    double rt[PF], ry[PF][M];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        if (j < len) {
            rt[j] = mesh[start + 1 + j];
#pragma unroll
            for (int c = 0; c < M; ++c) ry[j][c] = y[(size_t)(start + 1 + j) * M + c];
        }
    }
    for (int i0 = 0; i0 < len; i0 += PF) {
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            const int i = i0 + j;
            if (i < len) {
                const double tj = rt[j];
                double yj[M];
#pragma unroll
                for (int c = 0; c < M; ++c) yj[c] = ry[j][c];
                if (i + PF < len) {
                    const int nd = start + 1 + i + PF;
                    rt[j] = mesh[nd];
#pragma unroll
                    for (int c = 0; c < M; ++c) ry[j][c] = y[(size_t)nd * M + c];
                }
                body(start + i, ti, tj, yi, yj);
                ti = tj;
#pragma unroll
                for (int c = 0; c < M; ++c) yi[c] = yj[c];
            }
        }
    }
}

// Nodes in flight per chunk walk: deeper for the light passes, shallow where the
// dual-number pass sets the kernel's register peak.
template <int M> constexpr int kPfPrimal = M <= 2 ? 3 : 2;
template <int M> constexpr int kPfDefect = M <= 2 ? 4 : 2;
template <int M> constexpr int kPfDual = M <= 2 ? 2 : 1;
template <int M> constexpr int kPfBack = M <= 2 ? 2 : 1;

// interval(mesh, t) = clamp(searchsortedfirst(mesh, t) - 1, 1, N), 0-based (Core/utils.jl:119-121)
__device__ inline int interval_of(const double *mesh, int N, double t) {
    int lo = 0, hi = N + 1;  // first index with mesh[i] >= t, in [0, N+1]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (mesh[mid] < t) lo = mid + 1; else hi = mid;
    }
    int iv = lo;  // 1-based interval = lo (searchsortedfirst - 1 with 1-based +1)
    iv = iv < 1 ? 1 : (iv > N ? N : iv);
    return iv - 1;
}

// ------------------------------------------------------------------ stages
// K_r = f(Y_r, t + c_r h), Y_r = (1-v_r) yi + v_r yj + h sum_{j<r} x_rj K_j (MIRK/collocation.jl:60-66)
template <int M, int METHOD, typename T>
__device__ inline void stages(int fam, const Params<M> &P, double ti, double h, const T *yi, const T *yj,
                              T (&K)[Tab<METHOD>::s][M]) {
    using TB = Tab<METHOD>;
#pragma unroll
    for (int r = 0; r < TB::s; ++r) {
        T Y[M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            T a = (1.0 - TB::v(r)) * yi[c];
            Y[c] = a + TB::v(r) * yj[c];
        }
        if (r > 0) {
#pragma unroll
            for (int c = 0; c < M; ++c) {
                T acc = ConstOf<T>::mk(0.0);
                bool any = false;
#pragma unroll
                for (int j = 0; j < r; ++j) {
                    if (TB::x(r, j) != 0.0) {
                        if (!any) { acc = TB::x(r, j) * K[j][c]; any = true; }
                        else acc = acc + TB::x(r, j) * K[j][c];
                    }
                }
                if (any) Y[c] = Y[c] + h * acc;
            }
        }
        f_eval<M, T>(fam, P, ti + TB::c(r) * h, Y, K[r]);
    }
}

// Phi = (yj - yi) - h sum_r b_r K_r   (MIRK/collocation.jl:69-70)
template <int M, int METHOD, typename T>
__device__ inline void phi_of(double h, const T *yi, const T *yj, const T (&K)[Tab<METHOD>::s][M], T *phi) {
    using TB = Tab<METHOD>;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        T acc = TB::b(0) * K[0][c];
#pragma unroll
        for (int r = 1; r < TB::s; ++r) acc = acc + TB::b(r) * K[r][c];
        phi[c] = (yj[c] - yi[c]) - h * acc;
    }
}

// extra interpolation stages k*_r (MIRK/interpolation.jl:337-372), primal
template <int M, int METHOD>
__device__ inline void interp_stages(int fam, const Params<M> &P, double ti, double h, const double *yi,
                                     const double *yj, const double (&K)[Tab<METHOD>::s][M],
                                     double (&KI)[Tab<METHOD>::sstar - Tab<METHOD>::s][M]) {
    using TB = Tab<METHOD>;
    constexpr int NS = TB::sstar - TB::s;
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        double Y[M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < TB::s; ++j)
                if (TB::xs(r, j) != 0.0) a += K[j][c] * TB::xs(r, j);
#pragma unroll
            for (int j = 0; j < r; ++j)
                if (TB::xs(r, TB::s + j) != 0.0) a += KI[j][c] * TB::xs(r, TB::s + j);
            Y[c] = a * h + (1.0 - TB::vs(r)) * yi[c] + TB::vs(r) * yj[c];
        }
        f_eval<M, double>(fam, P, ti + TB::cs(r) * h, Y, KI[r]);
    }
}

// z = yi + h (K w + KI w*), z' = K w' + KI w'*   (MIRK/adaptivity.jl:623-666)
template <int M, int METHOD>
__device__ inline void sum_stages(double h, const double *yi, const double (&K)[Tab<METHOD>::s][M],
                                  const double (&KI)[Tab<METHOD>::sstar - Tab<METHOD>::s][M], const double *w,
                                  const double *wp, double *z, double *zp) {
    using TB = Tab<METHOD>;
    constexpr int NS = TB::sstar - TB::s;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        double a = 0.0, ap = 0.0;
#pragma unroll
        for (int r = 0; r < TB::s; ++r) { a += K[r][c] * w[r]; if (zp) ap += K[r][c] * wp[r]; }
#pragma unroll
        for (int r = 0; r < NS; ++r) { a += KI[r][c] * w[TB::s + r]; if (zp) ap += KI[r][c] * wp[TB::s + r]; }
        z[c] = a * h + yi[c];
        if (zp) zp[c] = ap;
    }
}

// Residual and Jacobian blocks of one interval with dual numbers.
template <int M, int METHOD>
__device__ inline void interval_dual(int fam, const Params<M> &P, double ti, double h, const double *yi,
                                     const double *yj, double *phi, double (&S)[M][M], double (&R)[M][M],
                                     double (&Kp)[Tab<METHOD>::s][M]) {
    using D = Dual<2 * M>;
    D a[M], b[M];
#pragma unroll
    for (int c = 0; c < M; ++c) {
        a[c] = dconst<2 * M>(yi[c]);
        a[c].d[c] = 1.0;
        b[c] = dconst<2 * M>(yj[c]);
        b[c].d[M + c] = 1.0;
    }
    D K[Tab<METHOD>::s][M];
    stages<M, METHOD, D>(fam, P, ti, h, a, b, K);
    D ph[M];
    phi_of<M, METHOD, D>(h, a, b, K, ph);
#pragma unroll
    for (int i = 0; i < M; ++i) {
        phi[i] = ph[i].v;
#pragma unroll
        for (int j = 0; j < M; ++j) { S[i][j] = ph[i].d[j]; R[i][j] = ph[i].d[M + j]; }
    }
#pragma unroll
    for (int r = 0; r < Tab<METHOD>::s; ++r)
#pragma unroll
        for (int c = 0; c < M; ++c) Kp[r][c] = K[r][c].v;
}

// f and its Jacobian df/dy at (t, y) for the families the specialised M = 2
// kernels serve; f is the same arithmetic as f_eval's.
template <int FAM>
__device__ inline void f_jac2(const Params<2> &P, double t, const double *y, double *f, double (&J)[2][2]) {
    static_assert(FAM == BVP_FAMILY_BRATU || FAM == BVP_FAMILY_SINGPERT, "family without an analytic Jacobian");
    if constexpr (FAM == BVP_FAMILY_BRATU) {  // (y2, -lambda e^y1)
        const double e = exp(y[0]);
        f[0] = y[1];
        f[1] = -P.p[0] * e;
        J[0][0] = 0.0; J[0][1] = 1.0;
        J[1][0] = -P.p[0] * e; J[1][1] = 0.0;
    } else {  // (y2, (y1 - t y2) / eps)
        f[0] = y[1];
        f[1] = (y[0] - t * y[1]) / P.p[0];
        J[0][0] = 0.0; J[0][1] = 1.0;
        J[1][0] = 1.0 / P.p[0]; J[1][1] = -t / P.p[0];
    }
}

// Residual and Jacobian blocks of one interval by the chain rule through the
// stages (the specialised kernels): dK_r/dy = J_r dY_r/dy with
// dY_r/dy_i = (1 - v_r) I + h sum_{j<r} x_rj dK_j/dy_i (and v_r I + ... for y_j),
// S = -I - h sum_r b_r dK_r/dy_i, R = I - h sum_r b_r dK_r/dy_j -- the same
// blocks interval_dual forms with dual numbers (MIRK/mirk.jl:697-862), in
// M x M products instead of 2M-direction arithmetic through every operation.
// The stage values and Phi are the same arithmetic as stages<> and phi_of<>.
template <int M, int METHOD, int FAM>
__device__ inline void interval_chain(const Params<M> &P, double ti, double h, const double *yi, const double *yj,
                                      double *phi, double (&S)[M][M], double (&R)[M][M],
                                      double (&Kp)[Tab<METHOD>::s][M]) {
    using TB = Tab<METHOD>;
    double Di[TB::s][M][M], Dj[TB::s][M][M];
#pragma unroll
    for (int r = 0; r < TB::s; ++r) {
        double Y[M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            double a = (1.0 - TB::v(r)) * yi[c];
            Y[c] = a + TB::v(r) * yj[c];
        }
        double Yi[M][M], Yj[M][M];
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
            for (int b = 0; b < M; ++b) {
                Yi[a][b] = (a == b) ? (1.0 - TB::v(r)) : 0.0;
                Yj[a][b] = (a == b) ? TB::v(r) : 0.0;
            }
        if (r > 0) {
#pragma unroll
            for (int c = 0; c < M; ++c) {
                double acc = 0.0;
                bool any = false;
#pragma unroll
                for (int j = 0; j < r; ++j) {
                    if (TB::x(r, j) != 0.0) {
                        if (!any) { acc = TB::x(r, j) * Kp[j][c]; any = true; }
                        else acc = acc + TB::x(r, j) * Kp[j][c];
                    }
                }
                if (any) Y[c] = Y[c] + h * acc;
            }
#pragma unroll
            for (int j = 0; j < r; ++j) {
                if (TB::x(r, j) != 0.0) {
                    const double hx = h * TB::x(r, j);
#pragma unroll
                    for (int a = 0; a < M; ++a)
#pragma unroll
                        for (int b = 0; b < M; ++b) {
                            Yi[a][b] += hx * Di[j][a][b];
                            Yj[a][b] += hx * Dj[j][a][b];
                        }
                }
            }
        }
        double J[M][M];
        f_jac2<FAM>(P, ti + TB::c(r) * h, Y, Kp[r], J);
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
            for (int b = 0; b < M; ++b) {
                double si = 0.0, sj = 0.0;
#pragma unroll
                for (int q = 0; q < M; ++q) {
                    si += J[a][q] * Yi[q][b];
                    sj += J[a][q] * Yj[q][b];
                }
                Di[r][a][b] = si;
                Dj[r][a][b] = sj;
            }
    }
    phi_of<M, METHOD, double>(h, yi, yj, Kp, phi);
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = 0; b < M; ++b) {
            double ai = TB::b(0) * Di[0][a][b], aj = TB::b(0) * Dj[0][a][b];
#pragma unroll
            for (int r = 1; r < TB::s; ++r) {
                ai = ai + TB::b(r) * Di[r][a][b];
                aj = aj + TB::b(r) * Dj[r][a][b];
            }
            S[a][b] = ((a == b) ? -1.0 : 0.0) - h * ai;
            R[a][b] = ((a == b) ? 1.0 : 0.0) - h * aj;
        }
}

// ------------------------------------------------------------------ elimination
// Relation E dy_a + F dy_b = g between two nodes.
template <int M> struct Rel {
    double E[M][M], F[M][M], g[M];
};
// What back-substitution of one eliminated node needs: dy_b = h - X dy_a - Z dy_c.
template <int M> struct Back {
    double X[M][M], Z[M][M], h[M];
};
template <int M> constexpr int kRelFields = 2 * M * M + M;

// Eliminate the shared node b of  L: E1 dy_a + F1 dy_b = g1  and  S dy_b + R dy_c = r
// by Gaussian elimination with partial pivoting on the stacked 2M x M column
// block [F1; S].  Produces the relation for (a, c) and the back-substitution
// rows of b.
template <int M>
__device__ inline void eliminate(const Rel<M> &L, const double (&S)[M][M], const double (&R)[M][M],
                                 const double *r, Rel<M> &out, Back<M> &bk) {
    constexpr int M2 = 2 * M;
    double A[M2][M], B1[M2][M], B2[M2][M], q[M2];
#pragma unroll
    for (int i = 0; i < M; ++i) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            A[i][j] = L.F[i][j]; B1[i][j] = L.E[i][j]; B2[i][j] = 0.0;
            A[M + i][j] = S[i][j]; B1[M + i][j] = 0.0; B2[M + i][j] = R[i][j];
        }
        q[i] = L.g[i];
        q[M + i] = r[i];
    }
#pragma unroll
    for (int j = 0; j < M; ++j) {
        // pivot search over rows j..2M-1 (first maximum, as LAPACK's idamax)
        int p = j;
        double best = fabs(A[j][j]);
#pragma unroll
        for (int i = j + 1; i < M2; ++i) {
            const double v = fabs(A[i][j]);
            if (v > best) { best = v; p = i; }
        }
        // swap rows j and p by selects (keeps the arrays in registers)
#pragma unroll
        for (int i = j + 1; i < M2; ++i) {
            const bool sw = (i == p);
#pragma unroll
            for (int c = 0; c < M; ++c) {
                double t;
                t = A[j][c]; A[j][c] = sw ? A[i][c] : A[j][c]; A[i][c] = sw ? t : A[i][c];
                t = B1[j][c]; B1[j][c] = sw ? B1[i][c] : B1[j][c]; B1[i][c] = sw ? t : B1[i][c];
                t = B2[j][c]; B2[j][c] = sw ? B2[i][c] : B2[j][c]; B2[i][c] = sw ? t : B2[i][c];
            }
            double t = q[j];
            q[j] = sw ? q[i] : q[j];
            q[i] = sw ? t : q[i];
        }
        const double inv = 1.0 / A[j][j];
#pragma unroll
        for (int i = j + 1; i < M2; ++i) {
            const double m = A[i][j] * inv;
#pragma unroll
            for (int c = j + 1; c < M; ++c) A[i][c] -= m * A[j][c];
#pragma unroll
            for (int c = 0; c < M; ++c) { B1[i][c] -= m * B1[j][c]; B2[i][c] -= m * B2[j][c]; }
            q[i] -= m * q[j];
        }
    }
    // back-substitution rows: solve U [X Z h] = [B1 B2 q] (top M rows)
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
        const double inv = 1.0 / A[i][i];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            double x = B1[i][c], z = B2[i][c];
#pragma unroll
            for (int k = i + 1; k < M; ++k) { x -= A[i][k] * bk.X[k][c]; z -= A[i][k] * bk.Z[k][c]; }
            bk.X[i][c] = x * inv;
            bk.Z[i][c] = z * inv;
        }
        double hh = q[i];
#pragma unroll
        for (int k = i + 1; k < M; ++k) hh -= A[i][k] * bk.h[k];
        bk.h[i] = hh * inv;
    }
#pragma unroll
    for (int i = 0; i < M; ++i) {
#pragma unroll
        for (int c = 0; c < M; ++c) { out.E[i][c] = B1[M + i][c]; out.F[i][c] = B2[M + i][c]; }
        out.g[i] = q[M + i];
    }
}

template <int M>
__device__ inline void back_apply(const Back<M> &bk, const double *ya, const double *yc, double *yb) {
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double v = bk.h[i];
#pragma unroll
        for (int c = 0; c < M; ++c) v -= bk.X[i][c] * ya[c] + bk.Z[i][c] * yc[c];
        yb[i] = v;
    }
}

// scratch I/O of Back records; lane-interleaved: [slot][field][64]
// STRIDE = WAVE: chunk records of all lanes interleaved; STRIDE = 1: compact tree records.
template <int M, int STRIDE = WAVE>
__device__ inline void store_back(double *base, int slot, int lane, const Back<M> &bk) {
    constexpr int NF = kRelFields<M>;
    double *p = base + (size_t)slot * NF * STRIDE + lane;
    int f = 0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) p[(f++) * STRIDE] = bk.X[i][c];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) p[(f++) * STRIDE] = bk.Z[i][c];
#pragma unroll
    for (int i = 0; i < M; ++i) p[(f++) * STRIDE] = bk.h[i];
}
template <int M, int STRIDE = WAVE>
__device__ inline void load_back(const double *base, int slot, int lane, Back<M> &bk) {
    constexpr int NF = kRelFields<M>;
    const double *p = base + (size_t)slot * NF * STRIDE + lane;
    int f = 0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) bk.X[i][c] = p[(f++) * STRIDE];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) bk.Z[i][c] = p[(f++) * STRIDE];
#pragma unroll
    for (int i = 0; i < M; ++i) bk.h[i] = p[(f++) * STRIDE];
}

// Chunk records stored in FP32 (X, Z and h): a slot holds 2 M^2 + M floats per
// lane, [field][64 lanes], every field a contiguous run of 64 values (256 bytes).
// The elimination itself runs in FP64; Newton's next residual, evaluated in FP64,
// corrects the rounding of the record (the correction is an inexact Newton step).
template <int M> constexpr int kRecFloatsF32 = 2 * M * M + M;
template <int M>
__device__ inline void store_back_f32(double *base, int slot, int lane, const Back<M> &bk) {
    float *q = reinterpret_cast<float *>(base) + (size_t)slot * kRecFloatsF32<M> * WAVE + lane;
    int f = 0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) q[(f++) * WAVE] = (float)bk.X[i][c];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) q[(f++) * WAVE] = (float)bk.Z[i][c];
#pragma unroll
    for (int i = 0; i < M; ++i) q[(f++) * WAVE] = (float)bk.h[i];
}
template <int M>
__device__ inline void load_back_f32(const double *base, int slot, int lane, Back<M> &bk) {
    const float *q = reinterpret_cast<const float *>(base) + (size_t)slot * kRecFloatsF32<M> * WAVE + lane;
    int f = 0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) bk.X[i][c] = (double)q[(f++) * WAVE];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) bk.Z[i][c] = (double)q[(f++) * WAVE];
#pragma unroll
    for (int i = 0; i < M; ++i) bk.h[i] = (double)q[(f++) * WAVE];
}
// chunk record of slot `slot` in the storage precision F32 (a template parameter of
// the solve path, so that each precision has its own register allocation)
template <int M, bool F32>
__device__ inline void put_rec(double *base, int slot, int lane, const Back<M> &bk) {
    if constexpr (F32) store_back_f32<M>(base, slot, lane, bk);
    else store_back<M>(base, slot, lane, bk);
}
template <int M, bool F32>
__device__ inline void get_rec(const double *base, int slot, int lane, Back<M> &bk) {
    if constexpr (F32) load_back_f32<M>(base, slot, lane, bk);
    else load_back<M>(base, slot, lane, bk);
}

// tree records live after the chunk records: index of the merge of lane l at level d
__device__ inline int tree_index(int d, int l) { return (WAVE - (WAVE >> d)) + (l >> (d + 1)); }

// Dense n x n solve with partial pivoting (n = 2M), single lane.
template <int NN>
__device__ inline bool dense_solve(double (&A)[NN][NN], double (&b)[NN]) {
#pragma unroll
    for (int j = 0; j < NN; ++j) {
        int p = j;
        double best = fabs(A[j][j]);
#pragma unroll
        for (int i = j + 1; i < NN; ++i) if (fabs(A[i][j]) > best) { best = fabs(A[i][j]); p = i; }
#pragma unroll
        for (int i = j + 1; i < NN; ++i) {
            const bool sw = (i == p);
#pragma unroll
            for (int c = 0; c < NN; ++c) {
                const double t = A[j][c];
                A[j][c] = sw ? A[i][c] : A[j][c];
                A[i][c] = sw ? t : A[i][c];
            }
            const double t = b[j];
            b[j] = sw ? b[i] : b[j];
            b[i] = sw ? t : b[i];
        }
        if (A[j][j] == 0.0) return false;
        const double inv = 1.0 / A[j][j];
#pragma unroll
        for (int i = j + 1; i < NN; ++i) {
            const double m = A[i][j] * inv;
#pragma unroll
            for (int c = j + 1; c < NN; ++c) A[i][c] -= m * A[j][c];
            b[i] -= m * b[j];
        }
    }
#pragma unroll
    for (int i = NN - 1; i >= 0; --i) {
        double v = b[i];
#pragma unroll
        for (int c = i + 1; c < NN; ++c) v -= A[i][c] * b[c];
        b[i] = v / A[i][i];
    }
    return true;
}

// ------------------------------------------------------------------ instance view
template <int M> struct Inst {
    int fam;
    int minc;  // chunk_of's minimum intervals per lane
    Params<M> P;
    int bc_kind;
    double bct[M], bcv[M];
    int bcc[M];
    double t0, t1;
};

template <int M>
__device__ inline void load_inst(const BatchDev &B, int b, Inst<M> &I) {
    I.fam = B.family[b];
    I.minc = B.min_chunk > 0 ? B.min_chunk : 1;
    const int po = B.param_off[b], pn = B.param_cnt[b];
#pragma unroll
    for (int k = 0; k < Params<M>::MAXP; ++k) I.P.p[k] = k < pn ? B.params[po + k] : 0.0;
    I.bc_kind = B.bc_kind[b];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        I.bct[k] = B.bc_t[(size_t)b * M + k];
        I.bcv[k] = B.bc_val[(size_t)b * M + k];
        I.bcc[k] = B.bc_comp[(size_t)b * M + k];
    }
    I.t0 = B.t0[b];
    I.t1 = B.t1[b];
}

// BC row resolution on the current mesh (Core/utils.jl:157-209; EvalSol
// MIRK/interpolation.jl:214-231): node carrying the Jacobian entry, and for an
// interior STANDARD row the interval and tau of the continuous extension.
struct BcLoc {
    int node;
    int iv;       // -1 when the value is read straight from a node
    double tau;
};

template <int M>
__device__ inline BcLoc bc_locate(const Inst<M> &I, int k, const double *mesh, int N) {
    BcLoc L;
    const double t = sel<M>(I.bct, k);
    L.iv = -1;
    L.tau = 0.0;
    if (I.bc_kind == BVP_BC_TWO_POINT) {
        L.node = (t == I.t0) ? 0 : N;
        return L;
    }
    if (t == mesh[0]) { L.node = 0; return L; }
    if (t == mesh[N]) { L.node = N; return L; }
    const int iv = interval_of(mesh, N, t);
    L.node = iv;
    L.iv = iv;
    L.tau = (t - mesh[iv]) / (mesh[iv + 1] - mesh[iv]);
    return L;
}

// ------------------------------------------------------------------ LDS layout
// Nodes whose state (or intervals whose shat) fit the staging area: all of the
// initial meshes of the M = 2 workloads; for larger M the area is what the tree
// leaves (the LDS per wave is not raised for them).
template <int M> constexpr int kStageNodes = M <= 2 ? 1025 : (2 * WAVE - 1) * kRelFields<M> / M;
template <int M> struct Shared {
    union {
        struct {
            double rel[WAVE][kRelFields<M>];          // relation exchange for the tree
            double tree[(WAVE - 1) * kRelFields<M>];  // back-substitution records of the tree merges
        };
        // after the tree: the corrected state of the chunk back-substitution, or the
        // shat of the defect pass, staged so that they reach global memory as whole
        // lines (a lane's own stores are 8-16 bytes at a stride of its chunk length)
        double stage[kStageNodes<M> * M];
    };
    double lo[WAVE][M], hi[WAVE][M];  // corrections at chunk ends
    double bcres[M];
    int bcnode[M];
    double root[2 * M];
    int flag;
    int item;
};

template <int M>
__device__ inline void rel_to_lds(Shared<M> &sh, int lane, const Rel<M> &R) {
    int f = 0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) sh.rel[lane][f++] = R.E[i][c];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) sh.rel[lane][f++] = R.F[i][c];
#pragma unroll
    for (int i = 0; i < M; ++i) sh.rel[lane][f++] = R.g[i];
}
template <int M>
__device__ inline void rel_from_lds(const Shared<M> &sh, int lane, Rel<M> &R) {
    int f = 0;
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) R.E[i][c] = sh.rel[lane][f++];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int c = 0; c < M; ++c) R.F[i][c] = sh.rel[lane][f++];
#pragma unroll
    for (int i = 0; i < M; ++i) R.g[i] = sh.rel[lane][f++];
}

// ------------------------------------------------------------------ residual pieces
// BC residual of row k at state y (primal).  For an interior STANDARD row the
// stages of interval iv are recomputed here.
template <int M, int METHOD>
__device__ inline double bc_residual(const Inst<M> &I, int k, const BcLoc &L, const double *mesh, const double *y) {
    const int comp = sel<M>(I.bcc, k);
    const double val = sel<M>(I.bcv, k);
    if (L.iv < 0) return y[(size_t)L.node * M + comp] - val;
    using TB = Tab<METHOD>;
    const int i = L.iv;
    const double ti = mesh[i], h = mesh[i + 1] - mesh[i];
    double yi[M], yj[M];
#pragma unroll
    for (int c = 0; c < M; ++c) { yi[c] = y[(size_t)i * M + c]; yj[c] = y[(size_t)(i + 1) * M + c]; }
    double K[TB::s][M], KI[TB::sstar - TB::s][M];
    stages<M, METHOD, double>(I.fam, I.P, ti, h, yi, yj, K);
    interp_stages<M, METHOD>(I.fam, I.P, ti, h, yi, yj, K, KI);
    double w[TB::sstar], z[M];
    interp_weights<METHOD>(L.tau, w, nullptr);
    sum_stages<M, METHOD>(h, yi, K, KI, w, nullptr, z, nullptr);
    return sel<M>(z, comp) - val;
}

// ------------------------------------------------------------------ Newton step
// One forward pass at y: residual, blocks, chunk elimination.  Returns the
// lane's max |Phi| (NaN-aware via `bad`), its chunk relation.
template <int M, int METHOD, bool F32 = false, int FAM = -1>
__device__ inline void forward_dual(const Inst<M> &I, const double *mesh, const double *y, int N, int lane,
                                    double *scr, Rel<M> &rel, double &mx, bool &bad, double *phi_out,
                                    double *S_out, double *R_out) {
    int P, start, len;
    chunk_of(N, lane, P, start, len, I.minc);
    mx = 0.0;
    bad = false;
    walk_chunk<M, kPfDual<M>>(mesh, y, start, len, [&](int k, double ti, double tj, const double *yi, const double *yj) {
        const double h = tj - ti;
        double phi[M], S[M][M], R[M][M], Kp[Tab<METHOD>::s][M];
        if constexpr (M == 2 && (FAM == BVP_FAMILY_BRATU || FAM == BVP_FAMILY_SINGPERT))
            interval_chain<M, METHOD, FAM>(I.P, ti, h, yi, yj, phi, S, R, Kp);
        else
            interval_dual<M, METHOD>(I.fam, I.P, ti, h, yi, yj, phi, S, R, Kp);
        double r[M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            r[c] = -phi[c];
            const double a = fabs(phi[c]);
            bad |= !(a <= 1.7976931348623157e308);
            mx = fmax(mx, a);
        }
        if (phi_out) {
#pragma unroll
            for (int c = 0; c < M; ++c) phi_out[(size_t)k * M + c] = phi[c];
        }
        if (S_out) {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    S_out[(size_t)k * M * M + i * M + j] = S[i][j];
                    R_out[(size_t)k * M * M + i * M + j] = R[i][j];
                }
        }
        if (scr) {
            if (k == start) {
#pragma unroll
                for (int i = 0; i < M; ++i) {
#pragma unroll
                    for (int j = 0; j < M; ++j) { rel.E[i][j] = S[i][j]; rel.F[i][j] = R[i][j]; }
                    rel.g[i] = r[i];
                }
            } else {
                Rel<M> nr;
                Back<M> bk;
                eliminate<M>(rel, S, R, r, nr, bk);
                put_rec<M, F32>(scr, k - start - 1, lane, bk);
                rel = nr;
            }
        }
    });
}

// Tree up-sweep over the chunk relations (lane l holds the relation of its chunk in
// `rel`), root solve on lane 0 (the boundary rows, in sh.bcnode / sh.bcres, and the
// end-to-end relation), tree down-sweep.  Leaves in sh.lo[l] / sh.hi[l] the
// corrections at the first and last node of lane l's chunk.  Returns false (all
// lanes) when the root system is singular.  scr holds the chunk records (read by the
// walk to an interior boundary node), F32 their storage precision.  ROW2: irow holds,
// per boundary row, [2M + 2] doubles: a row linear in the corrections of two adjacent
// nodes (sh.bcnode and the next; coefficients, constant) where its last entry is
// nonzero, else the row reads its node; rowfn() writes irow after the up-sweep.
struct NoRows {
    __device__ void operator()() const {}
};
template <int M, bool F32, bool ROW2 = false, class RowFn = NoRows>
__device__ __forceinline__ bool tree_solve(const Inst<M> &I, Shared<M> &sh, int N, int lane, int P, double *scr,
                                           Rel<M> &rel, const double *irow = nullptr, RowFn rowfn = RowFn()) {
    int D = 0;
    while ((1 << D) < P) ++D;
    double *tree = sh.tree;  // tree records stay on chip (read back level by level in the down-sweep)
    // ---- up-sweep
    if (lane < P) rel_to_lds<M>(sh, lane, rel);
    __syncthreads();
    for (int d = 0; d < D; ++d) {
        const int span = 1 << d;
        if ((lane & ((span << 1) - 1)) == 0 && lane + span < P) {
            Rel<M> right, nr;
            rel_from_lds<M>(sh, lane + span, right);
            Back<M> bk;
            eliminate<M>(rel, right.E, right.F, right.g, nr, bk);
            store_back<M, 1>(tree, tree_index(d, lane), 0, bk);
            rel = nr;
            rel_to_lds<M>(sh, lane, rel);
        }
        __syncthreads();
    }
    if constexpr (ROW2) {
        rowfn();
        __syncthreads();
    }
    // ---- root: 2M x 2M system of the boundary rows and the end-to-end relation
    if (lane == 0) {
        constexpr int NN = 2 * M;
        double A[NN][NN], b[NN];
#pragma unroll
        for (int i = 0; i < NN; ++i) {
#pragma unroll
            for (int c = 0; c < NN; ++c) A[i][c] = 0.0;
            b[i] = 0.0;
        }
        // dy_node affinely in (dy_0, dy_N) (columns: dy_0 | dy_N | 1), by walking the
        // elimination records from the root to the node (identity rows at nodes 0 and N)
        constexpr int W = 2 * M + 1;
        auto walk = [&](int node, double (&mt)[M][W]) {
                double mlo[M][W], mhi[M][W];
#pragma unroll
                for (int i = 0; i < M; ++i)
#pragma unroll
                    for (int c = 0; c < W; ++c) {
                        mlo[i][c] = (c == i) ? 1.0 : 0.0;
                        mhi[i][c] = (c == M + i) ? 1.0 : 0.0;
                    }
                int l = 0;
                bool found = false;
                for (int d = D - 1; d >= 0 && !found; --d) {
                    const int span = 1 << d;
                    if (l + span >= P) continue;
                    int ps, plen, pP;
                    chunk_of(N, l + span, pP, ps, plen, I.minc);
                    Back<M> bk;
                    load_back<M, 1>(tree, tree_index(d, l), 0, bk);
                    double mb[M][W];
#pragma unroll
                    for (int i = 0; i < M; ++i)
#pragma unroll
                        for (int c = 0; c < W; ++c) {
                            double v = (c == 2 * M) ? bk.h[i] : 0.0;
#pragma unroll
                            for (int q = 0; q < M; ++q) v -= bk.X[i][q] * mlo[q][c] + bk.Z[i][q] * mhi[q][c];
                            mb[i][c] = v;
                        }
                    if (node == ps) {
#pragma unroll
                        for (int i = 0; i < M; ++i)
#pragma unroll
                            for (int c = 0; c < W; ++c) mt[i][c] = mb[i][c];
                        found = true;
                    } else if (node < ps) {
#pragma unroll
                        for (int i = 0; i < M; ++i)
#pragma unroll
                            for (int c = 0; c < W; ++c) mhi[i][c] = mb[i][c];
                    } else {
#pragma unroll
                        for (int i = 0; i < M; ++i)
#pragma unroll
                            for (int c = 0; c < W; ++c) mlo[i][c] = mb[i][c];
                        l += span;
                    }
                }
                if (!found) {
                    int ls, llen, lP;
                    chunk_of(N, l, lP, ls, llen, I.minc);
                    if (node == ls) {
#pragma unroll
                        for (int i = 0; i < M; ++i)
#pragma unroll
                            for (int c = 0; c < W; ++c) mt[i][c] = mlo[i][c];
                    } else {
                        // backwards through the chunk of lane l, from its end to `node`
                        double mn[M][W];
#pragma unroll
                        for (int i = 0; i < M; ++i)
#pragma unroll
                            for (int c = 0; c < W; ++c) mn[i][c] = mhi[i][c];
                        for (int kk = ls + llen - 1; kk >= node; --kk) {
                            Back<M> bk;
                            get_rec<M, F32>(scr, kk - ls - 1, l, bk);
                            double mk[M][W];
#pragma unroll
                            for (int i = 0; i < M; ++i)
#pragma unroll
                                for (int c = 0; c < W; ++c) {
                                    double v = (c == 2 * M) ? bk.h[i] : 0.0;
#pragma unroll
                                    for (int q = 0; q < M; ++q) v -= bk.X[i][q] * mlo[q][c] + bk.Z[i][q] * mn[q][c];
                                    mk[i][c] = v;
                                }
#pragma unroll
                            for (int i = 0; i < M; ++i)
#pragma unroll
                                for (int c = 0; c < W; ++c) mn[i][c] = mk[i][c];
                        }
#pragma unroll
                        for (int i = 0; i < M; ++i)
#pragma unroll
                            for (int c = 0; c < W; ++c) mt[i][c] = mn[i][c];
                    }
                }
        };
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const int node = sh.bcnode[k];
            const int comp = sel<M>(I.bcc, k);
            b[k] = -sh.bcres[k];
            if (ROW2 && irow[k * (W + 1) + W] != 0.0) {
                // a row linear in dy_node and dy_node+1: coefficients irow[k][0..M) and
                // [M..2M), constant irow[k][2M]
                const double *ir = irow + k * (W + 1);
                double ma[M][W], mb[M][W];
                walk(node, ma);
                walk(node + 1, mb);
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    double v = 0.0;
#pragma unroll
                    for (int i = 0; i < M; ++i) v += ir[i] * ma[i][c] + ir[M + i] * mb[i][c];
                    if (c < 2 * M) A[k][c] = v;
                    else b[k] -= v + ir[2 * M];
                }
            } else if (I.bc_kind == BVP_BC_TWO_POINT || node == 0 || node == N) {
#pragma unroll
                for (int c = 0; c < M; ++c) {
                    A[k][c] = (node == 0 && c == comp) ? 1.0 : 0.0;
                    A[k][M + c] = (node == N && c == comp) ? 1.0 : 0.0;
                }
            } else {
                double mt[M][W];
                walk(node, mt);
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    if (i == comp) {
#pragma unroll
                        for (int c = 0; c < 2 * M; ++c) A[k][c] = mt[i][c];
                        b[k] -= mt[i][2 * M];
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
#pragma unroll
            for (int c = 0; c < M; ++c) { A[M + i][c] = rel.E[i][c]; A[M + i][M + c] = rel.F[i][c]; }
            b[M + i] = rel.g[i];
        }
        const bool ok = dense_solve<NN>(A, b);
        sh.flag = ok ? 1 : 0;
#pragma unroll
        for (int c = 0; c < M; ++c) { sh.lo[0][c] = b[c]; sh.hi[0][c] = b[M + c]; }
    }
    __syncthreads();
    if (!sh.flag) return false;
    // ---- down-sweep
    for (int d = D - 1; d >= 0; --d) {
        const int span = 1 << d;
        if ((lane & ((span << 1) - 1)) == 0 && lane + span < P) {
            Back<M> bk;
            load_back<M, 1>(tree, tree_index(d, lane), 0, bk);
            double ya[M], yc[M], yb[M];
#pragma unroll
            for (int c = 0; c < M; ++c) { ya[c] = sh.lo[lane][c]; yc[c] = sh.hi[lane][c]; }
            back_apply<M>(bk, ya, yc, yb);
#pragma unroll
            for (int c = 0; c < M; ++c) {
                sh.lo[lane + span][c] = yb[c];
                sh.hi[lane + span][c] = yc[c];
                sh.hi[lane][c] = yb[c];
            }
        }
        __syncthreads();
    }
    return true;
}

// Tree up-sweep, root solve (lane 0), tree down-sweep and chunk back-substitution.
// Writes the correction into `delta` (point-major) when non-null and, when
// `apply` is set, adds it to y.
// FUSE (with apply and no delta): while the chunk is walked backwards the corrected
// values of both ends of each interval are in registers, so the residual of the new
// state is evaluated there -- *fn_out = max|F| of the corrected state (BC rows
// included; NaN if not finite), the convergence check of the next Newton iteration
// without another pass over HBM -- and with FUSE == 2 also the defect estimate,
// *dn_out = max_i |d_i| (MIRK/adaptivity.jl:416-461; NaN if not finite).
template <int M, int METHOD, bool F32 = false, int FUSE = 0>
__device__ inline bool solve_correction(const Inst<M> &I, Shared<M> &sh, const double *mesh, double *y, int N,
                                        int lane, double *scr, size_t tree_base, Rel<M> &rel, double *delta,
                                        bool apply, double *fn_out = nullptr, double *dn_out = nullptr) {
    int P, start, len;
    chunk_of(N, lane, P, start, len, I.minc);
    (void)tree_base;
    if (!tree_solve<M, F32>(I, sh, N, lane, P, scr, rel)) return false;
    // ---- chunk back-substitution
    // The corrected state goes to the LDS staging area (the tree records are no
    // longer needed) and then to y in whole lines, when it fits.
    const bool staged = apply && delta == nullptr && N + 1 <= kStageNodes<M>;
    double *ynew = staged ? sh.stage : y;
    using TB = Tab<METHOD>;
    double fmx = 0.0, dmx = 0.0;
    bool fbad = false, dbad = false;
    // residual (and defect) of interval [ti, tj] at the corrected state yi, yj
    auto check = [&](double ti, double tj, const double *yi, const double *yj) {
        const double h = tj - ti;
        double K[TB::s][M], ph[M];
        stages<M, METHOD, double>(I.fam, I.P, ti, h, yi, yj, K);
        phi_of<M, METHOD, double>(h, yi, yj, K, ph);
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const double a = fabs(ph[c]);
            fbad |= !(a <= 1.7976931348623157e308);
            fmx = fmax(fmx, a);
        }
        if constexpr (FUSE == 2) {
            double w[TB::sstar], wp[TB::sstar];
            double KI[TB::sstar - TB::s][M], z[M], zp[M], f[M];
            interp_stages<M, METHOD>(I.fam, I.P, ti, h, yi, yj, K, KI);
            double e1 = 0.0, e2 = 0.0, d1[M], d2[M];
            interp_weights<METHOD>(TB::tau(), w, wp);
            sum_stages<M, METHOD>(h, yi, K, KI, w, wp, z, zp);
            f_eval<M, double>(I.fam, I.P, ti + TB::tau() * h, z, f);
#pragma unroll
            for (int c = 0; c < M; ++c) { d1[c] = (zp[c] - f[c]) / (fabs(f[c]) + 1.0); e1 = fmax(e1, fabs(d1[c])); }
            interp_weights<METHOD>(1.0 - TB::tau(), w, wp);
            sum_stages<M, METHOD>(h, yi, K, KI, w, wp, z, zp);
            f_eval<M, double>(I.fam, I.P, ti + (1.0 - TB::tau()) * h, z, f);
#pragma unroll
            for (int c = 0; c < M; ++c) { d2[c] = (zp[c] - f[c]) / (fabs(f[c]) + 1.0); e2 = fmax(e2, fabs(d2[c])); }
            const bool first = e1 > e2;
            double ei = 0.0;
#pragma unroll
            for (int c = 0; c < M; ++c) {
                const double d = first ? d1[c] : d2[c];
                dbad |= !(fabs(d) <= 1.7976931348623157e308);
                ei = fmax(ei, fabs(d));
            }
            dmx = fmax(dmx, ei);
        }
    };
    if (len > 0) {
        double ya[M], yn[M];
#pragma unroll
        for (int c = 0; c < M; ++c) { ya[c] = sh.lo[lane][c]; yn[c] = sh.hi[lane][c]; }
        // FUSE: t and corrected state of the node right of the next interval to check
        // (the chunk's end node: read before any lane writes the state)
        double tr = 0.0, yr[M];
        if constexpr (FUSE > 0) {
            tr = mesh[start + len];
#pragma unroll
            for (int c = 0; c < M; ++c) yr[c] = y[(size_t)(start + len) * M + c] + yn[c];
        }
        if (lane == P - 1) {
#pragma unroll
            for (int c = 0; c < M; ++c) {
                if (delta) delta[(size_t)N * M + c] = yn[c];
                if (apply) ynew[(size_t)N * M + c] = y[(size_t)N * M + c] + yn[c];
            }
        }
        // nodes k = start+len-1 down to start+1; step i is node start+len-1-i, whose record
        // is slot len-2-i.  Records (and y, and with FUSE the node times) are read PB
        // steps ahead into a register ring.
        constexpr int PB = kPfBack<M>;
        const int ns = len - 1;
        Back<M> rb[PB];
        double ry[PB][M], rt[PB];
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            if (j < ns) {
                get_rec<M, F32>(scr, len - 2 - j, lane, rb[j]);
                if (apply) {
#pragma unroll
                    for (int c = 0; c < M; ++c) ry[j][c] = y[(size_t)(start + len - 1 - j) * M + c];
                }
                if constexpr (FUSE > 0) rt[j] = mesh[start + len - 1 - j];
            }
        }
        for (int i0 = 0; i0 < ns; i0 += PB) {
#pragma unroll
            for (int j = 0; j < PB; ++j) {
                const int i = i0 + j;
                if (i < ns) {
                    const int k = start + len - 1 - i;
                    const Back<M> bk = rb[j];
                    double ycur[M];
#pragma unroll
                    for (int c = 0; c < M; ++c) ycur[c] = ry[j][c];
                    const double tk = rt[j];
                    if (i + PB < ns) {
                        get_rec<M, F32>(scr, len - 2 - (i + PB), lane, rb[j]);
                        if (apply) {
#pragma unroll
                            for (int c = 0; c < M; ++c) ry[j][c] = y[(size_t)(k - PB) * M + c];
                        }
                        if constexpr (FUSE > 0) rt[j] = mesh[k - PB];
                    }
                    double yk[M], yl[M];
                    back_apply<M>(bk, ya, yn, yk);
#pragma unroll
                    for (int c = 0; c < M; ++c) {
                        yl[c] = ycur[c] + yk[c];
                        if (delta) delta[(size_t)k * M + c] = yk[c];
                        if (apply) ynew[(size_t)k * M + c] = yl[c];
                        yn[c] = yk[c];
                    }
                    if constexpr (FUSE > 0) {
                        check(tk, tr, yl, yr);
                        tr = tk;
#pragma unroll
                        for (int c = 0; c < M; ++c) yr[c] = yl[c];
                    }
                }
            }
        }
        double yl[M];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            yl[c] = y[(size_t)start * M + c] + ya[c];
            if (delta) delta[(size_t)start * M + c] = ya[c];
            if (apply) ynew[(size_t)start * M + c] = yl[c];
        }
        if constexpr (FUSE > 0) check(mesh[start], tr, yl, yr);
    }
    __syncthreads();
    if (staged) {
        for (int q = lane; q < (N + 1) * M; q += WAVE) y[q] = sh.stage[q];
        __syncthreads();
    }
    if constexpr (FUSE > 0) {
        if (lane < M) {
            const BcLoc L = bc_locate<M>(I, lane, mesh, N);
            const double a = fabs(bc_residual<M, METHOD>(I, lane, L, mesh, y));
            fbad |= !(a <= 1.7976931348623157e308);
            fmx = fmax(fmx, a);
        }
        const bool fb = wave_any(fbad);
        fmx = uniform(wave_max(fmx));
        *fn_out = fb ? NAN : fmx;
        if constexpr (FUSE == 2) {
            const bool db = wave_any(dbad);
            dmx = uniform(wave_max(dmx));
            *dn_out = db ? NAN : dmx;
        }
        __syncthreads();
    }
    return true;
}

// Residual evaluation: forward pass + BC rows.  Leaves BC residuals and their
// nodes in LDS; returns max|F| (NaN if not finite) in all lanes.
template <int M, int METHOD, bool F32 = false, int FAM = -1>
__device__ inline double residual_and_blocks(const Inst<M> &I, Shared<M> &sh, const double *mesh, const double *y,
                                             int N, int lane, double *scr, Rel<M> &rel) {
    double mx;
    bool bad;
    forward_dual<M, METHOD, F32, FAM>(I, mesh, y, N, lane, scr, rel, mx, bad, nullptr, nullptr, nullptr);
    if (lane < M) {
        const BcLoc L = bc_locate<M>(I, lane, mesh, N);
        sh.bcnode[lane] = L.node;
        sh.bcres[lane] = bc_residual<M, METHOD>(I, lane, L, mesh, y);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double a = fabs(sh.bcres[k]);
        bad |= !(a <= 1.7976931348623157e308);
        mx = fmax(mx, a);
    }
    const bool anybad = wave_any(bad);
    mx = wave_max(mx);
    __syncthreads();
    return anybad ? NAN : mx;
}

// Residual only (primal stages, no Jacobian blocks, no elimination): max|F| (NaN
// if not finite) in all lanes.  Newton's convergence check after a step uses it,
// so an iterate that has converged does not pay for the dual-number pass.
template <int M, int METHOD>
__device__ inline double residual_primal(const Inst<M> &I, const double *mesh, const double *y, int N, int lane) {
    int P, start, len;
    chunk_of(N, lane, P, start, len, I.minc);
    double mx = 0.0;
    bool bad = false;
    walk_chunk<M, kPfPrimal<M>>(mesh, y, start, len, [&](int, double ti, double tj, const double *yi, const double *yj) {
        const double h = tj - ti;
        double K[Tab<METHOD>::s][M], ph[M];
        stages<M, METHOD, double>(I.fam, I.P, ti, h, yi, yj, K);
        phi_of<M, METHOD, double>(h, yi, yj, K, ph);
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const double a = fabs(ph[c]);
            bad |= !(a <= 1.7976931348623157e308);
            mx = fmax(mx, a);
        }
    });
    if (lane < M) {
        const BcLoc L = bc_locate<M>(I, lane, mesh, N);
        const double a = fabs(bc_residual<M, METHOD>(I, lane, L, mesh, y));
        bad |= !(a <= 1.7976931348623157e308);
        mx = fmax(mx, a);
    }
    const bool anybad = wave_any(bad);
    mx = wave_max(mx);
    __syncthreads();
    return anybad ? NAN : mx;
}

// Primal residual statistics at state y: returns sum of squares of all residual
// rows (Phi and BC rows); *mx receives max|F|.  Both NaN if a value is not finite.
template <int M, int METHOD>
__device__ inline double residual_stats(const Inst<M> &I, const double *mesh, const double *y, int N, int lane,
                                        double *mx_out) {
    int P, start, len;
    chunk_of(N, lane, P, start, len, I.minc);
    double acc = 0.0, mx = 0.0;
    walk_chunk<M, kPfPrimal<M>>(mesh, y, start, len, [&](int, double ti, double tj, const double *yi, const double *yj) {
        const double h = tj - ti;
        double K[Tab<METHOD>::s][M], ph[M];
        stages<M, METHOD, double>(I.fam, I.P, ti, h, yi, yj, K);
        phi_of<M, METHOD, double>(h, yi, yj, K, ph);
#pragma unroll
        for (int c = 0; c < M; ++c) {
            acc += ph[c] * ph[c];
            mx = fmax(mx, fabs(ph[c]));
        }
    });
    if (lane < M) {
        const BcLoc L = bc_locate<M>(I, lane, mesh, N);
        const double r = bc_residual<M, METHOD>(I, lane, L, mesh, y);
        acc += r * r;
        mx = fmax(mx, fabs(r));
    }
    const bool bad = wave_any(!(acc <= 1.7976931348623157e308));
    acc = uniform(wave_sum(acc));
    mx = uniform(wave_max(mx));
    __syncthreads();
    if (mx_out) *mx_out = bad ? NAN : mx;
    return bad ? NAN : acc;
}

// ytmp = y + a*d1 + b*d2 (d2 may be null), then the residual statistics there.
template <int M, int METHOD>
__device__ inline double residual_sq_trial(const Inst<M> &I, const double *mesh, const double *y, const double *d1,
                                           double a, const double *d2, double b, int N, int lane, double *ytmp,
                                           double *mx_out = nullptr) {
    for (int q = lane; q < (N + 1) * M; q += WAVE) {
        double v = y[q] + a * d1[q];
        if (d2) v += b * d2[q];
        ytmp[q] = v;
    }
    __syncthreads();
    return residual_stats<M, METHOD>(I, mesh, ytmp, N, lane, mx_out);
}

// Phi_k and the Jacobian blocks of one interval (dual numbers, or the chain rule
// in the kernels specialised to one family).
template <int M, int METHOD, int FAM>
__device__ inline void interval_blocks(const Inst<M> &I, double ti, double h, const double *yi, const double *yj,
                                       double (&phi)[M], double (&S)[M][M], double (&R)[M][M]) {
    double Kp[Tab<METHOD>::s][M];
    if constexpr (M == 2 && (FAM == BVP_FAMILY_BRATU || FAM == BVP_FAMILY_SINGPERT))
        interval_chain<M, METHOD, FAM>(I.P, ti, h, yi, yj, phi, S, R, Kp);
    else
        interval_dual<M, METHOD>(I.fam, I.P, ti, h, yi, yj, phi, S, R, Kp);
}

// g = J' F at y (point-major, M*(N+1)); the BC rows' residuals and nodes are the
// ones residual_and_blocks left in LDS (frozen-stage unit rows).
template <int M, int METHOD, int FAM>
__device__ inline void jt_times_f(const Inst<M> &I, Shared<M> &sh, const double *mesh, const double *y, int N,
                                  int lane, double *g) {
    int P, start, len;
    chunk_of(N, lane, P, start, len, I.minc);
    double carry[M];
#pragma unroll
    for (int c = 0; c < M; ++c) carry[c] = 0.0;
    walk_chunk<M, kPfDual<M>>(mesh, y, start, len, [&](int k, double ti, double tj, const double *yi, const double *yj) {
        double phi[M], S[M][M], R[M][M];
        interval_blocks<M, METHOD, FAM>(I, ti, tj - ti, yi, yj, phi, S, R);
#pragma unroll
        for (int b = 0; b < M; ++b) {
            double gs = carry[b], gr = 0.0;
#pragma unroll
            for (int a = 0; a < M; ++a) {
                gs += S[a][b] * phi[a];
                gr += R[a][b] * phi[a];
            }
            g[(size_t)k * M + b] = gs;
            carry[b] = gr;
        }
    });
    // the carry of a chunk belongs to the first node of the next one
    if (len > 0) {
#pragma unroll
        for (int c = 0; c < M; ++c) sh.lo[lane][c] = carry[c];
    }
    __syncthreads();
    if (len > 0 && lane > 0) {
#pragma unroll
        for (int c = 0; c < M; ++c) g[(size_t)start * M + c] += sh.lo[lane - 1][c];
    }
    if (lane == P - 1) {
#pragma unroll
        for (int c = 0; c < M; ++c) g[(size_t)N * M + c] = carry[c];
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < M; ++k) g[(size_t)sh.bcnode[k] * M + sel<M>(I.bcc, k)] += sh.bcres[k];
    }
    __syncthreads();
}

// |J v|^2 at y (Phi rows and the frozen-stage BC rows); all lanes.
template <int M, int METHOD, int FAM>
__device__ inline double jac_norm2(const Inst<M> &I, const Shared<M> &sh, const double *mesh, const double *y,
                                   const double *v, int N, int lane) {
    int P, start, len;
    chunk_of(N, lane, P, start, len, I.minc);
    double acc = 0.0;
    double vi[M];
    if (len > 0) {
#pragma unroll
        for (int c = 0; c < M; ++c) vi[c] = v[(size_t)start * M + c];
    }
    walk_chunk<M, kPfDual<M>>(mesh, y, start, len, [&](int k, double ti, double tj, const double *yi, const double *yj) {
        double phi[M], S[M][M], R[M][M], vj[M];
        interval_blocks<M, METHOD, FAM>(I, ti, tj - ti, yi, yj, phi, S, R);
#pragma unroll
        for (int c = 0; c < M; ++c) vj[c] = v[(size_t)(k + 1) * M + c];
#pragma unroll
        for (int a = 0; a < M; ++a) {
            double r = 0.0;
#pragma unroll
            for (int b = 0; b < M; ++b) r += S[a][b] * vi[b] + R[a][b] * vj[b];
            acc += r * r;
        }
#pragma unroll
        for (int c = 0; c < M; ++c) vi[c] = vj[c];
    });
    if (lane < M) {
        const double r = v[(size_t)sh.bcnode[lane] * M + sel<M>(I.bcc, lane)];
        acc += r * r;
    }
    acc = uniform(wave_sum(acc));
    __syncthreads();
    return acc;
}

// What the plain Newton step evaluates at the corrected state inside its
// back-substitution walk: 1 the residual max|F|, 2 also the defect -- the defect
// only when the step is predicted to converge: from max|F| = fn, quadratic
// convergence C fn^2 <= tol with C taken as 1e-4 (fn <= 100 sqrt(tol); on the
// Bratu family at tol 1e-8 nine in ten first steps from the learned guess start above
// sqrt(tol) and still converge).  A step that converges without the defect is
// followed by a separate defect pass, which computes the same values.
constexpr int kFusePlain = 2;

struct NewtonOut {
    int ret, iters, stage, warm_fallback;
    double fn;  // max|F| at the state Newton returns
    double dn;  // defect estimate of that state when dn_valid (evaluated with its last correction)
    int dn_valid;
};

// LineSearches.jl's BackTracking, order 3 (the restatement the oracle follows):
// phi(a) = |F(y + a d)|^2 / 2, Armijo constant 1e-4, quadratic then cubic
// interpolation, each new step kept within [0.1, 0.5] of the previous one, a
// non-finite trial halved first (at most 52 times), at most 1000 reductions.
// phi returns the same value in every lane, so every branch here is uniform.
// Returns false when the reductions ran out.
template <typename Phi>
__device__ inline bool backtracking(Phi &&phi, double phi0, double dphi0, double &alpha) {
    const double c1 = 1e-4, rho_hi = 0.5, rho_lo = 0.1;
    double a1 = 1.0, a2 = 1.0;
    double fx0 = phi0, fx1 = phi(a1);
    for (int f = 0; !(fabs(fx1) <= 1.7976931348623157e308) && f < 52; ++f) {
        a1 = a2;
        a2 = a1 * 0.5;
        fx1 = phi(a2);
    }
    for (int iteration = 1; fx1 > phi0 + c1 * a2 * dphi0; ++iteration) {
        if (iteration > 1000) { alpha = a2; return false; }
        double at;
        if (iteration == 1) {
            at = -(dphi0 * a2 * a2) / (2.0 * (fx1 - phi0 - dphi0 * a2));
        } else {
            const double div = 1.0 / (a1 * a1 * a2 * a2 * (a2 - a1));
            const double e1 = fx1 - phi0 - dphi0 * a2, e0 = fx0 - phi0 - dphi0 * a1;
            const double a = (a1 * a1 * e1 - a2 * a2 * e0) * div;
            const double b = (-a1 * a1 * a1 * e1 + a2 * a2 * a2 * e0) * div;
            if (fabs(a) <= 2.220446049250313e-16) {
                at = dphi0 / (2.0 * b);
            } else {
                const double d = fmax(b * b - 3.0 * a * dphi0, 0.0);
                at = (-b + sqrt(d)) / (3.0 * a);
            }
        }
        a1 = a2;
        const double hi = a2 * rho_hi, lo = a2 * rho_lo;
        at = isnan(at) ? hi : fmin(at, hi);
        a2 = isnan(at) ? lo : fmax(at, lo);
        fx0 = fx1;
        fx1 = phi(a2);
    }
    alpha = a2;
    return true;
}

// The TrustRegion stage (dogleg step, "Simple" radius update; the restatement the
// oracle follows): from y, with buffers dn (Newton step), g (J'F) and ytmp (trial).
template <int M, int METHOD, bool F32, int FAM>
__device__ inline int trust_region(const BatchDev &B, const Inst<M> &I, Shared<M> &sh, const double *mesh, double *y,
                                   double *dn, double *g, double *ytmp, int N, int lane, double *scr,
                                   size_t tree_base, NewtonOut &out) {
    double mx;
    double ss = residual_stats<M, METHOD>(I, mesh, y, N, lane, &mx);
    double ymax = -INFINITY, ymin = INFINITY;
    for (int q = lane; q < (N + 1) * M; q += WAVE) {
        ymax = fmax(ymax, y[q]);
        ymin = fmin(ymin, y[q]);
    }
    ymax = uniform(wave_max(ymax));
    ymin = uniform(-wave_max(-ymin));
    __syncthreads();
    double radius = fmax(sqrt(ss), ymax - ymin) / 11.0;
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
            residual_and_blocks<M, METHOD, F32, FAM>(I, sh, mesh, y, N, lane, scr, rel);
            if (!solve_correction<M, METHOD, F32>(I, sh, mesh, y, N, lane, scr, tree_base, rel, dn, false))
                return BVP_RET_FAILURE;
            jt_times_f<M, METHOD, FAM>(I, sh, mesh, y, N, lane, g);
            double a = 0.0, b = 0.0, c = 0.0;
            for (int q = lane; q < (N + 1) * M; q += WAVE) {
                a += dn[q] * dn[q];
                b += g[q] * g[q];
                c += g[q] * dn[q];
            }
            dn2 = uniform(wave_sum(a));
            l2 = uniform(wave_sum(b));
            gdn = uniform(wave_sum(c));
            __syncthreads();
            jg2 = jac_norm2<M, METHOD, FAM>(I, sh, mesh, y, g, N, lane);
            fresh = false;
        }
        out.iters++;
        // step d = cn * dn + cg * g and the predicted reduction of the linear model
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
                // |a|^2, 2 dc.a for a = dn - dc, dc = -beta g
                const double A2 = dn2 + 2.0 * beta * gdn + beta * beta * l2;
                const double B2 = -2.0 * (beta * gdn + beta * beta * l2);
                const double C2 = dcauchy * dcauchy - radius * radius;
                const double tau = (-B2 + sqrt(fmax(0.0, B2 * B2 - 4.0 * A2 * C2))) / (2.0 * A2);
                cn = tau; cg = -(1.0 - tau) * beta;
                pred = ss - (1.0 - tau) * (1.0 - tau) * (ss - l2 * l2 / jg2);
            }
        }
        double mxt;
        const double sst = residual_sq_trial<M, METHOD>(I, mesh, y, dn, cn, g, cg, N, lane, ytmp, &mxt);
        const double rho = (ss - sst) / pred;
        if (rho > 1e-4) {
            for (int q = lane; q < (N + 1) * M; q += WAVE) y[q] = ytmp[q];
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

// The default polyalgorithm NewtonRaphson -> NewtonRaphson + BackTracking ->
// TrustRegion (Core/default_internal_solve.jl:37-43), stages selected by the bit
// mask B.stages, each restarting from the same start; stop when max|F| <= tol,
// checked before each step.
// restore() writes the starting state into y again; it regenerates the state from
// its sources instead of keeping a copy of it.
template <int M, int METHOD, bool F32 = false, int FAM = -1, typename Restore>
__device__ __forceinline__ NewtonOut newton(const BatchDev &B, const Inst<M> &I, Shared<M> &sh, const double *mesh, double *y,
                            Restore &&restore, double *delta, double *ytmp, double *gbuf, int N, int lane, double *scr,
                            size_t tree_base, const double *u0c) {
    NewtonOut out{BVP_RET_MAXITERS, 0, -1, 0, NAN, NAN, 0};
    // u0c != null: y holds a learned warm start; when the first Newton step from it
    // does not contract max|F| (or the guess is unusable), restart from the constant
    // guess u0c, as the plain method would have started.
    bool warm = u0c != nullptr;
    double fn_warm = 0.0;
    bool start_u0c = false;  // the first stage fell back to u0c: later stages start there too
    bool first = true;
    for (int stage = 0; stage < 3; ++stage) {
        if (!((B.stages >> stage) & 1)) continue;
        if (!first) {
            if (start_u0c) {
                for (int k = lane; k <= N; k += WAVE)
#pragma unroll
                    for (int c = 0; c < M; ++c) y[(size_t)k * M + c] = u0c[c];
            } else {
                restore();
            }
            __syncthreads();
        }
        first = false;
        out.stage = stage;
        int ret = BVP_RET_MAXITERS;
        out.dn_valid = 0;
        if (stage == 2) {
            ret = trust_region<M, METHOD, F32, FAM>(B, I, sh, mesh, y, delta, gbuf, ytmp, N, lane, scr, tree_base, out);
        } else {
            // after a plain step max|F| (and the defect) of the new state come from its
            // back-substitution; after a line-search step from a primal pass.  The dual
            // pass that builds the blocks runs only when another step is taken.
            double fn_next = NAN, dn_next = NAN;
            bool dn_fused = false;  // the last plain step evaluated the defect of its result
            const double fuse_defect_below = 100.0 * sqrt(B.newton_tol);
            for (int it = 0;; ++it) {
                Rel<M> rel;
                bool blocks = (it == 0);
                double fn = blocks ? residual_and_blocks<M, METHOD, F32, FAM>(I, sh, mesh, y, N, lane, scr, rel)
                                   : (stage == 0 ? fn_next : residual_primal<M, METHOD>(I, mesh, y, N, lane));
                out.dn_valid = (!blocks && stage == 0 && dn_fused) ? 1 : 0;
                out.dn = dn_next;
                if (warm) {
                    bool fall;
                    if (it == 0) {
                        fn_warm = fn;
                        fall = !(fn <= 1.7976931348623157e308);
                    } else {
                        fall = !(fn < fn_warm);
                        warm = false;
                    }
                    if (fall) {
                        warm = false;
                        out.warm_fallback = 1;
                        for (int k = lane; k <= N; k += WAVE)
#pragma unroll
                            for (int c = 0; c < M; ++c) y[(size_t)k * M + c] = u0c[c];
                        start_u0c = true;
                        __syncthreads();
                        fn = residual_and_blocks<M, METHOD, F32, FAM>(I, sh, mesh, y, N, lane, scr, rel);
                        blocks = true;
                        out.dn_valid = 0;
                        it = 0;  // the restart has the full iteration budget, as the plain method
                    }
                }
                out.fn = fn;  // every exit below leaves y at the state fn was computed for
                if (!(fn <= 1.7976931348623157e308)) { ret = BVP_RET_UNSTABLE; break; }
                if (fn <= B.newton_tol) { ret = BVP_RET_SUCCESS; break; }
                if (it >= B.max_newton) { ret = BVP_RET_MAXITERS; break; }
                if (!blocks) residual_and_blocks<M, METHOD, F32, FAM>(I, sh, mesh, y, N, lane, scr, rel);
                const bool plain = (stage == 0);
                dn_fused = plain && kFusePlain == 2 && fn <= fuse_defect_below;
                dn_next = NAN;
                const bool solved =
                    !plain ? solve_correction<M, METHOD, F32>(I, sh, mesh, y, N, lane, scr, tree_base, rel, delta, false)
                    : dn_fused ? solve_correction<M, METHOD, F32, 2>(I, sh, mesh, y, N, lane, scr, tree_base, rel,
                                                                     nullptr, true, &fn_next, &dn_next)
                               : solve_correction<M, METHOD, F32, 1>(I, sh, mesh, y, N, lane, scr, tree_base, rel,
                                                                     nullptr, true, &fn_next, nullptr);
                if (!solved) {
                    if (warm) {  // singular at the warm start: restart from the constant guess
                        warm = false;
                        out.warm_fallback = 1;
                        out.iters++;
                        for (int k = lane; k <= N; k += WAVE)
#pragma unroll
                            for (int c = 0; c < M; ++c) y[(size_t)k * M + c] = u0c[c];
                        start_u0c = true;
                        __syncthreads();
                        it = -1;
                        continue;
                    }
                    ret = BVP_RET_FAILURE;
                    break;
                }
                out.iters++;
                if (!plain) {
                    // phi(0) with the same primal code as the trials; phi'(0) = -|F|^2
                    // for the Newton direction
                    const double ss = residual_stats<M, METHOD>(I, mesh, y, N, lane, nullptr);
                    double alpha = 1.0;
                    auto phi = [&](double a) {
                        return 0.5 * residual_sq_trial<M, METHOD>(I, mesh, y, delta, a, nullptr, 0.0, N, lane, ytmp);
                    };
                    if (!backtracking(phi, 0.5 * ss, -ss, alpha)) { ret = BVP_RET_FAILURE; break; }
                    for (int q = lane; q < (N + 1) * M; q += WAVE) y[q] += alpha * delta[q];
                    __syncthreads();
                }
            }
        }
        out.ret = ret;
        if (ret == BVP_RET_SUCCESS) break;
    }
    return out;
}

// max_i |d_i| / max(1, |y_i|) for the Newton correction d at (mesh, y): the
// first-order distance of y from the discrete solution on this mesh (FP64
// residual and elimination).  NaN if the matrix is singular.
template <int M, int METHOD, bool F32, int FAM>
__device__ inline double correction_norm(const Inst<M> &I, Shared<M> &sh, const double *mesh, double *y, double *d,
                                         int N, int lane, double *scr, size_t tree_base) {
    Rel<M> rel;
    residual_and_blocks<M, METHOD, F32, FAM>(I, sh, mesh, y, N, lane, scr, rel);
    if (!solve_correction<M, METHOD, F32>(I, sh, mesh, y, N, lane, scr, tree_base, rel, d, false)) return NAN;
    double e = 0.0;
    for (int q = lane; q < (N + 1) * M; q += WAVE) e = fmax(e, fabs(d[q]) / fmax(1.0, fabs(y[q])));
    e = wave_max(e);
    __syncthreads();
    return e;
}

// ------------------------------------------------------------------ defect
// error_estimate!(::DefectControl): per interval the defect vector with the
// larger max-abs of the two sample points; returns defect norm (NaN-aware) and
// fills shat (= (max|d_i|/abstol)^(1/(p+1)), MIRK/adaptivity.jl:39-40).
// forced inline: the specialised kernels fold the family into it
//
// stage != null: shat is written through that LDS area (N values) and copied to
// shat in whole lines at the end.  lazy: shat, r1 and r2 are formed only when the
// defect norm exceeds abstol (only then does mesh selection read them; r1 = r2 =
// 0 otherwise), from the interval errors kept meanwhile where shat goes.
template <int M, int METHOD>
__device__ __forceinline__ double defect_pass(const Inst<M> &I, const double *mesh, const double *y, int N, int lane,
                                     double abstol, double *shat, double *errors, double &r1, double &r2,
                                     double *stage = nullptr, bool lazy = false) {
    using TB = Tab<METHOD>;
    int P, start, len;
    chunk_of(N, lane, P, start, len, I.minc);
    double w1[TB::sstar], w1p[TB::sstar], w2[TB::sstar], w2p[TB::sstar];
    interp_weights<METHOD>(TB::tau(), w1, w1p);
    interp_weights<METHOD>(1.0 - TB::tau(), w2, w2p);
    double dmax = 0.0, smax = 0.0, ssum = 0.0;
    bool bad = false;
    const double ex = 1.0 / (TB::order + 1);
    walk_chunk<M, kPfDefect<M>>(mesh, y, start, len, [&](int k, double ti, double tj, const double *yi, const double *yj) {
        const double h = tj - ti;
        double K[TB::s][M], KI[TB::sstar - TB::s][M];
        stages<M, METHOD, double>(I.fam, I.P, ti, h, yi, yj, K);
        interp_stages<M, METHOD>(I.fam, I.P, ti, h, yi, yj, K, KI);
        double z[M], zp[M], f1[M], f2[M], d1[M], d2[M];
        sum_stages<M, METHOD>(h, yi, K, KI, w1, w1p, z, zp);
        f_eval<M, double>(I.fam, I.P, ti + TB::tau() * h, z, f1);
        double e1 = 0.0, e2 = 0.0;
#pragma unroll
        for (int c = 0; c < M; ++c) { d1[c] = (zp[c] - f1[c]) / (fabs(f1[c]) + 1.0); e1 = fmax(e1, fabs(d1[c])); }
        sum_stages<M, METHOD>(h, yi, K, KI, w2, w2p, z, zp);
        f_eval<M, double>(I.fam, I.P, ti + (1.0 - TB::tau()) * h, z, f2);
#pragma unroll
        for (int c = 0; c < M; ++c) { d2[c] = (zp[c] - f2[c]) / (fabs(f2[c]) + 1.0); e2 = fmax(e2, fabs(d2[c])); }
        const bool first = e1 > e2;
        double ei = 0.0;
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const double d = first ? d1[c] : d2[c];
            if (errors) errors[(size_t)k * M + c] = d;
            bad |= !(fabs(d) <= 1.7976931348623157e308);
            ei = fmax(ei, fabs(d));
        }
        dmax = fmax(dmax, ei);
        if (lazy) {
            (stage ? stage : shat)[k] = ei;
        } else {
            const double sh = pow(ei / abstol, ex);
            if (shat) (stage ? stage : shat)[k] = sh;
            smax = fmax(smax, sh);
            ssum += sh;
        }
    });
    const bool anybad = wave_any(bad);
    dmax = wave_max(dmax);
    bool need = shat != nullptr;
    if (lazy) {
        need = !anybad && dmax > abstol;
        if (need) {  // shat from the kept errors, in the same order as above
            double *e = stage ? stage : shat;
            for (int k = start; k < start + len; ++k) {
                const double sh = pow(e[k] / abstol, ex);
                e[k] = sh;
                smax = fmax(smax, sh);
                ssum += sh;
            }
        }
    }
    if (need && stage) {
        __syncthreads();
        for (int k = lane; k < N; k += WAVE) shat[k] = stage[k];
        __syncthreads();
    }
    r1 = wave_max(smax);
    r2 = wave_sum(ssum);
    return anybad ? NAN : dmax;
}

// ------------------------------------------------------------------ mesh selection
// Julia round(Int, x): nearest, ties to even.
__device__ inline long long round_even(double x) { return (long long)rint(x); }

// half_mesh! (MIRK/adaptivity.jl:287-304) into `out`
__device__ inline void half_mesh_into(const double *mesh, int N, double *out, int lane) {
    for (int i = lane; i <= N; i += WAVE) out[2 * i] = mesh[i];
    for (int i = lane; i < N; i += WAVE) out[2 * i + 1] = (mesh[i + 1] + mesh[i]) / 2.0;
}

// mesh_selector! + redistribute! (MIRK/adaptivity.jl:23-243, 250-278).  shat
// holds the controller's indicator per interval on entry ((max|d|/abstol)^(1/(p+1))
// for DefectControl; it is divided by h in place), r1 and r2 its max and sum; rho
// the halving threshold (1 for DefectControl, 2 for the global-error controllers).
// Writes the new mesh into new_mesh and returns the new N, or -1 on failure (cap
// exceeded).
template <typename SH>
__device__ inline int mesh_select(const double *mesh, int N, double *shat, double r1, double r2, int max_sub,
                                  double *new_mesh, int lane, SH &sh, double rho = 1.0) {
    const int n = N;
    const double r3 = r2 / (double)n;
    long long npred = round_even(1.3 * r2 + 1.0);
    const double n_ = 0.1 * n;
    if (fabs((double)(npred - n)) < n_) npred = round_even((double)n + n_);
    if (r1 <= rho * r3) {
        const long long Ns = 2LL * n;
        if (Ns > max_sub) return -1;
        half_mesh_into(mesh, N, new_mesh, lane);
        __syncthreads();
        return (int)Ns;
    }
    const long long lb = (n + 1) / 2, ub = 4LL * n;
    const long long Ns = npred < lb ? lb : (npred > ub ? ub : npred);
    if (Ns > max_sub) return -1;
    // s_hat ./= h, then zeta = sum(s_hat .* h) / Ns
    int P, start, len;
    chunk_of(N, lane, P, start, len);
    double acc = 0.0;
    for (int k = start; k < start + len; ++k) {
        const double h = mesh[k + 1] - mesh[k];
        const double s = shat[k] / h;
        shat[k] = s;
        acc += s * h;
    }
    const double zeta = wave_sum(acc) / (double)Ns;
    __syncthreads();
    if (lane == 0) {
        // sequential equidistribution walk (MIRK/adaptivity.jl:254-273)
        new_mesh[0] = mesh[0];
        double t = mesh[0], integral = 0.0;
        int k = 0;
        long long i = 0;
        int ok = 1;
        while (k < n) {
            const double next_piece = shat[k] * (mesh[k + 1] - t);
            const double int_next = integral + next_piece;
            if (int_next > zeta) {
                const double tn = (zeta - integral) / shat[k] + t;
                if (i + 1 > Ns) { ok = 0; break; }
                new_mesh[i + 1] = tn;
                t = tn;
                ++i;
                integral = 0.0;
            } else {
                integral = int_next;
                t = mesh[k + 1];
                ++k;
            }
        }
        new_mesh[Ns] = mesh[n];
        sh.flag = ok;
    }
    __syncthreads();
    return sh.flag ? (int)Ns : -1;
}

// Solution transfer from the old continuous extension to the new nodes
// (MIRK/mirk.jl:366-370; MIRK/adaptivity.jl:6-13, 606-621).  REFERENCE mode
// reproduces the reference's in-place update: the base value of node j is the
// already-written new value when the old interval index ii < j.
template <int M, int METHOD>
__device__ inline void transfer(const Inst<M> &I, const double *mesh, const double *y, int N, const double *nm,
                                double *ny, int Nn, int mode, int lane, double *iiscr) {
    using TB = Tab<METHOD>;
    for (int j = lane; j <= Nn; j += WAVE) {
        const double m = nm[j];
        const int ii = interval_of(mesh, N, m);
        const double ti = mesh[ii], dt = mesh[ii + 1] - ti;
        const double tau = (m - ti) / dt;
        double yi[M], yj[M];
#pragma unroll
        for (int c = 0; c < M; ++c) { yi[c] = y[(size_t)ii * M + c]; yj[c] = y[(size_t)(ii + 1) * M + c]; }
        double K[TB::s][M], KI[TB::sstar - TB::s][M];
        stages<M, METHOD, double>(I.fam, I.P, ti, dt, yi, yj, K);
        interp_stages<M, METHOD>(I.fam, I.P, ti, dt, yi, yj, K, KI);
        double w[TB::sstar];
        interp_weights<METHOD>(tau, w, nullptr);
        if (mode == BVP_TRANSFER_CLEAN || ii >= j) {
            double z[M];
            sum_stages<M, METHOD>(dt, yi, K, KI, w, nullptr, z, nullptr);
#pragma unroll
            for (int c = 0; c < M; ++c) ny[(size_t)j * M + c] = z[c];
            iiscr[j] = -1.0;
        } else {
            double zero[M], z[M];
#pragma unroll
            for (int c = 0; c < M; ++c) zero[c] = 0.0;
            sum_stages<M, METHOD>(dt, zero, K, KI, w, nullptr, z, nullptr);
#pragma unroll
            for (int c = 0; c < M; ++c) ny[(size_t)j * M + c] = z[c];  // increment only
            iiscr[j] = (double)ii;
        }
    }
    __syncthreads();
    if (mode == BVP_TRANSFER_REFERENCE && lane == 0) {
        for (int j = 0; j <= Nn; ++j) {
            const double iid = iiscr[j];
            if (iid >= 0.0) {
                const int ii = (int)iid;
#pragma unroll
                for (int c = 0; c < M; ++c) ny[(size_t)j * M + c] += ny[(size_t)ii * M + c];
            }
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------ global error
// Julia's max: NaN if either argument is NaN
__device__ inline double jmax(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : fmax(a, b); }

// error_estimate!(::GlobalErrorControl) (MIRK/adaptivity.jl:511-538) with
// global_error + GE_subinterval! (:540-556) at (mesh, y).  RE: the same method,
// non-adaptive, on the halved mesh from halve_sol (:311-333) of y; HO (MIRK4 only):
// MIRK6 on the same mesh from y.  The nonlinear solve is the same polyalgorithm with
// the same tolerance, its return code is not consulted.  Per node e = (high - low)
// / (1 + |low|); per interval the end-node vector with the larger max-abs (the
// left one on ties; the right one when a max-abs is NaN), whose max-abs goes to
// ind[i] and, when errs is set, whose components go to errs[i][*].  Returns the
// max over the intervals (NaN-propagating), RE: times 2^p / (2^p - 1).
// ge: the slot's global-error buffers (wave_ge_doubles).
template <int M, int METHOD>
__device__ double global_error(const BatchDev &B, const Inst<M> &I, Shared<M> &sh, const double *mesh,
                               const double *y, int N, int lane, double *scr, size_t tree_base, double *ge,
                               double *ind, double *errs, long long &cnt_it) {
    const size_t c2 = (size_t)2 * B.cap;
    double *gm = ge, *gy = ge + c2, *dl = gy + c2 * M, *yt = dl + c2 * M, *gb = yt + c2 * M;
    int hs = 1;  // node i of the current mesh is node hs * i of the high-order solution
    if (B.global_method == BVP_GLOBAL_RE) {
        half_mesh_into(mesh, N, gm, lane);
        auto fill = [&]() {
            for (int i = lane; i <= N; i += WAVE)
#pragma unroll
                for (int c = 0; c < M; ++c) gy[(size_t)2 * i * M + c] = y[(size_t)i * M + c];
            for (int i = lane; i < N; i += WAVE)
#pragma unroll
                for (int c = 0; c < M; ++c)
                    gy[(size_t)(2 * i + 1) * M + c] = (y[(size_t)(i + 1) * M + c] + y[(size_t)i * M + c]) / 2.0;
            __syncthreads();
        };
        fill();
        const NewtonOut nl = newton<M, METHOD>(B, I, sh, gm, gy, fill, dl, yt, gb, 2 * N, lane, scr, tree_base, nullptr);
        cnt_it += 2LL * N * nl.iters;
        hs = 2;
    } else if constexpr (METHOD == 4) {
        auto fill = [&]() {
            for (int q = lane; q < (N + 1) * M; q += WAVE) gy[q] = y[q];
            __syncthreads();
        };
        fill();
        const NewtonOut nl = newton<M, 6>(B, I, sh, mesh, gy, fill, dl, yt, gb, N, lane, scr, tree_base, nullptr);
        cnt_it += (long long)N * nl.iters;
    }
    double nrm = 0.0;
    bool bad = false;
    for (int i = lane; i < N; i += WAVE) {
        double ea[M], eb[M], a = 0.0, b = 0.0;
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const double lo = y[(size_t)i * M + c], lo1 = y[(size_t)(i + 1) * M + c];
            ea[c] = (gy[(size_t)hs * i * M + c] - lo) / (1.0 + fabs(lo));
            eb[c] = (gy[(size_t)hs * (i + 1) * M + c] - lo1) / (1.0 + fabs(lo1));
            a = jmax(a, fabs(ea[c]));
            b = jmax(b, fabs(eb[c]));
        }
        const bool left = a >= b;
        const double e = left ? a : b;
        ind[i] = e;
        if (errs)
#pragma unroll
            for (int c = 0; c < M; ++c) errs[(size_t)i * M + c] = left ? ea[c] : eb[c];
        bad |= isnan(e);
        nrm = fmax(nrm, e);
    }
    const bool anybad = wave_any(bad);
    nrm = wave_max(nrm);
    __syncthreads();
    if (anybad) return NAN;
    if (B.global_method == BVP_GLOBAL_RE) {
        const double p2 = pow(2.0, (double)Tab<METHOD>::order);
        return nrm * p2 / (p2 - 1.0);
    }
    return nrm;
}

// ind[i] -> (ind[i] / abstol)^ex (add: added to what shat holds), r1 and r2 the max
// and sum over the intervals (chunk order, as the defect pass)
__device__ inline void indicator(double *shat, const double *ind, int N, int lane, double abstol, double ex, bool add,
                                 int minc, double &r1, double &r2) {
    int P, start, len;
    chunk_of(N, lane, P, start, len, minc);
    double smax = 0.0, ssum = 0.0;
    for (int k = start; k < start + len; ++k) {
        double v = pow(ind[k] / abstol, ex);
        if (add) v = shat[k] + v;
        shat[k] = v;
        smax = fmax(smax, v);
        ssum += v;
    }
    r1 = wave_max(smax);
    r2 = wave_sum(ssum);
    __syncthreads();
}

// error_estimate! of the global-error controllers at a converged state
// (MIRK/adaptivity.jl:463-525, Core/calc_errors.jl:67-163) and the indicator their
// mesh_selector! equidistributes (MIRK/adaptivity.jl:77-243; halving threshold 2):
//   GLOBAL      shat = (ge / abstol)^(1/p)
//   SEQUENTIAL  the defect first (its threshold test sets info); when it is <= abstol
//               the global error replaces it; shat = (error / abstol)^(1/(p+1))
//   HYBRID      DE * defect + GE * global error; shat = (d / abstol)^(1/(p+1)) +
//               (ge / abstol)^(1/(p+1)); no threshold test
// Returns the error norm the outer loop compares with abstol.
template <int M, int METHOD>
__device__ double control_estimate(const BatchDev &B, const Inst<M> &I, Shared<M> &sh, const double *mesh,
                                   const double *y, int N, int lane, double *scr, size_t tree_base, double *shat,
                                   double *ge, double &last_defect, int &info, double &r1, double &r2,
                                   long long &cnt_it, long long &cnt_def) {
    constexpr int p = Tab<METHOD>::order;
    const double abstol = B.abstol;
    // interval indicators of the global error: past the nested solve's buffers
    double *ind = ge + (size_t)2 * B.cap * (3 * M + 1);
    if (B.controller == BVP_CONTROL_GLOBAL) {
        const double gn = global_error<M, METHOD>(B, I, sh, mesh, y, N, lane, scr, tree_base, ge, ind, nullptr, cnt_it);
        indicator(shat, ind, N, lane, abstol, 1.0 / p, false, I.minc, r1, r2);
        return gn;
    }
    const double dn = defect_pass<M, METHOD>(I, mesh, y, N, lane, abstol, shat, nullptr, r1, r2);
    cnt_def += N;
    last_defect = dn;
    __syncthreads();
    if (B.controller == BVP_CONTROL_SEQUENTIAL) {
        info = (dn > B.defect_threshold || isnan(dn)) ? BVP_RET_FAILURE : BVP_RET_SUCCESS;
        if (info == BVP_RET_SUCCESS && dn <= abstol) {
            const double gn = global_error<M, METHOD>(B, I, sh, mesh, y, N, lane, scr, tree_base, ge, ind, nullptr, cnt_it);
            indicator(shat, ind, N, lane, abstol, 1.0 / (p + 1), false, I.minc, r1, r2);
            return gn;
        }
        return dn;
    }
    const double gn = global_error<M, METHOD>(B, I, sh, mesh, y, N, lane, scr, tree_base, ge, ind, nullptr, cnt_it);
    indicator(shat, ind, N, lane, abstol, 1.0 / (p + 1), true, I.minc, r1, r2);
    return B.hybrid_de * dn + B.hybrid_ge * gn;
}

// ------------------------------------------------------------------ instance buffers
template <int M> struct Bufs {
    double *mesh[2];
    double *y[2];
    // mesh 0 as read: the buffer, or the packed initial mesh until a mesh is
    // written into buffer 0 (the solve path does not copy the initial mesh)
    const double *m0;
    // buffer c (0 or 1) chosen with a select, so that the array is never indexed
    // at run time and the pointers keep their global address space; meshes are
    // written through mesh[] only
    __device__ const double *mesh_of(int c) const { return c ? mesh[1] : m0; }
    __device__ double *y_of(int c) const { return c ? y[1] : y[0]; }
};

template <int M>
__device__ inline Bufs<M> bufs_of(const BatchDev &B, int b) {
    Bufs<M> r;
    const size_t cap = (size_t)B.cap;
    r.mesh[0] = as_global(B.mesh + (size_t)b * 2 * cap);
    r.mesh[1] = r.mesh[0] + cap;
    r.y[0] = as_global(B.y + (size_t)b * 2 * cap * M);
    r.y[1] = r.y[0] + cap * M;
    r.m0 = r.mesh[0];
    return r;
}

// ------------------------------------------------------------------ full solve of one instance
// FAM >= 0: the batch holds one family only, known at compile time, and only
// two-point boundary conditions; every f evaluation then folds to that family's
// code and the interior-node boundary walk drops out, which lowers the register
// peak the other paths (sin/cos reduction and the like) impose.
// GE: the batch runs a global-error controller (B.controller != DEFECT) with the
// slot's global-error buffers ge.
template <int M, int METHOD, int FAM = -1, bool F32 = false, bool GE = false>
__device__ __forceinline__ void solve_instance(const BatchDev &B, int b, Shared<M> &sh, int lane, double *scr, size_t tree_base,
                               double *aux, double *ge = nullptr) {
    Inst<M> I;
    load_inst<M>(B, b, I);
    if constexpr (FAM >= 0) {
        I.fam = FAM;
        I.bc_kind = BVP_BC_TWO_POINT;  // the specialised kernels serve two-point batches only
    }
    Bufs<M> bf = bufs_of<M>(B, b);
    int cur = 0;
    int N = B.N0[b];
    // initial guess from the packed inputs (or the learned guess from the anchors);
    // the initial mesh is read where it was packed
    const long long off0 = B.init_off[b];
    bf.m0 = as_global(B.init_mesh + off0);
    auto initial_state = [&]() {
        const long long off = off0;
        // U nodes per lane per round: every load of a round is issued before its stores
        constexpr int U = 4;
        if constexpr (M == 2) {
            if (B.anc) {
                const double t0 = B.t0[b], t1 = B.t1[b];
                const int K = B.anc_K;
                if (K <= 4) {
                    // all anchor records of the instance in registers, loaded together
                    // with the first node times (no load waits on a node time)
                    double A[4][4];
                    const double *ab = as_global(B.anc) + (size_t)b * K * 4;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int r = 0; r < 4; ++r) A[q][r] = q < K ? ab[q * 4 + r] : 0.0;
                    for (int k0 = lane; k0 <= N; k0 += U * WAVE) {
                        double t[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int k = k0 + u * WAVE;
                            t[u] = k <= N ? B.init_mesh[off + k] : t0;
                        }
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int k = k0 + u * WAVE;
                            if (k <= N) {
                                const int j = anchor_interval(K, t0, t1, t[u]);  // 0 .. K-2 <= 2
                                double rec[8];
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    rec[r] = j == 0 ? A[0][r] : (j == 1 ? A[1][r] : A[2][r]);
                                    rec[4 + r] = j == 0 ? A[1][r] : (j == 1 ? A[2][r] : A[3][r]);
                                }
                                                                anchor_eval(rec, K, t0, t1, t[u], bf.y[0] + (size_t)k * M);
                            }
                        }
                    }
                } else
                for (int k0 = lane; k0 <= N; k0 += U * WAVE) {
                    double t[U], rec[U][8];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = k0 + u * WAVE;
                        t[u] = k <= N ? B.init_mesh[off + k] : t0;
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const double *a = as_global(B.anc) + ((size_t)b * K + anchor_interval(K, t0, t1, t[u])) * 4;
#pragma unroll
                        for (int q = 0; q < 8; ++q) rec[u][q] = a[q];
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = k0 + u * WAVE;
                        if (k <= N) {
                                                        anchor_eval(rec[u], K, t0, t1, t[u], bf.y[0] + (size_t)k * M);
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (!(M == 2 && B.anc)) {
            for (int k0 = lane; k0 <= N; k0 += U * WAVE) {
                double t[U], v[U][M];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = k0 + u * WAVE;
                    if (k <= N) {
                        t[u] = B.init_mesh[off + k];
#pragma unroll
                        for (int c = 0; c < M; ++c) v[u][c] = B.init_y[(off + k) * M + c];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = k0 + u * WAVE;
                    if (k <= N) {
                        #pragma unroll
                        for (int c = 0; c < M; ++c) bf.y[0][(size_t)k * M + c] = v[u][c];
                    }
                }
            }
            __syncthreads();
        }
    };
    initial_state();
    double *shat = as_global(B.shat + (size_t)b * B.cap);
    double *delta = aux;                               // [cap*M]
    double *ytmp = aux + (size_t)B.cap * M;            // [cap*M]
    double *iiscr = aux + (size_t)2 * B.cap * M;       // [cap]
    double *gbuf = aux + (size_t)2 * B.cap * M + B.cap;  // [cap*M] J'F of the trust-region stage
    int info = BVP_RET_SUCCESS, nl_ret = BVP_RET_SUCCESS;
    int outer = 0, iters = 0, fb = 0;
    double last_defect = NAN;
    long long cnt_it = 0, cnt_def = 0, cnt_tr = 0, cnt_chk = 0;
    double rn = NAN;
    bool rn_valid = false;  // rn is max|F| at the current (mesh, y)
    // where the current Newton solve starts: 0 the initial state, 1 the transfer from
    // the other buffer (prevN intervals), 2 zeros (halving); restore() rebuilds it
    int start_kind = 0, prevN = 0;
    auto restore = [&]() {
        double *yc = bf.y_of(cur);
        if (start_kind == 0) {
            initial_state();
        } else if (start_kind == 1) {
            transfer<M, METHOD>(I, bf.mesh_of(1 - cur), bf.y_of(1 - cur), prevN, bf.mesh_of(cur), yc, N,
                                B.transfer_mode, lane, iiscr);
        } else {
            for (int k = lane; k <= N; k += WAVE)
#pragma unroll
                for (int c = 0; c < M; ++c) yc[(size_t)k * M + c] = 0.0;
        }
    };
    for (;;) {
        ++outer;
        const double *mesh = bf.mesh_of(cur);
        double *y = bf.y_of(cur);
        // the other buffer, as written (a new mesh goes to the buffer itself)
        double *mesh_o = cur ? bf.mesh[0] : bf.mesh[1], *y_o = bf.y_of(1 - cur);
        const double *u0c = (B.warm && outer == 1) ? as_global(B.u0c + (size_t)b * M) : nullptr;
        NewtonOut nl = newton<M, METHOD, F32, FAM>(B, I, sh, mesh, y, restore, delta, ytmp, gbuf, N, lane, scr,
                                                   tree_base, u0c);
        iters += nl.iters;
        rn = nl.fn;
        rn_valid = true;
        cnt_it += (long long)N * nl.iters;
        // separate primal residual passes: the checks of the line-search and trust-region
        // stages (the plain stage checks inside its back-substitution)
        if (nl.stage > 0) cnt_chk += (long long)N * nl.iters;
        if (nl.stage == 1) fb |= 1;
        if (nl.stage == 2) fb |= 16;
        if (nl.warm_fallback) fb |= 2;
        nl_ret = nl.ret;
        double error_norm = 2.0 * B.abstol;
        if (!B.adaptive) { info = nl.ret; break; }
        info = nl.ret;
        double r1 = 0.0, r2 = 0.0, rho = 1.0;
        if constexpr (GE) {
            if (info == BVP_RET_SUCCESS) {
                error_norm = control_estimate<M, METHOD>(B, I, sh, mesh, y, N, lane, scr, tree_base, shat, ge,
                                                         last_defect, info, r1, r2, cnt_it, cnt_def);
                rho = 2.0;
            }
        } else if (info == BVP_RET_SUCCESS) {
            if (nl.dn_valid && !(nl.dn > B.abstol)) {
                // the defect came with the last correction and no mesh selection follows
                error_norm = nl.dn;
            } else {
                error_norm = defect_pass<M, METHOD>(I, mesh, y, N, lane, B.abstol, shat, nullptr, r1, r2,
                                                    N <= kStageNodes<M> * M ? sh.stage : nullptr, true);
                cnt_def += N;
            }
            last_defect = error_norm;
            info = (error_norm > B.defect_threshold || isnan(error_norm)) ? BVP_RET_FAILURE : BVP_RET_SUCCESS;
        }
        __syncthreads();
        if (info == BVP_RET_SUCCESS) {
            if (error_norm > B.abstol) {
                if (cur == 1) bf.m0 = bf.mesh[0];  // buffer 0 receives the new mesh
                const int Nn = mesh_select(mesh, N, shat, r1, r2, B.max_sub, mesh_o, lane, sh, rho);
                if (Nn < 0) {
                    info = BVP_RET_FAILURE;
                } else {
                    transfer<M, METHOD>(I, mesh, y, N, mesh_o, y_o, Nn, B.transfer_mode, lane, iiscr);
                    cnt_tr += N + Nn;
                    prevN = N;
                    start_kind = 1;
                    N = Nn;
                    cur = 1 - cur;
                    rn_valid = false;
                }
            }
        } else {
            if (2 * N > B.max_sub) {
                info = BVP_RET_FAILURE;
            } else {
                if (cur == 1) bf.m0 = bf.mesh[0];  // buffer 0 receives the halved mesh
                half_mesh_into(mesh, N, mesh_o, lane);
                for (int k = lane; k <= 2 * N; k += WAVE)
#pragma unroll
                    for (int c = 0; c < M; ++c) y_o[(size_t)k * M + c] = 0.0;
                __syncthreads();
                N = 2 * N;
                cur = 1 - cur;
                start_kind = 2;
                rn_valid = false;
                info = BVP_RET_SUCCESS;
            }
        }
        if (!(info == BVP_RET_SUCCESS && error_norm > B.abstol)) break;
        if (outer >= MAX_OUTER) { info = BVP_RET_FAILURE; break; }
    }
    // final residual norm: Newton's last check unless the mesh changed after it
    if (!rn_valid) rn = residual_primal<M, METHOD>(I, bf.mesh_of(cur), bf.y_of(cur), N, lane);
    // distance of the returned state from the FP64 discrete solution on its mesh
    const double err = B.report_err ? correction_norm<M, METHOD, F32, FAM>(I, sh, bf.mesh_of(cur), bf.y_of(cur), delta,
                                                                            N, lane, scr, tree_base)
                                    : NAN;
    if (lane == 0) {
        B.err[b] = err;
        B.retcode[b] = (nl_ret == BVP_RET_SUCCESS) ? info : nl_ret;
        B.N_final[b] = N;
        B.newton_iters[b] = iters;
        B.outer_iters[b] = outer;
        B.fallback[b] = fb;
        // 2: the final mesh is the packed initial mesh (state in buffer 0)
        B.cur_buf[b] = (cur == 0 && bf.m0 != bf.mesh[0]) ? 2 : cur;
        B.defect_norm[b] = last_defect;
        B.resid_norm[b] = rn;
        B.cnt[(size_t)b * 4 + 0] = cnt_it;
        B.cnt[(size_t)b * 4 + 1] = cnt_def;
        B.cnt[(size_t)b * 4 + 2] = cnt_tr;
        B.cnt[(size_t)b * 4 + 3] = cnt_chk;
    }
    __syncthreads();
}

// ------------------------------------------------------------------ kernels
template <int M>
__device__ inline int next_item(unsigned *counter, Shared<M> &sh, int lane) {
    if (lane == 0) sh.item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = sh.item;
    __syncthreads();
    return it;
}

// Persistent fused solve: each wave pulls instances from the queue (ordered by
// decreasing size) and runs the whole residual -> Jacobian -> LU -> update ->
// defect -> mesh loop for it.  WPE = waves per SIMD the register allocation is
// held to (1: no spills; more: more latency hiding, some spills) -- a variant
// the autotuner picks by measurement.
template <int M, int METHOD, int WPE, int FAM = -1, bool F32 = false>
__global__ __launch_bounds__(64, WPE) void k_solve(BatchDev B) {
    __shared__ Shared<M> sh;
    const int lane = threadIdx.x;
    double *scr = as_global(B.scratch + (size_t)blockIdx.x * B.scratch_per_slot);
    double *aux = as_global(B.aux + (size_t)blockIdx.x * B.aux_per_slot);
    for (;;) {
        const int item = next_item<M>(B.counter, sh, lane);
        if (item >= B.n) break;
        solve_instance<M, METHOD, FAM, F32>(B, B.order[item], sh, lane, scr, B.tree_base, aux);
    }
}

// The persistent solve under a global-error controller: the generic family code,
// FP64 records, the register allocation of one wave per SIMD.
template <int M, int METHOD>
__global__ __launch_bounds__(64, 1) void k_solve_ge(BatchDev B) {
    __shared__ Shared<M> sh;
    const int lane = threadIdx.x;
    double *scr = as_global(B.scratch + (size_t)blockIdx.x * B.scratch_per_slot);
    double *aux = as_global(B.aux + (size_t)blockIdx.x * B.aux_per_slot);
    double *ge = as_global(B.ge + (size_t)blockIdx.x * B.ge_per_slot);
    for (;;) {
        const int item = next_item<M>(B.counter, sh, lane);
        if (item >= B.n) break;
        solve_instance<M, METHOD, -1, false, true>(B, B.order[item], sh, lane, scr, B.tree_base, aux, ge);
    }
}

// The global-error estimate of each instance at its packed state (mesh and guess):
// out0 the interval error vectors at the instance's interval offset, out1 the norm.
template <int M, int METHOD>
__global__ __launch_bounds__(64, 1) void k_global_error(BatchDev B) {
    __shared__ Shared<M> sh;
    const int lane = threadIdx.x;
    double *scr = as_global(B.scratch + (size_t)blockIdx.x * B.scratch_per_slot);
    double *ge = as_global(B.ge + (size_t)blockIdx.x * B.ge_per_slot);
    for (;;) {
        const int b = next_item<M>(B.counter, sh, lane);
        if (b >= B.n) break;
        Inst<M> I;
        load_inst<M>(B, b, I);
        const int N = B.N0[b];
        const long long off = B.init_off[b];
        const double *mesh = as_global(B.init_mesh + off);
        const double *y = as_global(B.init_y + off * M);
        const long long ioff = off - b;  // interval offset = node offset - instance index
        double *ind = ge + (size_t)2 * B.cap * (3 * M + 1);
        long long cnt = 0;
        const double g = global_error<M, METHOD>(B, I, sh, mesh, y, N, lane, scr, B.tree_base, ge, ind,
                                                 B.out0 + ioff * M, cnt);
        if (lane == 0) B.out1[b] = g;
        __syncthreads();
    }
}

// Component kernels: one instance per queue item, state read from the packed
// inputs (mesh and guess), outputs packed by the offsets in the descriptor.
template <int M, int METHOD>
__global__ __launch_bounds__(64) void k_component(BatchDev B, int op) {
    __shared__ Shared<M> sh;
    const int lane = threadIdx.x;
    double *scr = as_global(B.scratch + (size_t)blockIdx.x * B.scratch_per_slot);
    double *aux = as_global(B.aux + (size_t)blockIdx.x * B.aux_per_slot);
    for (;;) {
        const int b = next_item<M>(B.counter, sh, lane);
        if (b >= B.n) break;
        Inst<M> I;
        load_inst<M>(B, b, I);
        const int N = B.N0[b];
        const long long off = B.init_off[b];
        // private working copy of the state (the solve may update y)
        Bufs<M> bf = bufs_of<M>(B, b);
        for (int k = lane; k <= N; k += WAVE) {
            bf.mesh[0][k] = B.init_mesh[off + k];
#pragma unroll
            for (int c = 0; c < M; ++c) bf.y[0][(size_t)k * M + c] = B.init_y[(off + k) * M + c];
        }
        __syncthreads();
        const double *mesh = bf.mesh[0];
        double *y = bf.y[0];
        if (op == OP_RESIDUAL) {
            // reference layout: TWO_POINT [bca; Phi; bcb], STANDARD [bc; Phi]
            double *out = B.out0 + off * M;
            int La = 0;
            if (I.bc_kind == BVP_BC_TWO_POINT) {
#pragma unroll
                for (int k = 0; k < M; ++k) La += (I.bct[k] == I.t0) ? 1 : 0;
            } else {
                La = M;
            }
            Rel<M> rel;
            double mx;
            bool bad;
            forward_dual<M, METHOD>(I, mesh, y, N, lane, nullptr, rel, mx, bad, out + La, nullptr, nullptr);
            if (lane == 0) {
                int ia = 0, ib = 0;
                for (int k = 0; k < M; ++k) {
                    const BcLoc L = bc_locate<M>(I, k, mesh, N);
                    const double r = bc_residual<M, METHOD>(I, k, L, mesh, y);
                    if (I.bc_kind == BVP_BC_TWO_POINT && sel<M>(I.bct, k) != I.t0) out[La + (size_t)N * M + (ib++)] = r;
                    else out[ia++] = r;
                }
            }
        } else if (op == OP_JACOBIAN) {
            Rel<M> rel;
            double mx;
            bool bad;
            const long long ioff = off - b;  // interval offset = node offset - instance index
            forward_dual<M, METHOD>(I, mesh, y, N, lane, nullptr, rel, mx, bad, nullptr, B.out0 + ioff * M * M,
                                    B.out1 + ioff * M * M);
            if (lane < M) B.iout[(size_t)b * M + lane] = bc_locate<M>(I, lane, mesh, N).node;
        } else if (op == OP_NEWTON_STEP) {
            Rel<M> rel;
            bool ok;
            if (B.f32rec) {
                (void)residual_and_blocks<M, METHOD, true>(I, sh, mesh, y, N, lane, scr, rel);
                ok = solve_correction<M, METHOD, true>(I, sh, mesh, y, N, lane, scr, B.tree_base, rel,
                                                       B.out0 + off * M, false);
            } else {
                (void)residual_and_blocks<M, METHOD>(I, sh, mesh, y, N, lane, scr, rel);
                ok = solve_correction<M, METHOD>(I, sh, mesh, y, N, lane, scr, B.tree_base, rel,
                                                 B.out0 + off * M, false);
            }
            if (lane == 0) B.iout[b] = ok ? 0 : 1;
        } else if (op == OP_DEFECT) {
            double r1, r2;
            const long long ioff = off - b;
            const double dn = defect_pass<M, METHOD>(I, mesh, y, N, lane, B.abstol, nullptr, B.out0 + ioff * M, r1, r2);
            if (lane == 0) B.out1[b] = dn;
        } else if (op == OP_MESH_SELECT) {
            double r1, r2;
            double *shat = as_global(B.shat + (size_t)b * B.cap);
            const double dn = defect_pass<M, METHOD>(I, mesh, y, N, lane, B.abstol, shat, nullptr, r1, r2);
            (void)dn;
            __syncthreads();
            const int Nn = mesh_select(mesh, N, shat, r1, r2, B.max_sub, bf.mesh[1], lane, sh);
            if (Nn < 0) {
                if (lane == 0) { B.iout[(size_t)b * 2] = N; B.iout[(size_t)b * 2 + 1] = BVP_RET_FAILURE; }
            } else {
                double *iiscr = aux + (size_t)2 * B.cap * M;
                transfer<M, METHOD>(I, mesh, y, N, bf.mesh[1], bf.y[1], Nn, B.transfer_mode, lane, iiscr);
                double *om = B.out0 + (size_t)b * B.cap, *oy = B.out1 + (size_t)b * B.cap * M;
                for (int k = lane; k <= Nn; k += WAVE) {
                    om[k] = bf.mesh[1][k];
#pragma unroll
                    for (int c = 0; c < M; ++c) oy[(size_t)k * M + c] = bf.y[1][(size_t)k * M + c];
                }
                if (lane == 0) { B.iout[(size_t)b * 2] = Nn; B.iout[(size_t)b * 2 + 1] = BVP_RET_SUCCESS; }
            }
        }
        __syncthreads();
    }
}

// Gather the final state of every instance into packed outputs (offsets given
// per instance in out_off, in nodes); one block per instance.
template <int M>
__global__ __launch_bounds__(256) void k_gather(BatchDev B, const long long *out_off, double *out_mesh,
                                                double *out_y) {
    const int b = blockIdx.x;
    if (b >= B.n) return;
    const int N = B.N_final[b];
    int cur = B.cur_buf[b];
    Bufs<M> bf = bufs_of<M>(B, b);
    if (cur == 2) {  // the mesh never left the packed inputs
        bf.m0 = B.init_mesh + B.init_off[b];
        cur = 0;
    }
    const long long o = out_off[b];
    for (int k = threadIdx.x; k <= N; k += blockDim.x) {
        out_mesh[o + k] = bf.mesh_of(cur)[k];
#pragma unroll
        for (int c = 0; c < M; ++c) out_y[(o + k) * M + c] = bf.y_of(cur)[(size_t)k * M + c];
    }
}

// Dense output (MIRKInterpolation, MIRK/interpolation.jl:17-206): one thread per
// query point.  The solution of instance b is its packed (mesh, y); point q lies
// in interval(mesh, t) (a node belongs to the interval on its left; points
// outside [t0, t1] extrapolate the end intervals), tau = (t - t_i)/h, and
//   u(t) = y_i + h sum_r w_r(tau) K_r,   u'(t) = sum_r w'_r(tau) K_r
// over the discrete stages K at (y_i, y_{i+1}) and the extra interpolation stages.
template <int M, int METHOD>
__global__ __launch_bounds__(256) void k_interp(BatchDev B, const int *qinst, const double *tq, double *uo,
                                                double *duo, long long nq) {
    using TB = Tab<METHOD>;
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    const int b = qinst[q];
    Inst<M> I;
    load_inst<M>(B, b, I);
    const long long off = B.init_off[b];
    const double *mesh = as_global(B.init_mesh + off);
    const double *y = as_global(B.init_y + off * M);
    const int N = B.N0[b];
    const double t = tq[q];
    const int i = interval_of(mesh, N, t);
    const double ti = mesh[i], h = mesh[i + 1] - ti, tau = (t - ti) / h;
    double yi[M], yj[M];
#pragma unroll
    for (int c = 0; c < M; ++c) { yi[c] = y[(size_t)i * M + c]; yj[c] = y[(size_t)(i + 1) * M + c]; }
    double K[TB::s][M], KI[TB::sstar - TB::s][M];
    stages<M, METHOD, double>(I.fam, I.P, ti, h, yi, yj, K);
    interp_stages<M, METHOD>(I.fam, I.P, ti, h, yi, yj, K, KI);
    double w[TB::sstar], wp[TB::sstar], z[M], zp[M];
    interp_weights<METHOD>(tau, w, wp);
    sum_stages<M, METHOD>(h, yi, K, KI, w, wp, z, zp);
#pragma unroll
    for (int c = 0; c < M; ++c) {
        uo[q * M + c] = z[c];
        if (duo) duo[q * M + c] = zp[c];
    }
}

// ------------------------------------------------------------------ host launchers
template <int M, int METHOD, int WPE, int FAM = -1, bool F32 = false>
hipError_t launch_solve_t(const BatchDev &B, int blocks, hipStream_t s) {
    k_solve<M, METHOD, WPE, FAM, F32><<<blocks, 64, 0, s>>>(B);
    return hipGetLastError();
}
template <int M, int METHOD>
hipError_t launch_component_t(const BatchDev &B, int op, int blocks, hipStream_t s) {
    k_component<M, METHOD><<<blocks, 64, 0, s>>>(B, op);
    return hipGetLastError();
}
template <int M, int METHOD>
hipError_t launch_solve_ge_t(const BatchDev &B, int blocks, hipStream_t s) {
    k_solve_ge<M, METHOD><<<blocks, 64, 0, s>>>(B);
    return hipGetLastError();
}
template <int M, int METHOD>
hipError_t launch_global_error_t(const BatchDev &B, int blocks, hipStream_t s) {
    k_global_error<M, METHOD><<<blocks, 64, 0, s>>>(B);
    return hipGetLastError();
}

template <int M, int METHOD, int WPE, int FAM = -1, bool F32 = false>
int occupancy_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_solve<M, METHOD, WPE, FAM, F32>, 64, 0) != hipSuccess)
        return 4 * WPE;
    return nb;
}

}  // namespace bvpdev
