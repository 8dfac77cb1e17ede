// mirk_block_impl.h -- MIRK collocation on gfx950 for larger state dimensions
// (M = 8, 16, 32; family LINEAR_TANH f = A y + c tanh(y)), one 64-lane wavefront
// per BVP instance.
//
// Where the small-M kernel gives each lane a chunk of intervals, here each
// interval is a block operation done by the whole wave, and the intervals are
// walked in order (a batch of thousands of instances fills the GPU):
//   * FP64 residual, defect and Jacobian-vector products with lanes over the M
//     components (matrix-vector products read their operand from LDS);
//   * Jacobian blocks [S_k | R_k] as 2M tangent columns, one per lane, pushed
//     through the stage recursion (forward-mode differentiation with 2M
//     directions, in the factorisation precision LT);
//   * block elimination with partial pivoting, one row of the stacked 2M x
//     (3M+1) system per lane; pivot rows are not swapped, the pivot lane is
//     recorded instead, and the rows that remain form the relation carried to
//     the next interval;
//   * LT = float gives the FP32 factorisation; FP64 iterative refinement then
//     solves J d = -F to full accuracy with residuals r = -F - J d computed by a
//     one-direction dual (J d) in FP64.
// Reference counterparts as in mirk_wave_impl.h (MIRK/collocation.jl:44-102,
// MIRK/mirk.jl:697-862 and 927-1002, MIRK/adaptivity.jl:23-75, 250-304, 370-461).
#pragma once

#include "mirk_wave_impl.h"

namespace bvpdev {

template <int M, int METHOD, typename LT> struct BShared {
    static constexpr int S = Tab<METHOD>::s;
    static constexpr int NS = Tab<METHOD>::sstar - Tab<METHOD>::s;
    static constexpr int W = 3 * M + 1;  // stacked row: pivot block | y_0 block | y_{k+1} block | rhs
    double At[M][M];                     // A transposed (At[q][i] = A[i][q]): conflict-free lane reads
    LT Al[M][M];                         // A in the factorisation precision, row-major (broadcast reads)
    double c;
    double va[M], vb[M], vd[M];          // node vectors (state or corrections)
    double Y[S][M], K[S][M], KI[NS][M];  // stage arguments and values
    double Yd[S][M], Kd[S][M];           // their tangents (J v products)
    LT sech2[S][M];
    LT G[M][2 * M + 1];                  // rows [S | R | r] of the new block row
    LT rel[M][2 * M + 1];                // relation rows [E | F | g] for the next interval
    LT prow[W];                          // pivot row broadcast
    double drow[2 * M + 1];              // root pivot row broadcast
    double bcres[M];
    int piv[2 * M];
    int flag, item;
};

// ------------------------------------------------------------------ FP64 stage work (lanes = components)
// K_r = f(Y_r), Y_r = (1-v_r) a + v_r b + h sum_{j<r} x_rj K_j ; lanes c < M.
template <int M, int METHOD, typename LT>
__device__ inline void bstages(BShared<M, METHOD, LT> &sh, int lane, double h) {
    using TB = Tab<METHOD>;
#pragma unroll
    for (int r = 0; r < TB::s; ++r) {
        if (lane < M) {
            double y = (1.0 - TB::v(r)) * sh.va[lane] + TB::v(r) * sh.vb[lane];
            double acc = 0.0;
            bool any = false;
#pragma unroll
            for (int j = 0; j < r; ++j)
                if (TB::x(r, j) != 0.0) { acc += TB::x(r, j) * sh.K[j][lane]; any = true; }
            if (any) y += h * acc;
            sh.Y[r][lane] = y;
        }
        __syncthreads();
        if (lane < M) {
            double acc = 0.0;
#pragma unroll 8
            for (int q = 0; q < M; ++q) acc += sh.At[q][lane] * sh.Y[r][q];
            const double th = tanh(sh.Y[r][lane]);
            sh.K[r][lane] = acc + sh.c * th;
            sh.sech2[r][lane] = (LT)(1.0 - th * th);
        }
        __syncthreads();
    }
}

// Phi_k component of this lane: (b - a) - h sum b_r K_r
template <int M, int METHOD, typename LT>
__device__ inline double bphi(const BShared<M, METHOD, LT> &sh, int lane, double h) {
    using TB = Tab<METHOD>;
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < TB::s; ++r) acc += TB::b(r) * sh.K[r][lane];
    return (sh.vb[lane] - sh.va[lane]) - h * acc;
}

// Directional derivative of this lane's component of Phi_k along (da, db) at the
// stages already in sh.Y/K: S_k da + R_k db, exactly in FP64 (a one-direction dual).
template <int M, int METHOD, typename LT>
__device__ inline double bphi_tangent(BShared<M, METHOD, LT> &sh, int lane, double h, double da, double db) {
    using TB = Tab<METHOD>;
#pragma unroll
    for (int r = 0; r < TB::s; ++r) {
        if (lane < M) {
            double y = (1.0 - TB::v(r)) * da + TB::v(r) * db;
            double acc = 0.0;
            bool any = false;
#pragma unroll
            for (int j = 0; j < r; ++j)
                if (TB::x(r, j) != 0.0) { acc += TB::x(r, j) * sh.Kd[j][lane]; any = true; }
            if (any) y += h * acc;
            sh.Yd[r][lane] = y;
        }
        __syncthreads();
        if (lane < M) {
            double acc = 0.0;
#pragma unroll 8
            for (int q = 0; q < M; ++q) acc += sh.At[q][lane] * sh.Yd[r][q];
            const double th = tanh(sh.Y[r][lane]);
            sh.Kd[r][lane] = acc + sh.c * (1.0 - th * th) * sh.Yd[r][lane];
        }
        __syncthreads();
    }
    double acc = 0.0;
    if (lane < M) {
#pragma unroll
        for (int r = 0; r < TB::s; ++r) acc += TB::b(r) * sh.Kd[r][lane];
    }
    return (db - da) - h * acc;
}

// Extra interpolation stages (MIRK/interpolation.jl:337-372) from sh.K, sh.va, sh.vb.
template <int M, int METHOD, typename LT>
__device__ inline void binterp_stages(BShared<M, METHOD, LT> &sh, int lane, double h) {
    using TB = Tab<METHOD>;
    constexpr int NS = TB::sstar - TB::s;
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        if (lane < M) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < TB::s; ++j)
                if (TB::xs(r, j) != 0.0) a += sh.K[j][lane] * TB::xs(r, j);
#pragma unroll
            for (int j = 0; j < r; ++j)
                if (TB::xs(r, TB::s + j) != 0.0) a += sh.KI[j][lane] * TB::xs(r, TB::s + j);
            sh.vd[lane] = a * h + (1.0 - TB::vs(r)) * sh.va[lane] + TB::vs(r) * sh.vb[lane];
        }
        __syncthreads();
        if (lane < M) {
            double acc = 0.0;
#pragma unroll 8
            for (int q = 0; q < M; ++q) acc += sh.At[q][lane] * sh.vd[q];
            sh.KI[r][lane] = acc + sh.c * tanh(sh.vd[lane]);
        }
        __syncthreads();
    }
}

// f(z) component of this lane with z in sh.vd
template <int M, int METHOD, typename LT>
__device__ inline double bf_of_vd(BShared<M, METHOD, LT> &sh, int lane) {
    double acc = 0.0;
#pragma unroll 8
    for (int q = 0; q < M; ++q) acc += sh.At[q][lane] * sh.vd[q];
    return acc + sh.c * tanh(sh.vd[lane]);
}

// ------------------------------------------------------------------ tangent columns: [S | R]
// Lane l < 2M computes column l of G = sum_r b_r dK_r with dK_r = J_r dY_r, then
// writes S = -I - h G[:, :M], R = I - h G[:, M:] into sh.G (row-major rows).
template <int M, int METHOD, typename LT>
__device__ inline void btangent_blocks(BShared<M, METHOD, LT> &sh, int lane, double hd) {
    using TB = Tab<METHOD>;
    const LT h = (LT)hd;
    const bool left = lane < M;
    const int col = left ? lane : lane - M;  // column of the identity seed inside its half
    LT g[M];   // column l of G
    LT q4[M];  // accumulated x_{s-1,j} dK_j (j >= 2) for the last stage of MIRK6
#pragma unroll
    for (int i = 0; i < M; ++i) { g[i] = (LT)0; q4[i] = (LT)0; }
    if (lane < 2 * M) {
        // stages 0 and 1: dK_0 = [J_0 | 0], dK_1 = [0 | J_1]; column l of J_r is A[:, col] + c sech2_r[col] e_col
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int r0 = left ? 0 : 1;
            LT jv = sh.Al[i][col] + ((i == col) ? (LT)sh.c * sh.sech2[r0][col] : (LT)0);
            g[i] = (LT)TB::b(r0) * jv;
        }
    }
#pragma unroll
    for (int r = 2; r < TB::s; ++r) {
        if (lane < 2 * M) {
            // dY_r column: seed (1-v_r) or v_r on e_col, plus h (x_r0 dK_0 + x_r1 dK_1) + h q (r == s-1)
            LT dy[M];
            const LT seed = left ? (LT)(1.0 - TB::v(r)) : (LT)TB::v(r);
            const LT xr = left ? (LT)TB::x(r, 0) : (LT)TB::x(r, 1);
            const int r0 = left ? 0 : 1;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                LT jv = sh.Al[i][col] + ((i == col) ? (LT)sh.c * sh.sech2[r0][col] : (LT)0);
                dy[i] = ((i == col) ? seed : (LT)0) + h * xr * jv;
                if (r == TB::s - 1 && TB::s > 3) dy[i] += h * q4[i];
            }
            // dK_r = J_r dy
            LT dk[M];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                LT acc = (LT)sh.c * sh.sech2[r][i] * dy[i];
#pragma unroll
                for (int qq = 0; qq < M; ++qq) acc += sh.Al[i][qq] * dy[qq];
                dk[i] = acc;
            }
#pragma unroll
            for (int i = 0; i < M; ++i) {
                g[i] += (LT)TB::b(r) * dk[i];
                if (TB::s > 3 && r < TB::s - 1) q4[i] += (LT)TB::x(TB::s - 1, r) * dk[i];
            }
        }
    }
    if (lane < 2 * M) {
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const LT e = (i == col) ? (left ? (LT)-1 : (LT)1) : (LT)0;
            sh.G[i][lane] = e - h * g[i];
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------ row elimination
// Gaussian elimination with partial pivoting on NC pivot columns of a 64-row
// system held one row per lane (rows of lanes >= nrows are absent).  The pivot
// row of column j is lane piv[j]; it is broadcast through LDS and not swapped.
// On return `done` marks pivot lanes.  When `mult` is non-null, the multiplier
// used on this lane for column j is stored at mult[j * 64].
template <int NC, int NW, typename T>
__device__ inline void eliminate_rows(T (&row)[NW], bool &done, int lane, int nrows, T *prow, int *piv, T *mult) {
    done = lane >= nrows;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        // argmax |row[j]| over not-done lanes (ties: lowest lane)
        double v = done ? -1.0 : (double)fabs((double)row[j]);
        int who = lane;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(v, o);
            const int ow = __shfl_xor(who, o);
            if (ov > v || (ov == v && ow < who)) { v = ov; who = ow; }
        }
        const int p = who;
        // (loops run over the full row with compile-time conditions so that they
        // unroll completely and the row stays in registers)
        if (lane == p) {
#pragma unroll
            for (int c = 0; c < NW; ++c)
                if (c >= j) prow[c] = row[c];
            piv[j] = p;
            done = true;
        }
        __syncthreads();
        const T pv = prow[j];
        if (!done) {
            const T m = row[j] / pv;
            row[j] = (T)0;
#pragma unroll
            for (int c = 0; c < NW; ++c)
                if (c > j) row[c] -= m * prow[c];
            if (mult) mult[(size_t)j * WAVE] = m;
        } else if (mult) {
            mult[(size_t)j * WAVE] = (T)0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ records
// Per interval k (interior node k of the single chunk): for the pivot rows in
// pivot order j: U_j (M), X_j (M), Z_j (M), h_j; with the FP32 factorisation also
// the multipliers [j][lane] and the pivot lanes, so that new right-hand sides
// can be pushed through the same elimination.
template <int M> struct BRec {
    static constexpr size_t U = 0, X = (size_t)M * M, Z = 2 * (size_t)M * M, H = 3 * (size_t)M * M;
    static constexpr size_t MULT = H + M;                    // [M][64]
    static constexpr size_t PIV = MULT + (size_t)M * WAVE;   // [M] (stored as T)
    static constexpr size_t SIZE = PIV + M;
};

// ------------------------------------------------------------------ instance
template <int M> struct BInst {
    int bc_node_is_end[M];  // 0: row at node 0, 1: row at node N
    int bcc[M];
    double bcv[M];
};

template <int M, int METHOD, typename LT>
__device__ inline void bload_inst(const BatchDev &B, int b, BShared<M, METHOD, LT> &sh, BInst<M> &I, int lane) {
    const int po = B.param_off[b];
    const double *A = B.params + po + 1;
    for (int e = lane; e < M * M; e += WAVE) {
        const int i = e / M, q = e % M;
        sh.At[q][i] = A[e];
        sh.Al[i][q] = (LT)A[e];
    }
    if (lane == 0) sh.c = B.params[po];
    const double t0 = B.t0[b];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        I.bc_node_is_end[k] = (B.bc_t[(size_t)b * M + k] == t0) ? 0 : 1;
        I.bcc[k] = B.bc_comp[(size_t)b * M + k];
        I.bcv[k] = B.bc_val[(size_t)b * M + k];
    }
    __syncthreads();
}

// FP64 residual of all intervals at y (+ alpha*delta if delta): returns max|F|
// (NaN if not finite) and, when ss != null, the sum of squares.  BC residuals
// are left in sh.bcres.
template <int M, int METHOD, typename LT>
__device__ double bresidual(const BInst<M> &I, BShared<M, METHOD, LT> &sh, const double *mesh, const double *y,
                            const double *delta, double alpha, int N, int lane, double *ss,
                            double *phi_out = nullptr) {
    double mx = 0.0, acc = 0.0;
    bool bad = false;
    for (int k = 0; k < N; ++k) {
        const double h = mesh[k + 1] - mesh[k];
        if (lane < M) {
            double a = y[(size_t)k * M + lane], bv = y[(size_t)(k + 1) * M + lane];
            if (delta) { a += alpha * delta[(size_t)k * M + lane]; bv += alpha * delta[(size_t)(k + 1) * M + lane]; }
            sh.va[lane] = a;
            sh.vb[lane] = bv;
        }
        __syncthreads();
        bstages<M, METHOD, LT>(sh, lane, h);
        if (lane < M) {
            const double ph = bphi<M, METHOD, LT>(sh, lane, h);
            if (phi_out) phi_out[(size_t)k * M + lane] = ph;
            bad |= !(fabs(ph) <= 1.7976931348623157e308);
            mx = fmax(mx, fabs(ph));
            acc += ph * ph;
        }
        __syncthreads();
    }
    if (lane < M) {
        const int node = sel<M>(I.bc_node_is_end, lane) ? N : 0;
        const int comp = sel<M>(I.bcc, lane);
        double v = y[(size_t)node * M + comp];
        if (delta) v += alpha * delta[(size_t)node * M + comp];
        const double r = v - sel<M>(I.bcv, lane);
        sh.bcres[lane] = r;
        bad |= !(fabs(r) <= 1.7976931348623157e308);
        mx = fmax(mx, fabs(r));
        acc += r * r;
    }
    __syncthreads();
    const bool anybad = wave_any(bad);
    mx = wave_max(mx);
    acc = wave_sum(acc);
    if (ss) *ss = anybad ? NAN : acc;
    return anybad ? NAN : mx;
}

// Factorise J(y) interval by interval (single chunk from node 0) and solve
// J d = rhs where rhs = -F(y) (FP64, computed on the fly) when first == true,
// or rhs taken from `rin` (point-major residual layout: Phi_k at k*M, BC rows at
// N*M + k) when first == false (a refinement solve through the stored factors).
// The correction is written to dout (point-major, M*(N+1)).
template <int M, int METHOD, typename LT>
__device__ bool bsolve(const BInst<M> &I, BShared<M, METHOD, LT> &sh, const double *mesh, const double *y, int N,
                       int lane, LT *rec, bool factor, const double *rin, double *dout) {
    constexpr int W = 3 * M + 1;
    constexpr bool KEEP = true;  // multipliers are stored so that refinement can reuse the factors
    // ---- forward sweep
    LT row[W];
    bool done;
    for (int k = 0; k < N; ++k) {
        const double h = mesh[k + 1] - mesh[k];
        LT *R = rec + (size_t)k * BRec<M>::SIZE;
        if (factor) {
            if (lane < M) {
                sh.va[lane] = y[(size_t)k * M + lane];
                sh.vb[lane] = y[(size_t)(k + 1) * M + lane];
            }
            __syncthreads();
            bstages<M, METHOD, LT>(sh, lane, h);
            double rk = 0.0;
            if (lane < M) rk = -bphi<M, METHOD, LT>(sh, lane, h);
            btangent_blocks<M, METHOD, LT>(sh, lane, h);
            if (lane < M) sh.G[lane][2 * M] = (LT)rk;
            __syncthreads();
        } else {
            if (lane < M) sh.G[lane][2 * M] = (LT)rin[(size_t)k * M + lane];
            __syncthreads();
        }
        if (k == 0) {
            // the first block row is the relation itself: E = S_0 (on y_0), F = R_0 (on y_1)
            if (lane < M) {
                if (factor) {
#pragma unroll
                    for (int c = 0; c < M; ++c) {
                        sh.rel[lane][c] = sh.G[lane][c];
                        sh.rel[lane][M + c] = sh.G[lane][M + c];
                    }
                }
                sh.rel[lane][2 * M] = sh.G[lane][2 * M];
            }
            __syncthreads();
            continue;
        }
        if (factor) {
            // stacked rows: lanes < M hold the relation [F | E | 0 | g], lanes M..2M-1 hold [S | 0 | R | r]
#pragma unroll
            for (int c = 0; c < W; ++c) row[c] = (LT)0;
            if (lane < M) {
#pragma unroll
                for (int c = 0; c < M; ++c) { row[c] = sh.rel[lane][M + c]; row[M + c] = sh.rel[lane][c]; }
                row[3 * M] = sh.rel[lane][2 * M];
            } else if (lane < 2 * M) {
                const int i = lane - M;
#pragma unroll
                for (int c = 0; c < M; ++c) { row[c] = sh.G[i][c]; row[2 * M + c] = sh.G[i][M + c]; }
                row[3 * M] = sh.G[i][2 * M];
            }
            __syncthreads();
            eliminate_rows<M, W, LT>(row, done, lane, 2 * M, sh.prow, sh.piv, KEEP ? R + BRec<M>::MULT + lane : nullptr);
            // records of the pivot rows, in pivot order
            if (done && lane < 2 * M) {
                int j = 0;
#pragma unroll
                for (int q = 0; q < M; ++q) j = (sh.piv[q] == lane) ? q : j;
#pragma unroll
                for (int c = 0; c < M; ++c) {
                    R[BRec<M>::U + (size_t)c * M + j] = row[c];
                    R[BRec<M>::X + (size_t)c * M + j] = row[M + c];
                    R[BRec<M>::Z + (size_t)c * M + j] = row[2 * M + c];
                }
                R[BRec<M>::H + j] = row[3 * M];
            }
            if (lane < M) R[BRec<M>::PIV + lane] = (LT)sh.piv[lane];
            // compact the remaining rows into the next relation [E' | F' | g']
            const unsigned long long rest = __ballot(!done && lane < 2 * M);
            if (!done && lane < 2 * M) {
                const int slot = __popcll(rest & ((1ull << lane) - 1ull));
#pragma unroll
                for (int c = 0; c < M; ++c) { sh.rel[slot][c] = row[M + c]; sh.rel[slot][M + c] = row[2 * M + c]; }
                sh.rel[slot][2 * M] = row[3 * M];
            }
            __syncthreads();
        } else {
            // push the new right-hand side through the stored elimination
            LT q = (LT)0;
            if (lane < M) q = sh.rel[lane][2 * M];
            else if (lane < 2 * M) q = sh.G[lane - M][2 * M];
            int pv[M];
#pragma unroll
            for (int j = 0; j < M; ++j) pv[j] = (int)R[BRec<M>::PIV + j];
            bool dn = lane >= 2 * M;
#pragma unroll
            for (int j = 0; j < M; ++j) {
                const LT qp = __shfl(q, pv[j]);
                if (lane == pv[j]) { dn = true; R[BRec<M>::H + j] = q; }
                else if (!dn) q -= R[BRec<M>::MULT + (size_t)j * WAVE + lane] * qp;
            }
            const unsigned long long rest = __ballot(!dn && lane < 2 * M);
            __syncthreads();
            if (!dn && lane < 2 * M) {
                const int slot = __popcll(rest & ((1ull << lane) - 1ull));
                sh.rel[slot][2 * M] = q;
            }
            __syncthreads();
        }
    }
    // ---- root: BC rows and the relation E d_0 + F d_N = g (2M x 2M, FP64, row per lane)
    double rr[2 * M + 1];
#pragma unroll
    for (int c = 0; c < 2 * M + 1; ++c) rr[c] = 0.0;
    if (lane < M) {
        const bool end = sel<M>(I.bc_node_is_end, lane) != 0;
        const int comp = sel<M>(I.bcc, lane);
#pragma unroll
        for (int c = 0; c < M; ++c) {
            rr[c] = (!end && c == comp) ? 1.0 : 0.0;
            rr[M + c] = (end && c == comp) ? 1.0 : 0.0;
        }
        rr[2 * M] = factor ? -sh.bcres[lane] : rin[(size_t)N * M + lane];
    } else if (lane < 2 * M) {
        const int i = lane - M;
#pragma unroll
        for (int c = 0; c < 2 * M; ++c) rr[c] = (double)sh.rel[i][c];
        rr[2 * M] = (double)sh.rel[i][2 * M];
    }
    __syncthreads();
    bool rdone;
    eliminate_rows<2 * M, 2 * M + 1, double>(rr, rdone, lane, 2 * M, sh.drow, sh.piv, (double *)nullptr);
    // back substitution over the pivot order; the solution goes to LDS (drow)
#pragma unroll
    for (int j = 2 * M - 1; j >= 0; --j) {
        const int p = sh.piv[j];
        double xj = 0.0;
        if (lane == p) xj = rr[2 * M] / rr[j];
        xj = __shfl(xj, p);
        if (lane == 0) sh.drow[j] = xj;
        if (rdone && lane != p) rr[2 * M] -= rr[j] * xj;
    }
    __syncthreads();
    bool ok = true;
    if (lane < 2 * M) ok = (sh.drow[lane] <= 1.7976931348623157e308 && sh.drow[lane] >= -1.7976931348623157e308);
    if (lane < M) {
        dout[lane] = sh.drow[lane];
        dout[(size_t)N * M + lane] = sh.drow[M + lane];
    }
    // ---- backward sweep: d_k = U^-1 (h - X d_0 - Z d_{k+1})
    if (lane < M) { sh.va[lane] = sh.drow[lane]; sh.vb[lane] = sh.drow[M + lane]; }
    __syncthreads();
    for (int k = N - 1; k >= 1; --k) {
        const LT *R = rec + (size_t)k * BRec<M>::SIZE;
        double v = 0.0;
        if (lane < M) {
            const int j = lane;
            v = (double)R[BRec<M>::H + j];
#pragma unroll 8
            for (int c = 0; c < M; ++c)
                v -= (double)R[BRec<M>::X + (size_t)c * M + j] * sh.va[c] + (double)R[BRec<M>::Z + (size_t)c * M + j] * sh.vb[c];
        }
        // triangular solve U d = v, U_j in pivot order: U[j][c] = R[U + c*M + j]
        double dk = 0.0;
#pragma unroll
        for (int j = M - 1; j >= 0; --j) {
            double dj = 0.0;
            if (lane == j) dj = v / (double)R[BRec<M>::U + (size_t)j * M + j];
            dj = __shfl(dj, j);
            if (lane == j) dk = dj;
            if (lane < j) v -= (double)R[BRec<M>::U + (size_t)j * M + lane] * dj;
        }
        __syncthreads();
        if (lane < M) {
            dout[(size_t)k * M + lane] = dk;
            sh.vb[lane] = dk;
            ok &= (dk <= 1.7976931348623157e308 && dk >= -1.7976931348623157e308);
        }
        __syncthreads();
    }
    return !wave_any(!ok);
}

// r = -F(y) - J(y) d  (FP64), point-major residual layout as used by bsolve
template <int M, int METHOD, typename LT>
__device__ void brefine_residual(const BInst<M> &I, BShared<M, METHOD, LT> &sh, const double *mesh, const double *y,
                                 const double *d, int N, int lane, double *rout) {
    for (int k = 0; k < N; ++k) {
        const double h = mesh[k + 1] - mesh[k];
        if (lane < M) {
            sh.va[lane] = y[(size_t)k * M + lane];
            sh.vb[lane] = y[(size_t)(k + 1) * M + lane];
        }
        __syncthreads();
        bstages<M, METHOD, LT>(sh, lane, h);
        double da = 0.0, db = 0.0, ph = 0.0;
        if (lane < M) {
            da = d[(size_t)k * M + lane];
            db = d[(size_t)(k + 1) * M + lane];
            ph = bphi<M, METHOD, LT>(sh, lane, h);
        }
        const double jd = bphi_tangent<M, METHOD, LT>(sh, lane, h, da, db);
        if (lane < M) rout[(size_t)k * M + lane] = -ph - jd;
        __syncthreads();
    }
    if (lane < M) {
        const int node = sel<M>(I.bc_node_is_end, lane) ? N : 0;
        const int comp = sel<M>(I.bcc, lane);
        const double v = y[(size_t)node * M + comp] - sel<M>(I.bcv, lane);
        rout[(size_t)N * M + lane] = -v - d[(size_t)node * M + comp];
    }
    __syncthreads();
}

// ------------------------------------------------------------------ Newton (block)
template <int M, int METHOD, typename LT>
__device__ NewtonOut bnewton(const BatchDev &B, const BInst<M> &I, BShared<M, METHOD, LT> &sh, const double *mesh,
                             double *y, double *ybak, double *delta, double *rbuf, double *ytmp, int N, int lane,
                             LT *rec) {
    NewtonOut out{BVP_RET_MAXITERS, 0, 0};
    const int attempts = B.linesearch ? 2 : 1;
    const bool fp32 = sizeof(LT) == 4;
    const int refine = fp32 ? (B.refine_iters > 0 ? B.refine_iters : 0) : 0;
    if (B.linesearch) {
        for (int e = lane; e < (N + 1) * M; e += WAVE) ybak[e] = y[e];
        __syncthreads();
    }
    for (int attempt = 0; attempt < attempts; ++attempt) {
        if (attempt == 1) {
            for (int e = lane; e < (N + 1) * M; e += WAVE) y[e] = ybak[e];
            __syncthreads();
        }
        int ret = BVP_RET_MAXITERS;
        for (int it = 0;; ++it) {
            double ss;
            const double fn = bresidual<M, METHOD, LT>(I, sh, mesh, y, nullptr, 0.0, N, lane, &ss);
            if (!(fn <= 1.7976931348623157e308)) { ret = BVP_RET_UNSTABLE; break; }
            if (fn <= B.newton_tol) { ret = BVP_RET_SUCCESS; break; }
            if (it >= B.max_newton) { ret = BVP_RET_MAXITERS; break; }
            bool ok = bsolve<M, METHOD, LT>(I, sh, mesh, y, N, lane, rec, true, nullptr, delta);
            for (int rI = 0; ok && rI < refine; ++rI) {
                // FP64 iterative refinement: r = -F - J d, solve J e = r with the FP32 factors, d += e
                brefine_residual<M, METHOD, LT>(I, sh, mesh, y, delta, N, lane, rbuf);
                ok = bsolve<M, METHOD, LT>(I, sh, mesh, y, N, lane, rec, false, rbuf, ytmp);
                for (int e = lane; e < (N + 1) * M; e += WAVE) delta[e] += ytmp[e];
                __syncthreads();
            }
            if (!ok) { ret = BVP_RET_FAILURE; break; }
            out.iters++;
            double alpha = 1.0;
            if (attempt == 1) {
                for (int ls = 0; ls < 12; ++ls) {
                    double ph;
                    bresidual<M, METHOD, LT>(I, sh, mesh, y, delta, alpha, N, lane, &ph);
                    if (ph <= (1.0 - 2e-4 * alpha) * ss) break;
                    alpha *= 0.5;
                }
            }
            for (int e = lane; e < (N + 1) * M; e += WAVE) y[e] += alpha * delta[e];
            __syncthreads();
        }
        out.ret = ret;
        out.fallback = attempt;
        if (ret == BVP_RET_SUCCESS) break;
    }
    return out;
}

// ------------------------------------------------------------------ defect (block)
template <int M, int METHOD, typename LT>
__device__ double bdefect(BShared<M, METHOD, LT> &sh, const double *mesh, const double *y, int N, int lane,
                          double abstol, double *shat, double *errors, double &r1, double &r2) {
    using TB = Tab<METHOD>;
    double w1[TB::sstar], w1p[TB::sstar], w2[TB::sstar], w2p[TB::sstar];
    interp_weights<METHOD>(TB::tau(), w1, w1p);
    interp_weights<METHOD>(1.0 - TB::tau(), w2, w2p);
    const double ex = 1.0 / (TB::order + 1);
    double dmax = 0.0, smax = 0.0, ssum = 0.0;
    bool bad = false;
    for (int k = 0; k < N; ++k) {
        const double ti = mesh[k], h = mesh[k + 1] - ti;
        if (lane < M) {
            sh.va[lane] = y[(size_t)k * M + lane];
            sh.vb[lane] = y[(size_t)(k + 1) * M + lane];
        }
        __syncthreads();
        bstages<M, METHOD, LT>(sh, lane, h);
        binterp_stages<M, METHOD, LT>(sh, lane, h);
        double d[2];
        double e[2];
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const double *w = side ? w2 : w1;
            const double *wp = side ? w2p : w1p;
            double zp = 0.0;
            if (lane < M) {
                double a = 0.0;
#pragma unroll
                for (int r = 0; r < TB::s; ++r) { a += sh.K[r][lane] * w[r]; zp += sh.K[r][lane] * wp[r]; }
#pragma unroll
                for (int r = 0; r < TB::sstar - TB::s; ++r) {
                    a += sh.KI[r][lane] * w[TB::s + r];
                    zp += sh.KI[r][lane] * wp[TB::s + r];
                }
                sh.vd[lane] = a * h + sh.va[lane];
            }
            __syncthreads();
            double dd = 0.0;
            if (lane < M) {
                const double f = bf_of_vd<M, METHOD, LT>(sh, lane);
                dd = (zp - f) / (fabs(f) + 1.0);
            }
            __syncthreads();
            d[side] = dd;
            e[side] = wave_max(fabs(dd));
        }
        const bool first = e[0] > e[1];
        const double dv = first ? d[0] : d[1];
        const double ei = first ? e[0] : e[1];
        if (errors && lane < M) errors[(size_t)k * M + lane] = dv;
        bad |= !(fabs(dv) <= 1.7976931348623157e308);
        dmax = fmax(dmax, ei);
        const double shv = pow(ei / abstol, ex);
        if (shat && lane == 0) shat[k] = shv;
        smax = fmax(smax, shv);
        ssum += shv;
    }
    __syncthreads();
    r1 = smax;
    r2 = ssum;  // identical in all lanes (computed from wave-uniform values)
    return wave_any(bad) ? NAN : dmax;
}

// Transfer of the old continuous extension to the new nodes, node by node
// (MIRK/mirk.jl:366-370); REFERENCE mode reads the base value y0.u[ii] from the
// array being written, as the reference's in-place update does.
template <int M, int METHOD, typename LT>
__device__ void btransfer(BShared<M, METHOD, LT> &sh, const double *mesh, const double *y, int N, const double *nm,
                          double *ny, int Nn, int mode, int lane) {
    using TB = Tab<METHOD>;
    for (int j = 0; j <= Nn; ++j) {
        const double m = nm[j];
        const int ii = interval_of(mesh, N, m);
        const double ti = mesh[ii], dt = mesh[ii + 1] - ti;
        const double tau = (m - ti) / dt;
        if (lane < M) {
            sh.va[lane] = y[(size_t)ii * M + lane];
            sh.vb[lane] = y[(size_t)(ii + 1) * M + lane];
        }
        __syncthreads();
        bstages<M, METHOD, LT>(sh, lane, dt);
        binterp_stages<M, METHOD, LT>(sh, lane, dt);
        double w[TB::sstar];
        interp_weights<METHOD>(tau, w, nullptr);
        if (lane < M) {
            double a = 0.0;
#pragma unroll
            for (int r = 0; r < TB::s; ++r) a += sh.K[r][lane] * w[r];
#pragma unroll
            for (int r = 0; r < TB::sstar - TB::s; ++r) a += sh.KI[r][lane] * w[TB::s + r];
            const double base = (mode == BVP_TRANSFER_REFERENCE && ii < j) ? ny[(size_t)ii * M + lane] : sh.va[lane];
            ny[(size_t)j * M + lane] = a * dt + base;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ full solve (block)
template <int M, int METHOD, typename LT>
__device__ void bsolve_instance(const BatchDev &B, int b, BShared<M, METHOD, LT> &sh, int lane, LT *rec, double *aux) {
    BInst<M> I;
    bload_inst<M, METHOD, LT>(B, b, sh, I, lane);
    double *mesh_b[2], *y_b[2];
    {
        const size_t cap = (size_t)B.cap;
        mesh_b[0] = B.mesh + (size_t)b * 2 * cap;
        mesh_b[1] = mesh_b[0] + cap;
        y_b[0] = B.y + (size_t)b * 2 * cap * M;
        y_b[1] = y_b[0] + cap * M;
    }
    int cur = 0;
    int N = B.N0[b];
    {
        const long long off = B.init_off[b];
        for (int k = lane; k <= N; k += WAVE) mesh_b[0][k] = B.init_mesh[off + k];
        for (long long e = lane; e < (long long)(N + 1) * M; e += WAVE) y_b[0][e] = B.init_y[off * M + e];
        __syncthreads();
    }
    double *shat = B.shat + (size_t)b * B.cap;
    double *delta = aux, *rbuf = aux + (size_t)B.cap * M, *ytmp = aux + (size_t)2 * B.cap * M;
    int info = BVP_RET_SUCCESS, nl_ret = BVP_RET_SUCCESS;
    int outer = 0, iters = 0, fb = 0;
    double last_defect = NAN;
    long long cnt_it = 0, cnt_def = 0, cnt_tr = 0, cnt_chk = 0;
    for (;;) {
        ++outer;
        double *mesh = mesh_b[cur], *y = y_b[cur];
        double *mesh_o = mesh_b[1 - cur], *y_o = y_b[1 - cur];
        NewtonOut nl;
        if (N > B.rec_cap) {
            nl.ret = BVP_RET_FAILURE;
            nl.iters = 0;
            nl.fallback = 0;
            fb |= 8;  // records of this instance do not fit the per-wave factor storage
        } else {
            nl = bnewton<M, METHOD, LT>(B, I, sh, mesh, y, y_o, delta, rbuf, ytmp, N, lane, rec);
        }
        iters += nl.iters;
        cnt_it += (long long)N * nl.iters;
        cnt_chk += N;
        if (nl.fallback) fb |= 1;
        nl_ret = nl.ret;
        double error_norm = 2.0 * B.abstol;
        if (!B.adaptive || (fb & 8)) { info = nl.ret; break; }
        info = nl.ret;
        double r1 = 0.0, r2 = 0.0;
        if (info == BVP_RET_SUCCESS) {
            error_norm = bdefect<M, METHOD, LT>(sh, mesh, y, N, lane, B.abstol, shat, nullptr, r1, r2);
            cnt_def += N;
            last_defect = error_norm;
            info = (error_norm > B.defect_threshold || isnan(error_norm)) ? BVP_RET_FAILURE : BVP_RET_SUCCESS;
        }
        __syncthreads();
        if (info == BVP_RET_SUCCESS) {
            if (error_norm > B.abstol) {
                const int Nn = mesh_select(mesh, N, shat, r1, r2, B.max_sub, mesh_o, lane, sh);
                if (Nn < 0) {
                    info = BVP_RET_FAILURE;
                } else {
                    btransfer<M, METHOD, LT>(sh, mesh, y, N, mesh_o, y_o, Nn, B.transfer_mode, lane);
                    cnt_tr += N + Nn;
                    N = Nn;
                    cur = 1 - cur;
                }
            }
        } else {
            if (2 * N > B.max_sub) {
                info = BVP_RET_FAILURE;
            } else {
                half_mesh_into(mesh, N, mesh_o, lane);
                for (long long e = lane; e < (long long)(2 * N + 1) * M; e += WAVE) y_o[e] = 0.0;
                __syncthreads();
                N = 2 * N;
                cur = 1 - cur;
                info = BVP_RET_SUCCESS;
            }
        }
        if (!(info == BVP_RET_SUCCESS && error_norm > B.abstol)) break;
        if (outer >= MAX_OUTER) { info = BVP_RET_FAILURE; break; }
    }
    double ss;
    const double rn = bresidual<M, METHOD, LT>(I, sh, mesh_b[cur], y_b[cur], nullptr, 0.0, N, lane, &ss);
    if (lane == 0) {
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
        B.cnt[(size_t)b * 4 + 3] = cnt_chk;
    }
    __syncthreads();
}

template <int M, int METHOD, typename LT>
__global__ __launch_bounds__(64) void k_bsolve(BatchDev B) {
    __shared__ BShared<M, METHOD, LT> sh;
    const int lane = threadIdx.x;
    LT *rec = (LT *)B.scratch + (size_t)blockIdx.x * B.scratch_per_slot;
    double *aux = B.aux + (size_t)blockIdx.x * B.aux_per_slot;
    for (;;) {
        if (lane == 0) sh.item = (int)atomicAdd(B.counter, 1u);
        __syncthreads();
        const int item = sh.item;
        __syncthreads();
        if (item >= B.n) break;
        bsolve_instance<M, METHOD, LT>(B, B.order[item], sh, lane, rec, aux);
    }
}

// Component operations on a given state (parity entry points).
template <int M, int METHOD, typename LT>
__global__ __launch_bounds__(64) void k_bcomponent(BatchDev B, int op) {
    __shared__ BShared<M, METHOD, LT> sh;
    const int lane = threadIdx.x;
    LT *rec = (LT *)B.scratch + (size_t)blockIdx.x * B.scratch_per_slot;
    double *aux = B.aux + (size_t)blockIdx.x * B.aux_per_slot;
    for (;;) {
        if (lane == 0) sh.item = (int)atomicAdd(B.counter, 1u);
        __syncthreads();
        const int b = sh.item;
        __syncthreads();
        if (b >= B.n) break;
        BInst<M> I;
        bload_inst<M, METHOD, LT>(B, b, sh, I, lane);
        const int N = B.N0[b];
        const long long off = B.init_off[b];
        double *mesh = B.mesh + (size_t)b * 2 * B.cap;
        double *y = B.y + (size_t)b * 2 * B.cap * M;
        for (int k = lane; k <= N; k += WAVE) mesh[k] = B.init_mesh[off + k];
        for (long long e = lane; e < (long long)(N + 1) * M; e += WAVE) y[e] = B.init_y[off * M + e];
        __syncthreads();
        if (op == OP_RESIDUAL) {
            // two-point layout [bca; Phi; bcb] (rows at t0 first, rows at t1 last)
            double *out = B.out0 + off * M;
            int La = 0;
#pragma unroll
            for (int k = 0; k < M; ++k) La += I.bc_node_is_end[k] ? 0 : 1;
            double ss;
            bresidual<M, METHOD, LT>(I, sh, mesh, y, nullptr, 0.0, N, lane, &ss, out + La);
            if (lane == 0) {
                int ia = 0, ib = 0;
                for (int k = 0; k < M; ++k) {
                    if (I.bc_node_is_end[k]) out[La + (size_t)N * M + (ib++)] = sh.bcres[k];
                    else out[ia++] = sh.bcres[k];
                }
            }
        } else if (op == OP_JACOBIAN) {
            const long long ioff = off - b;
            for (int k = 0; k < N; ++k) {
                const double h = mesh[k + 1] - mesh[k];
                if (lane < M) { sh.va[lane] = y[(size_t)k * M + lane]; sh.vb[lane] = y[(size_t)(k + 1) * M + lane]; }
                __syncthreads();
                bstages<M, METHOD, LT>(sh, lane, h);
                btangent_blocks<M, METHOD, LT>(sh, lane, h);
                double *So = B.out0 + (ioff + k) * M * M, *Ro = B.out1 + (ioff + k) * M * M;
                for (int e = lane; e < M * M; e += WAVE) {
                    const int i = e / M, c = e % M;
                    So[e] = (double)sh.G[i][c];
                    Ro[e] = (double)sh.G[i][M + c];
                }
                __syncthreads();
            }
            if (lane < M) B.iout[(size_t)b * M + lane] = sel<M>(I.bc_node_is_end, lane) ? N : 0;
        } else if (op == OP_NEWTON_STEP) {
            double ss;
            bresidual<M, METHOD, LT>(I, sh, mesh, y, nullptr, 0.0, N, lane, &ss);
            double *d = B.out0 + off * M;
            bool ok = (N <= B.rec_cap) && bsolve<M, METHOD, LT>(I, sh, mesh, y, N, lane, rec, true, nullptr, d);
            const int refine = sizeof(LT) == 4 ? B.refine_iters : 0;
            double *rbuf = aux + (size_t)B.cap * M, *ytmp = aux + (size_t)2 * B.cap * M;
            for (int rI = 0; ok && rI < refine; ++rI) {
                brefine_residual<M, METHOD, LT>(I, sh, mesh, y, d, N, lane, rbuf);
                ok = bsolve<M, METHOD, LT>(I, sh, mesh, y, N, lane, rec, false, rbuf, ytmp);
                for (long long e = lane; e < (long long)(N + 1) * M; e += WAVE) d[e] += ytmp[e];
                __syncthreads();
            }
            if (lane == 0) B.iout[b] = ok ? 0 : 1;
        } else if (op == OP_DEFECT) {
            double r1, r2;
            const long long ioff = off - b;
            const double dn = bdefect<M, METHOD, LT>(sh, mesh, y, N, lane, B.abstol, nullptr, B.out0 + ioff * M, r1, r2);
            if (lane == 0) B.out1[b] = dn;
        } else if (op == OP_MESH_SELECT) {
            double r1, r2;
            double *shat = B.shat + (size_t)b * B.cap;
            bdefect<M, METHOD, LT>(sh, mesh, y, N, lane, B.abstol, shat, nullptr, r1, r2);
            __syncthreads();
            double *nm = mesh + B.cap, *ny = y + (size_t)B.cap * M;
            const int Nn = mesh_select(mesh, N, shat, r1, r2, B.max_sub, nm, lane, sh);
            if (Nn < 0) {
                if (lane == 0) { B.iout[(size_t)b * 2] = N; B.iout[(size_t)b * 2 + 1] = BVP_RET_FAILURE; }
            } else {
                btransfer<M, METHOD, LT>(sh, mesh, y, N, nm, ny, Nn, B.transfer_mode, lane);
                double *om = B.out0 + (size_t)b * B.cap, *oy = B.out1 + (size_t)b * B.cap * M;
                for (int k = lane; k <= Nn; k += WAVE) om[k] = nm[k];
                for (long long e = lane; e < (long long)(Nn + 1) * M; e += WAVE) oy[e] = ny[e];
                if (lane == 0) { B.iout[(size_t)b * 2] = Nn; B.iout[(size_t)b * 2 + 1] = BVP_RET_SUCCESS; }
            }
        }
        __syncthreads();
    }
}

template <int M, int METHOD, typename LT>
hipError_t launch_bsolve_t(const BatchDev &B, int blocks, hipStream_t s) {
    k_bsolve<M, METHOD, LT><<<blocks, 64, 0, s>>>(B);
    return hipGetLastError();
}
template <int M, int METHOD, typename LT>
hipError_t launch_bcomponent_t(const BatchDev &B, int op, int blocks, hipStream_t s) {
    k_bcomponent<M, METHOD, LT><<<blocks, 64, 0, s>>>(B, op);
    return hipGetLastError();
}
template <int M, int METHOD, typename LT>
int boccupancy_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bsolve<M, METHOD, LT>, 64, 0) != hipSuccess) return 2;
    return nb;
}

}  // namespace bvpdev
