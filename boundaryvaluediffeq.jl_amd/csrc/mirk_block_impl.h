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
    static constexpr int W = 3 * M + 1;   // stacked row: pivot block | y_0 block | y_{k+1} block | rhs
    static constexpr int SW = W | 1;      // odd row stride (spreads the rows over the LDS banks)
    static constexpr int RW = 2 * M + 1;  // root row: d_0 block | d_N block | rhs
    double At[M][M];                      // A transposed (At[q][i] = A[i][q]): conflict-free lane reads
    LT Al[sizeof(LT) == 4 ? M : 1][M];    // A in FP32 for the FP32 tangents (the FP64 ones read At)
    double c;
    double va[M], vb[M], vd[M];           // node vectors (state or corrections)
    double Y[S][M], K[S][M], KI[NS][M];   // stage arguments and values
    double Yd[S][M], Kd[S][M];            // their tangents (J v products)
    LT sech2[S][M];
    union {
        LT stk[2 * M][SW];                // stacked rows of one elimination step (rows 0..M-1: relation)
        double root[M >= 16 ? 1 : 2 * M][RW];  // the end-node system (M < 16; larger M: in registers)
    } w;
    double xr[2 * M];
    LT prow[W + 3];                       // pivot row broadcast of the register elimination
    double bcres[M];
    int piv[2 * M];
    int src[M];
    int bc_end[M], bcc[M];
    double bcv[M];
    int flag, item;
};

template <int M, int METHOD, typename LT>
__device__ inline LT aval(const BShared<M, METHOD, LT> &sh, int i, int q) {
    if constexpr (sizeof(LT) == 4) return sh.Al[i][q];
    else return sh.At[q][i];
}

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
// Lane l < 2M computes column l of G = sum_r b_r dK_r, dK_r = J_r dY_r (forward-mode
// derivative of the stage recursion along the unit direction l), and writes
// S = -I - h G[:, :M] / R = I - h G[:, M:] straight into the stacked rows in LDS:
//   first == false: rows M..2M-1 get [S | 0 | R | .]   (a new block row)
//   first == true : rows 0..M-1  get [R | S | 0 | .]   (interval 0 is the first relation)
// The same blocks for the FP32 factorisation at M >= 16 with the stage products
// J_r dY_r = A dY_r + c sech^2(Y_r) o dY_r on the matrix cores: the M x 2M tangent
// matrix D = dY_r is staged in LDS (the rows the new block row will occupy, free
// until the end) and A D runs as 16 x 16 x 4 FP32 MFMA tiles
// (v_mfma_f32_16x16x4_f32).  Every lane keeps its part of [S | R] and of the
// last-stage accumulator in the MFMA output layout -- lane l holds, for each
// 16 x 16 tile, rows 4 (l / 16) .. +3 of column l % 16 -- so the elementwise terms
// (the identity seeds, the first two stages, the sech^2 diagonal) are formed in
// that layout and the blocks are written to the stacked rows once at the end.
template <int M, int METHOD>
__device__ __noinline__ void btangent_blocks_mfma(BShared<M, METHOD, float> &sh, int lane, double hd, bool first) {
    using TB = Tab<METHOD>;
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int RT = M / 16, CT = 2 * M / 16, KS = M / 4;
    const float h = (float)hd;
    const float c = (float)sh.c;
    const int rb = first ? 0 : M;
    const int l16 = lane & 15, l4 = lane >> 4;
    float *D = &sh.w.stk[M][0];  // D[q][col] at D[q * SW + col]
    constexpr int SW = BShared<M, METHOD, float>::SW;
    // A operands: A[16 rt + l16][4 ks + l4]
    float aop[RT][KS];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) aop[rt][ks] = sh.Al[16 * rt + l16][4 * ks + l4];
    float G[RT][CT][4], q4[RT][CT][4];
    // first two stages: G = +-I - h b_r0 J_r0 (r0 = 0 for the y_k columns, 1 for y_{k+1})
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int col = 16 * ct + l16;
            const bool left = col < M;
            const int jc = left ? col : col - M;
            const int r0 = left ? 0 : 1;
            const float sj = c * sh.sech2[r0][jc];
            const float hb0 = h * (float)(left ? TB::b(0) : TB::b(1));
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 16 * rt + 4 * l4 + v;
                const float jv = sh.Al[i][jc] + ((i == jc) ? sj : 0.f);
                const float e = (i == jc) ? (left ? -1.f : 1.f) : 0.f;
                G[rt][ct][v] = e - hb0 * jv;
                q4[rt][ct][v] = 0.f;
            }
        }
#pragma unroll
    for (int r = 2; r < TB::s; ++r) {
        // dY_r in the output layout, into LDS
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const int col = 16 * ct + l16;
                const bool left = col < M;
                const int jc = left ? col : col - M;
                const int r0 = left ? 0 : 1;
                const float sj = c * sh.sech2[r0][jc];
                const float seed = left ? (float)(1.0 - TB::v(r)) : (float)TB::v(r);
                const float xr = left ? (float)TB::x(r, 0) : (float)TB::x(r, 1);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 16 * rt + 4 * l4 + v;
                    const float jv = sh.Al[i][jc] + ((i == jc) ? sj : 0.f);
                    float d = ((i == jc) ? seed : 0.f) + h * xr * jv;
                    if (r == TB::s - 1 && TB::s > 3) d += h * q4[rt][ct][v];
                    D[i * SW + col] = d;
                }
            }
        __syncthreads();
        // A dY_r on the matrix cores
        f4 acc[RT][CT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            float bop[CT];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) bop[ct] = D[(4 * ks + l4) * SW + 16 * ct + l16];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop[rt][ks], bop[ct], acc[rt][ct], 0, 0, 0);
        }
        const float hb = h * (float)TB::b(r);
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 16 * rt + 4 * l4 + v;
                    const float jd = acc[rt][ct][v] + c * sh.sech2[r][i] * D[i * SW + 16 * ct + l16];
                    G[rt][ct][v] -= hb * jd;
                    if (TB::s > 3 && r < TB::s - 1) q4[rt][ct][v] += (float)TB::x(TB::s - 1, r) * jd;
                }
        __syncthreads();  // D is read by every lane before the next stage overwrites it
    }
    // [S | R] into the stacked rows (target column as in btangent_blocks); the left
    // lanes' middle block is cleared
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int col = 16 * ct + l16;
            const bool left = col < M;
            const int jc = left ? col : col - M;
            const int tc = first ? (left ? M + jc : jc) : (left ? jc : 2 * M + jc);
            const int zc = first ? 2 * M + jc : M + jc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 16 * rt + 4 * l4 + v;
                sh.w.stk[rb + i][tc] = G[rt][ct][v];
                if (left) sh.w.stk[rb + i][zc] = 0.f;
            }
        }
    __syncthreads();
}

template <int M, int METHOD, typename LT>
__device__ inline void btangent_blocks(BShared<M, METHOD, LT> &sh, int lane, double hd, bool first) {
    if constexpr (sizeof(LT) == 4 && M >= 16) {
        btangent_blocks_mfma<M, METHOD>(sh, lane, hd, first);
        return;
    }
    using TB = Tab<METHOD>;
    using SHT = BShared<M, METHOD, LT>;
    const LT h = (LT)hd;
    const bool left = lane < M;
    const int col = left ? lane : lane - M;
    const int rb = first ? 0 : M;
    const int tc = first ? (left ? M + col : col) : (left ? col : 2 * M + col);  // target column
    const int zc = first ? 2 * M + col : M + col;                               // column to clear (left lanes)
    if (lane < 2 * M) {
        const int r0 = left ? 0 : 1;
        const LT sj = (LT)sh.c * sh.sech2[r0][col];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const LT jv = aval<M, METHOD, LT>(sh, i, col) + ((i == col) ? sj : (LT)0);
            const LT e = (i == col) ? (left ? (LT)-1 : (LT)1) : (LT)0;
            sh.w.stk[rb + i][tc] = e - h * (LT)TB::b(r0) * jv;
            if (left) sh.w.stk[rb + i][zc] = (LT)0;
        }
    }
    if (TB::s > 2 && lane < 2 * M) {
        LT q4[M];  // sum_{j>=2} x_{s-1,j} dK_j for the last stage (MIRK6)
#pragma unroll
        for (int i = 0; i < M; ++i) q4[i] = (LT)0;
        const int r0 = left ? 0 : 1;
        const LT sj0 = (LT)sh.c * sh.sech2[r0][col];
#pragma unroll
        for (int r = 2; r < TB::s; ++r) {
            LT dy[M];
            const LT seed = left ? (LT)(1.0 - TB::v(r)) : (LT)TB::v(r);
            const LT xr = left ? (LT)TB::x(r, 0) : (LT)TB::x(r, 1);
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const LT jv = aval<M, METHOD, LT>(sh, i, col) + ((i == col) ? sj0 : (LT)0);
                dy[i] = ((i == col) ? seed : (LT)0) + h * xr * jv;
                if (r == TB::s - 1 && TB::s > 3) dy[i] += h * q4[i];
            }
            const LT hb = h * (LT)TB::b(r);
#pragma unroll
            for (int i = 0; i < M; ++i) {
                LT acc = (LT)sh.c * sh.sech2[r][i] * dy[i];
#pragma unroll
                for (int qq = 0; qq < M; ++qq) acc += aval<M, METHOD, LT>(sh, i, qq) * dy[qq];
                sh.w.stk[rb + i][tc] -= hb * acc;
                if (TB::s > 3 && r < TB::s - 1) q4[i] += (LT)TB::x(TB::s - 1, r) * acc;
            }
        }
    }
    __syncthreads();
    (void)sizeof(SHT);
}

// ------------------------------------------------------------------ elimination in LDS
// Gaussian elimination with partial pivoting on the first NC columns of an
// NR-row system stored row-major in LDS (row stride ST); lane l owns row l and
// updates it in place; the pivot row of column j (lane piv[j]) is read by all
// lanes and not swapped.  When `mult` is non-null the multiplier used on this
// lane for column j is stored at mult[j * 64].  Returns whether this lane's row
// became a pivot row.
template <int NC, int NR, int NWID, int ST, typename T>
__device__ inline bool eliminate_lds(T *mat, int lane, int *piv, T *mult) {
    bool done = lane >= NR;
    T *rowl = mat + (size_t)(lane < NR ? lane : 0) * ST;
    for (int j = 0; j < NC; ++j) {
        double v = done ? -1.0 : (double)fabs((double)rowl[j]);
        int who = lane;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(v, o);
            const int ow = __shfl_xor(who, o);
            if (ov > v || (ov == v && ow < who)) { v = ov; who = ow; }
        }
        const int p = who;
        if (lane == 0) piv[j] = p;
        const T *rowp = mat + (size_t)p * ST;
        const T pv = rowp[j];
        if (lane == p) {
            done = true;
            if (mult) mult[(size_t)j * WAVE] = (T)0;
        } else if (!done) {
            const T m = rowl[j] / pv;
            rowl[j] = (T)0;
            // fixed trip count, predicated: unrolls completely so the LDS reads pipeline
#pragma unroll
            for (int c = 1; c < NWID; ++c)
                if (c > j) rowl[c] -= m * rowp[c];
            if (mult) mult[(size_t)j * WAVE] = m;
        } else if (mult) {
            mult[(size_t)j * WAVE] = (T)0;
        }
        __syncthreads();
    }
    return done && lane < NR;
}

// ------------------------------------------------------------------ elimination in registers
// Lane of the first maximum of v over the wave: an inclusive max-scan inside
// each row of 16 lanes and across rows with DPP moves (no LDS round trips), the
// maximum read from lane 63, the first lane holding it found by a ballot.
__device__ inline int wave_argmax_first(float v) {
    float m = v;
#define BVP_DPP_MAX(ctrl, rmask) \
    m = fmaxf(m, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(m), __float_as_int(m), ctrl, rmask, 0xf, false)))
    BVP_DPP_MAX(0x111, 0xf);  // row_shr:1
    BVP_DPP_MAX(0x112, 0xf);  // row_shr:2
    BVP_DPP_MAX(0x114, 0xf);  // row_shr:4
    BVP_DPP_MAX(0x118, 0xf);  // row_shr:8
    BVP_DPP_MAX(0x142, 0xa);  // row_bcast:15 into rows 1 and 3
    BVP_DPP_MAX(0x143, 0xc);  // row_bcast:31 into rows 2 and 3
#undef BVP_DPP_MAX
    const float mx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), 63));
    const unsigned long long b = __ballot(v == mx);
    return b ? __ffsll((long long)b) - 1 : 0;
}

template <typename T> __device__ inline T rdlane(T v, int l);
template <> __device__ inline float rdlane<float>(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <> __device__ inline double rdlane<double>(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The same elimination as eliminate_lds with row l of the NR x NW system held
// by lane l in registers.  The step loop over j is not unrolled, so every
// access to row[] uses a compile-time column index (selects and predicated
// updates instead of row[j]), which keeps the row in registers.  At step j the
// pivot lane (piv[j], the first maximum of |row[j]| over the rows not yet used,
// found with DPP moves) is read column by column with v_readlane (a scalar
// broadcast with no memory round trip).  Pivot rows are not swapped.  Returns the step at which this
// lane's row became a pivot row, or -1.
template <int NC, int NR, int NW, typename T, bool STORE_MULT = true>
__device__ inline int eliminate_reg(T (&row)[NW], int lane, int (&piv)[NR], T *mult, T *prow) {
    bool done = lane >= NR;
    int step = -1;
#pragma unroll 1
    for (int j = 0; j < NC; ++j) {
        T rj = (T)0;
#pragma unroll
        for (int c = 0; c < NC; ++c) rj = (c == j) ? row[c] : rj;
        // pivot: first maximum of |row[j]| over the rows not yet used (compared in FP32)
        const int p = __builtin_amdgcn_readfirstlane(wave_argmax_first(done ? -1.f : (float)fabs(rj)));
        if (lane == 0) piv[j] = p;
        const T pv = rdlane(rj, p);
        const bool act = !done && lane != p;
        const T m = act ? rj / pv : (T)0;
        if constexpr (STORE_MULT) mult[(size_t)j * WAVE] = m;
        if (lane == p) { done = true; step = j; }
#pragma unroll
        for (int c = 0; c < NW; ++c) {
            const T mc = (c > j) ? m : (T)0;
            T r = row[c] - mc * rdlane(row[c], p);
            if (c < NC) r = (c == j && act) ? (T)0 : r;
            row[c] = r;
        }
    }
    return step;
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
template <int M> struct BInst {  // boundary rows; the arrays live in LDS (BShared)
    const int *bc_node_is_end;   // 0: row at node 0, 1: row at node N
    const int *bcc;
    const double *bcv;
};

template <int M, int METHOD, typename LT>
__device__ inline void bload_inst(const BatchDev &B, int b, BShared<M, METHOD, LT> &sh, BInst<M> &I, int lane) {
    const int po = B.param_off[b];
    const double *A = B.params + po + 1;
    for (int e = lane; e < M * M; e += WAVE) {
        const int i = e / M, q = e % M;
        sh.At[q][i] = A[e];
        if constexpr (sizeof(LT) == 4) sh.Al[i][q] = (LT)A[e];
    }
    if (lane == 0) sh.c = B.params[po];
    const double t0 = B.t0[b];
    if (lane < M) {
        sh.bc_end[lane] = (B.bc_t[(size_t)b * M + lane] == t0) ? 0 : 1;
        sh.bcc[lane] = B.bc_comp[(size_t)b * M + lane];
        sh.bcv[lane] = B.bc_val[(size_t)b * M + lane];
    }
    I.bc_node_is_end = sh.bc_end;
    I.bcc = sh.bcc;
    I.bcv = sh.bcv;
    __syncthreads();
}

// The phases below (residual, factor/solve, refinement residual, defect,
// transfer) are kept out of line: each gets its own register allocation instead
// of one allocation for their union, which is what spilled when they were inlined.
//
// FP64 residual of all intervals at y (+ alpha*delta if delta): returns max|F|
// (NaN if not finite) and, when ss != null, the sum of squares.  BC residuals
// are left in sh.bcres.
template <int M, int METHOD, typename LT>
__device__ __noinline__ double bresidual(const BInst<M> &I, BShared<M, METHOD, LT> &sh, const double *mesh, const double *y,
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
        const int node = I.bc_node_is_end[lane] ? N : 0;
        const int comp = I.bcc[lane];
        double v = y[(size_t)node * M + comp];
        if (delta) v += alpha * delta[(size_t)node * M + comp];
        const double r = v - I.bcv[lane];
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
// J d = rhs where rhs = -F(y) (FP64, computed on the fly) when factor == true,
// or rhs taken from `rin` (point-major residual layout: Phi_k at k*M, BC rows at
// N*M + k) when factor == false (a refinement solve through the stored factors).
// The correction is written to dout (point-major, M*(N+1)).
template <int M, int METHOD, typename LT>
__device__ __noinline__ bool bsolve(const BInst<M> &I, BShared<M, METHOD, LT> &sh, const double *mesh, const double *y, int N,
                       int lane, LT *rec, bool factor, const double *rin, double *dout) {
    using SHT = BShared<M, METHOD, LT>;
    constexpr int W = SHT::W, SW = SHT::SW, RW = SHT::RW;
    LT *stk = &sh.w.stk[0][0];
    // ---- forward sweep
    for (int k = 0; k < N; ++k) {
        const double h = mesh[k + 1] - mesh[k];
        LT *R = rec + (size_t)k * BRec<M>::SIZE;
        const int rb = (k == 0) ? 0 : M;  // rows receiving this interval's block row
        if (factor) {
            if (lane < M) {
                sh.va[lane] = y[(size_t)k * M + lane];
                sh.vb[lane] = y[(size_t)(k + 1) * M + lane];
            }
            __syncthreads();
            bstages<M, METHOD, LT>(sh, lane, h);
            double rk = 0.0;
            if (lane < M) rk = -bphi<M, METHOD, LT>(sh, lane, h);
            btangent_blocks<M, METHOD, LT>(sh, lane, h, k == 0);
            if (lane < M) sh.w.stk[rb + lane][3 * M] = (LT)rk;
            __syncthreads();
            if (k == 0) continue;
            if constexpr (M >= 16) {
                // the stacked rows go to registers (lane l: row l) and are eliminated there
                LT row[W];
    #pragma unroll
                for (int c = 0; c < W; ++c) row[c] = (lane < 2 * M) ? sh.w.stk[lane][c] : (LT)0;
                __syncthreads();
                const int pstep = eliminate_reg<M, 2 * M, W, LT>(row, lane, sh.piv, R + BRec<M>::MULT + lane, sh.prow);
                (void)stk;
                // records of the pivot rows, in pivot order: U | X | Z | h
                if (pstep >= 0) {
    #pragma unroll
                    for (int c = 0; c < M; ++c) {
                        R[BRec<M>::U + (size_t)c * M + pstep] = row[c];
                        R[BRec<M>::X + (size_t)c * M + pstep] = row[M + c];
                        R[BRec<M>::Z + (size_t)c * M + pstep] = row[2 * M + c];
                    }
                    R[BRec<M>::H + pstep] = row[3 * M];
                }
                // the remaining rows become the next relation [F' | E' | 0 | g'] in rows 0..M-1
                const bool keep = pstep < 0 && lane < 2 * M;
                const unsigned long long rest = __ballot(keep);
                __syncthreads();
                if (lane < M) R[BRec<M>::PIV + lane] = (LT)sh.piv[lane];
                if (keep) {
                    const int to = __popcll(rest & ((1ull << lane) - 1ull));
    #pragma unroll
                    for (int c = 0; c < M; ++c) {
                        sh.w.stk[to][c] = row[2 * M + c];
                        sh.w.stk[to][M + c] = row[M + c];
                        sh.w.stk[to][2 * M + c] = (LT)0;
                    }
                    sh.w.stk[to][3 * M] = row[3 * M];
                }
                __syncthreads();
            } else {
                const bool piv_row = eliminate_lds<M, 2 * M, W, SW, LT>(stk, lane, sh.piv, R + BRec<M>::MULT + lane);
                // records of the pivot rows, in pivot order: U | X | Z | h
                if (piv_row) {
                    int j = 0;
                    for (int q = 0; q < M; ++q) j = (sh.piv[q] == lane) ? q : j;
                    const LT *row = stk + (size_t)lane * SW;
                    for (int c = 0; c < M; ++c) {
                        R[BRec<M>::U + (size_t)c * M + j] = row[c];
                        R[BRec<M>::X + (size_t)c * M + j] = row[M + c];
                        R[BRec<M>::Z + (size_t)c * M + j] = row[2 * M + c];
                    }
                    R[BRec<M>::H + j] = row[3 * M];
                }
                if (lane < M) R[BRec<M>::PIV + lane] = (LT)sh.piv[lane];
                // the remaining rows become the next relation [F' | E' | 0 | g'] in rows 0..M-1
                // (each remaining row is read whole into registers before any row is written)
                const unsigned long long rest = __ballot(!piv_row && lane < 2 * M);
                const bool keep = !piv_row && lane < 2 * M;
                LT cp[2 * M + 1];
                if (keep) {
                    const LT *from = stk + (size_t)lane * SW;
    #pragma unroll
                    for (int c = 0; c < M; ++c) { cp[c] = from[2 * M + c]; cp[M + c] = from[M + c]; }
                    cp[2 * M] = from[3 * M];
                }
                __syncthreads();
                if (keep) {
                    LT *to = stk + (size_t)__popcll(rest & ((1ull << lane) - 1ull)) * SW;
    #pragma unroll
                    for (int c = 0; c < 2 * M; ++c) to[c] = cp[c];
    #pragma unroll
                    for (int c = 0; c < M; ++c) to[2 * M + c] = (LT)0;
                    to[3 * M] = cp[2 * M];
                }
                __syncthreads();
            }
        } else {
            // push the new right-hand side through the stored elimination
            if (k == 0) {
                if (lane < M) sh.w.stk[lane][3 * M] = (LT)rin[lane];
                __syncthreads();
                continue;
            }
            LT q = (LT)0;
            if (lane < M) q = sh.w.stk[lane][3 * M];
            else if (lane < 2 * M) q = (LT)rin[(size_t)k * M + lane - M];
            bool dn = lane >= 2 * M;
            LT mj[M];
            int pv[M];
#pragma unroll
            for (int j = 0; j < M; ++j) {
                mj[j] = R[BRec<M>::MULT + (size_t)j * WAVE + lane];
                pv[j] = (int)R[BRec<M>::PIV + j];
            }
#pragma unroll
            for (int j = 0; j < M; ++j) {
                const int pj = pv[j];
                const LT qp = __shfl(q, pj);
                if (lane == pj) { dn = true; R[BRec<M>::H + j] = q; }
                else if (!dn) q -= mj[j] * qp;
            }
            const unsigned long long rest = __ballot(!dn && lane < 2 * M);
            __syncthreads();
            if (!dn && lane < 2 * M) sh.w.stk[__popcll(rest & ((1ull << lane) - 1ull))][3 * M] = q;
            __syncthreads();
        }
    }
    // ---- root: BC rows and the relation E d_0 + F d_N = g (2M x 2M, FP64, in LDS).
    // The factor pass keeps [E | F] of the final relation in record 0 (interval 0
    // has no elimination record) for the refinement passes, since the root
    // system overwrites the stacked rows in LDS.
    LT relrow[2 * M + 1];
    if (lane >= M && lane < 2 * M) {
        const int i = lane - M;
        const LT *row = stk + (size_t)i * SW;  // [F | E | 0 | g]
        if (factor) {
#pragma unroll
            for (int c = 0; c < M; ++c) { relrow[c] = row[M + c]; relrow[M + c] = row[c]; }
#pragma unroll
            for (int c = 0; c < 2 * M; ++c) rec[(size_t)c * M + i] = relrow[c];
        } else {
#pragma unroll
            for (int c = 0; c < 2 * M; ++c) relrow[c] = rec[(size_t)c * M + i];
        }
        relrow[2 * M] = row[3 * M];
    }
    __syncthreads();
    if constexpr (M >= 16) {
        // the 2M x (2M + 1) end-node system with row l in lane l's registers (FP64)
        double rrow[2 * M + 1];
        if (lane < M) {
            const bool end = I.bc_node_is_end[lane] != 0;
            const int comp = I.bcc[lane];
#pragma unroll
            for (int c = 0; c < M; ++c) {
                rrow[c] = (!end && c == comp) ? 1.0 : 0.0;
                rrow[M + c] = (end && c == comp) ? 1.0 : 0.0;
            }
            rrow[2 * M] = factor ? -sh.bcres[lane] : rin[(size_t)N * M + lane];
        } else {
#pragma unroll
            for (int c = 0; c < 2 * M + 1; ++c) rrow[c] = (lane < 2 * M) ? (double)relrow[c] : 0.0;
        }
        eliminate_reg<2 * M, 2 * M, 2 * M + 1, double, false>(rrow, lane, sh.piv, (double *)nullptr, (double *)nullptr);
        __syncthreads();
        // back substitution over the pivot order; x_j goes to LDS and into every row's right-hand side
#pragma unroll 1
        for (int j = 2 * M - 1; j >= 0; --j) {
            const int p = __builtin_amdgcn_readfirstlane(sh.piv[j]);
            double rj = 0.0;
#pragma unroll
            for (int c = 0; c < 2 * M; ++c) rj = (c == j) ? rrow[c] : rj;
            const double xj = rdlane(rrow[2 * M], p) / rdlane(rj, p);
            if (lane == 0) sh.xr[j] = xj;
            rrow[2 * M] -= rj * xj;
        }
        __syncthreads();
    } else {
        if (lane < M) {
            const bool end = I.bc_node_is_end[lane] != 0;
            const int comp = I.bcc[lane];
            for (int c = 0; c < M; ++c) {
                sh.w.root[lane][c] = (!end && c == comp) ? 1.0 : 0.0;
                sh.w.root[lane][M + c] = (end && c == comp) ? 1.0 : 0.0;
            }
            sh.w.root[lane][2 * M] = factor ? -sh.bcres[lane] : rin[(size_t)N * M + lane];
        } else if (lane < 2 * M) {
#pragma unroll
            for (int c = 0; c < 2 * M + 1; ++c) sh.w.root[lane][c] = (double)relrow[c];
        }
        __syncthreads();
        const bool rpiv =
            eliminate_lds<2 * M, 2 * M, RW, RW, double>(&sh.w.root[0][0], lane, sh.piv, (double *)nullptr);
        // back substitution over the pivot order
        for (int j = 2 * M - 1; j >= 0; --j) {
            const int p = sh.piv[j];
            if (lane == p) sh.xr[j] = sh.w.root[p][2 * M] / sh.w.root[p][j];
            __syncthreads();
            const double xj = sh.xr[j];
            if (rpiv && lane != p) sh.w.root[lane][2 * M] -= sh.w.root[lane][j] * xj;
            __syncthreads();
        }
    }
    bool ok = true;
    if (lane < 2 * M) ok = (sh.xr[lane] <= 1.7976931348623157e308 && sh.xr[lane] >= -1.7976931348623157e308);
    if (lane < M) {
        dout[lane] = sh.xr[lane];
        dout[(size_t)N * M + lane] = sh.xr[M + lane];
    }
    // ---- backward sweep: d_k = U^-1 (h - X d_0 - Z d_{k+1})
    if (lane < M) { sh.va[lane] = sh.xr[lane]; sh.vb[lane] = sh.xr[M + lane]; }
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
        // triangular solve U d = v, U_j in pivot order: U[j][c] = R[U + c*M + j];
        // the U block is staged in LDS first (independent loads, one latency)
        LT *Us = stk;
        for (int e = lane; e < M * M; e += WAVE) Us[e] = R[BRec<M>::U + e];
        __syncthreads();
        double dk = 0.0;
        for (int j = M - 1; j >= 0; --j) {
            double dj = 0.0;
            if (lane == j) dj = v / (double)Us[(size_t)j * M + j];
            dj = __shfl(dj, j);
            if (lane == j) dk = dj;
            if (lane < j) v -= (double)Us[(size_t)j * M + lane] * dj;
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
__device__ __noinline__ void brefine_residual(const BInst<M> &I, BShared<M, METHOD, LT> &sh, const double *mesh, const double *y,
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
        const int node = I.bc_node_is_end[lane] ? N : 0;
        const int comp = I.bcc[lane];
        const double v = y[(size_t)node * M + comp] - I.bcv[lane];
        rout[(size_t)N * M + lane] = -v - d[(size_t)node * M + comp];
    }
    __syncthreads();
}

// ------------------------------------------------------------------ Newton (block)
// NewtonRaphson, then NewtonRaphson + BackTracking from the same start (bits 0
// and 1 of B.stages; Core/default_internal_solve.jl:37-43).  The TrustRegion
// stage (bit 2) is run by the wave kernels only.
template <int M, int METHOD, typename LT>
__device__ NewtonOut bnewton(const BatchDev &B, const BInst<M> &I, BShared<M, METHOD, LT> &sh, const double *mesh,
                             double *y, double *ybak, double *delta, double *rbuf, double *ytmp, int N, int lane,
                             LT *rec) {
    NewtonOut out{BVP_RET_MAXITERS, 0, -1, 0, NAN};
    const bool fp32 = sizeof(LT) == 4;
    const int refine = fp32 ? (B.refine_iters > 0 ? B.refine_iters : 0) : 0;
    const int stages = B.stages & 3;
    const bool restart = stages == 3;
    if (restart) {
        for (int e = lane; e < (N + 1) * M; e += WAVE) ybak[e] = y[e];
        __syncthreads();
    }
    bool first = true;
    for (int stage = 0; stage < 2; ++stage) {
        if (!((stages >> stage) & 1)) continue;
        if (!first) {
            for (int e = lane; e < (N + 1) * M; e += WAVE) y[e] = ybak[e];
            __syncthreads();
        }
        first = false;
        out.stage = stage;
        int ret = BVP_RET_MAXITERS;
        for (int it = 0;; ++it) {
            double ss;
            const double fn = bresidual<M, METHOD, LT>(I, sh, mesh, y, nullptr, 0.0, N, lane, &ss);
            out.fn = fn;
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
            if (stage == 1) {
                // BackTracking on |F(y + a d)|^2 / 2 with phi'(0) = -|F|^2 (Newton direction)
                auto phi = [&](double a) {
                    double ph;
                    bresidual<M, METHOD, LT>(I, sh, mesh, y, delta, a, N, lane, &ph);
                    return 0.5 * ph;
                };
                if (!backtracking(phi, 0.5 * ss, -ss, alpha)) { ret = BVP_RET_FAILURE; break; }
            }
            for (int e = lane; e < (N + 1) * M; e += WAVE) y[e] += alpha * delta[e];
            __syncthreads();
        }
        out.ret = ret;
        if (ret == BVP_RET_SUCCESS) break;
    }
    return out;
}

// ------------------------------------------------------------------ defect (block)
template <int M, int METHOD, typename LT>
__device__ __noinline__ double bdefect(BShared<M, METHOD, LT> &sh, const double *mesh, const double *y, int N, int lane,
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
__device__ __noinline__ void btransfer(BShared<M, METHOD, LT> &sh, const double *mesh, const double *y, int N, const double *nm,
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
            nl.stage = -1;
            fb |= 8;  // records of this instance do not fit the per-wave factor storage
        } else {
            nl = bnewton<M, METHOD, LT>(B, I, sh, mesh, y, y_o, delta, rbuf, ytmp, N, lane, rec);
        }
        iters += nl.iters;
        cnt_it += (long long)N * nl.iters;
        cnt_chk += N;
        if (nl.stage == 1) fb |= 1;
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
    // distance of the returned state from the discrete solution on its mesh: one more
    // Newton correction (factors in the plan's precision, refined as in the solve)
    double err = NAN;
    if (B.report_err && N <= B.rec_cap) {
        const double *mesh = mesh_b[cur], *y = y_b[cur];
        bool ok = bsolve<M, METHOD, LT>(I, sh, mesh, y, N, lane, rec, true, nullptr, delta);
        const int refine = sizeof(LT) == 4 ? (B.refine_iters > 0 ? B.refine_iters : 0) : 0;
        for (int rI = 0; ok && rI < refine; ++rI) {
            brefine_residual<M, METHOD, LT>(I, sh, mesh, y, delta, N, lane, rbuf);
            ok = bsolve<M, METHOD, LT>(I, sh, mesh, y, N, lane, rec, false, rbuf, ytmp);
            for (int e = lane; e < (N + 1) * M; e += WAVE) delta[e] += ytmp[e];
            __syncthreads();
        }
        if (ok) {
            double e_ = 0.0;
            for (int e = lane; e < (N + 1) * M; e += WAVE) e_ = fmax(e_, fabs(delta[e]) / fmax(1.0, fabs(y[e])));
            err = wave_max(e_);
        }
        __syncthreads();
    }
    if (lane == 0) {
        B.err[b] = err;
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
                btangent_blocks<M, METHOD, LT>(sh, lane, h, false);
                double *So = B.out0 + (ioff + k) * M * M, *Ro = B.out1 + (ioff + k) * M * M;
                for (int e = lane; e < M * M; e += WAVE) {
                    const int i = e / M, c = e % M;
                    So[e] = (double)sh.w.stk[M + i][c];
                    Ro[e] = (double)sh.w.stk[M + i][2 * M + c];
                }
                __syncthreads();
            }
            if (lane < M) B.iout[(size_t)b * M + lane] = I.bc_node_is_end[lane] ? N : 0;
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
