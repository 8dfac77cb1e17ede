// mirk_wave.h -- device descriptor of a batch and the launchers of the
// wave-per-instance kernels (mirk_wave.hip), used by the C ABI (bvp_api.hip).
#pragma once

#include <hip/hip_runtime.h>

// operations of the component kernel
enum {
    OP_RESIDUAL = 0,
    OP_JACOBIAN = 1,
    OP_NEWTON_STEP = 2,
    OP_DEFECT = 3,
    OP_MESH_SELECT = 4,
};

// Tableau of a FIRK method (Radau IIA, Lobatto IIIA/B/C): A in Butcher order, b, c,
// Q the coefficients of the stage polynomial (Q[k][j]: tau^(k+1) in int_0^tau l_j),
// tau* the defect sample point, s stages, order.
struct FirkTab {
    double A[7][7], b[7], c[7], Q[7][7], tau;
    int s, order;
};

// Everything a kernel needs about a batch; all pointers are device pointers.
struct BatchDev {
    int n;    // instances
    int M;    // state dimension (shared by the batch)
    int cap;  // node capacity per instance

    // problem description, per instance
    const int *family;
    const double *t0, *t1;
    const int *param_off, *param_cnt;
    const double *params;
    const int *bc_kind;
    const double *bc_t;    // [n][M]
    const int *bc_comp;    // [n][M]
    const double *bc_val;  // [n][M]
    const int *N0;
    const long long *init_off;  // node offset of each instance in the packed inputs
    const double *init_mesh;    // packed [sum (N0+1)]
    const double *init_y;       // packed [sum (N0+1)][M]
    const int *order;           // queue order (largest first)

    // working state, per instance
    double *mesh;  // [n][2][cap]   (double-buffered)
    double *y;     // [n][2][cap][M]
    double *shat;  // [n][cap]

    // per-instance results
    int *retcode, *N_final, *newton_iters, *outer_iters, *fallback, *cur_buf;
    double *defect_norm, *resid_norm, *err;
    long long *cnt;  // [n][4]: interval-iterations, defect intervals, transfer nodes, check passes

    // per-slot (persistent wave) scratch
    double *scratch;
    size_t scratch_per_slot;  // doubles
    size_t tree_base;         // offset of the tree records inside a slot
    double *aux;              // [slot][cap*(3M+1)]: delta, trial state, transfer indices, J'F
                              //   (block kernels: their own layout)
    size_t aux_per_slot;
    unsigned *counter;        // work-queue head

    // component outputs
    double *out0, *out1;
    int *iout;

    // learned warm start: init_y holds the learned guess; the first Newton step on
    // the first mesh must contract max|F|, else the solve restarts from u0c
    int warm;
    const double *u0c;        // [n][M] constant guess of each instance
    const double *anc;        // learned anchors [n][anc_K][4] (M = 2): when set, the initial
    int anc_K;                //   guess is their Hermite interpolant instead of init_y

    int rec_cap;              // block kernels: intervals whose factors fit a wave's storage
    int refine_iters;         // FP64 refinement steps per linear solve with FP32 factors
    int f32rec;               // wave kernels: elimination factors stored in FP32 (mixed precision)
    int min_chunk;            // wave kernels: at least this many intervals per lane's chunk
                              //   (1 ... kMaxMinChunk; chosen by the autotuner)

    // error controller (MIRK/adaptivity.jl:77-243, 463-567): BVP_CONTROL_* and
    // BVP_GLOBAL_*; the global-error solve's buffers, per slot: the halved mesh
    // [2 cap], its state, Newton correction, trial state and J'F [2 cap][M] each
    int controller, global_method;
    double hybrid_de, hybrid_ge;
    double *ge;
    size_t ge_per_slot;

    // FIRK methods (expanded form): the tableau (a FirkTab in device memory), the stage
    // values of every instance [n][cap][s][M], and for the component kernels the packed
    // input stages [sum N][s][M] (interval offset = init_off[b] - b)
    const void *firk_tab;
    double *stK;
    const double *init_K;
    int firk_nested;          // the nested form: u0c holds each instance's inner-solve start k0
    double nest_tol;          // its inner tolerance
    int firk_irows;           // the general FIRK kernels: some STANDARD row lies inside the span
                              // (EvalSol rows), or the TrustRegion stage is asked for

    // options
    int adaptive, max_sub, max_newton, stages, transfer_mode;
    int report_err;           // fill err[] after each solve (one more Newton correction)
    double abstol, newton_tol, defect_threshold;
};

// largest interval-per-lane minimum the autotuner tries (the chunk records of a
// slot are sized for chunks of up to 2 * kMaxMinChunk intervals)
constexpr int kMaxMinChunk = 8;

hipError_t wave_reset_counter(unsigned *counter, hipStream_t s);
// variant v of the persistent solve kernel (register allocation held to v + 1
// waves per SIMD for M = 2; one variant otherwise)
int wave_variants(int M);
// fam: the batch's single family, or -1 for a mixed batch (a kernel specialised
// to one family exists for some (M, method, family))
hipError_t wave_launch_solve(const BatchDev &B, int method, int variant, int blocks, int fam, hipStream_t s);
hipError_t wave_launch_component(const BatchDev &B, int method, int op, int blocks, hipStream_t s);
// the persistent solve with a global-error controller (B.controller != DEFECT;
// generic family code, FP64 records, one variant)
hipError_t wave_launch_solve_ge(const BatchDev &B, int method, int blocks, hipStream_t s);
// the controller's global-error estimate of each instance at its packed state:
// out0 the interval errors [sum N][M], out1 the norm per instance
hipError_t wave_launch_global_error(const BatchDev &B, int method, int blocks, hipStream_t s);
// doubles per slot of the global-error buffers for node capacity cap
inline size_t wave_ge_doubles(int cap, int M) { return (size_t)2 * cap * (4 * M + 1); }
hipError_t wave_launch_gather(const BatchDev &B, const long long *out_off, double *out_mesh, double *out_y,
                              int max_nodes, hipStream_t s);
// dense output at nq points: qinst[q] instance of point q, tq[q] its t; u (and du
// when non-null) receive M values per point
hipError_t wave_launch_interp(const BatchDev &B, int method, const int *qinst, const double *tq, double *u,
                              double *du, long long nq, hipStream_t s);
bool wave_supports(int M, int family);
int wave_rel_fields(int M);

// FIRK kernels (firk_wave_impl.h): M = 1..4, S = 1, 2, 3, 4, 5, 7 stages
bool firk_supports(int M, int S);
// variant v: register allocation held to v + 1 waves per SIMD (v = 0, 1)
constexpr int kFirkVariants = 2;
hipError_t firk_launch_solve(const BatchDev &B, int S, int variant, int blocks, hipStream_t s);
hipError_t firk_launch_component(const BatchDev &B, int S, int op, int blocks, hipStream_t s);
hipError_t firk_launch_interp(const BatchDev &B, int S, const int *qinst, const double *tq, double *u, double *du,
                              long long nq, hipStream_t s);
hipError_t firk_launch_gather(const BatchDev &B, int S, const long long *out_off, double *out, hipStream_t s);
int firk_blocks_per_cu(int M, int S, int variant, int nested);  // resident solve waves per CU
// doubles per slot of the FIRK work area for node capacity cap
inline size_t firk_aux_doubles(int cap, int M, int S) {
    return (size_t)cap * (2 * M + 2 * S * M + S * M * (M + 1));
}

hipError_t launch_copy_probe(const double *src, double *dst, size_t n, hipStream_t s);

// block kernels (M = 8, 16, 32; LINEAR_TANH): precision 0 = FP64 factors, 1 = FP32 factors + FP64 refinement
bool block_supports(int M, int family);
hipError_t block_launch_solve(const BatchDev &B, int method, int precision, int blocks, hipStream_t s);
hipError_t block_launch_component(const BatchDev &B, int method, int precision, int op, int blocks, hipStream_t s);
int block_blocks_per_cu(int M, int method, int precision);
size_t block_record_values(int M);  // values (of the factor type) per interval record
