/*
 * bvp_mirk.h -- C ABI of the MI355X (gfx950) MIRK collocation library.
 *
 * This is the drop-in boundary for the MIRK4/MIRK6 collocation solve that
 * SciML/BoundaryValueDiffEq.jl runs behind
 *     solve(prob::BVProblem, MIRK4()/MIRK6(); dt, abstol)
 * Each entry point replaces one reference function (file:line cited on each
 * declaration; paths are relative to the reference repository root, with
 *   MIRK/ = lib/BoundaryValueDiffEqMIRK/src/  and
 *   Core/ = lib/BoundaryValueDiffEqCore/src/ ).
 *
 * Conventions
 *  - Plain C types only: int32_t, double, pointers and sizes.  No pointer is
 *    retained after a call returns; all host buffers are caller-owned.
 *  - Return value: 0 on success, a negative BVP_E* code on an API/argument or
 *    HIP error (message via bvp_last_error).  Numerical outcomes of a solve
 *    are reported per instance in bvp_result.retcode, as the reference
 *    reports them in ReturnCode (MIRK/mirk.jl:374-385, MIRK/adaptivity.jl:51-66).
 *  - State layout: node-major, component-minor ("point-major"): y[k*M + c]
 *    is component c at mesh node k, k = 0..N.  This is the layout of the
 *    reference's flattened unknown vector (Core/utils.jl:28-35, 59-79).
 *  - Batched entry points take n instances.  Per-instance outputs are packed
 *    back to back in instance order; the size of each slot is stated per
 *    entry point.
 *  - A context is not thread-safe: use one per host thread.
 */
#ifndef BVP_MIRK_H
#define BVP_MIRK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- methods (MIRK/algorithms.jl:55-61; tableaus MIRK/mirk_tableaus.jl:62-87, 120-152) */
#define BVP_MIRK4 4
#define BVP_MIRK6 6

/* ---- FIRK methods (lib/BoundaryValueDiffEqFIRK/src/, "FIRK/" below): bvp_opts.method
 * and the method argument of the bvp_firk_* entry points may name a fully implicit
 * Runge-Kutta collocation method instead of MIRK4/MIRK6:
 *   BVP_RADAU_IIA(s)     s = 1, 2, 3, 5, 7   order 2s - 1  (FIRK/radau_tableaus.jl:1-163)
 *   BVP_LOBATTO_IIIA(s)  s = 2, 3, 4, 5      order 2s - 2  (FIRK/lobatto_tableaus.jl:1-118)
 *   BVP_LOBATTO_IIIB(s)  s = 2, 3, 4, 5      order 2s - 2  (FIRK/lobatto_tableaus.jl:120-219)
 *   BVP_LOBATTO_IIIC(s)  s = 2, 3, 4, 5      order 2s - 2  (FIRK/lobatto_tableaus.jl:221-323)
 * They run in the expanded formulation, the reference's default (nested_nlsolve =
 * false; FIRK/firk.jl:290-453): the unknowns are the node values y_i and the s stage
 * values K_i1..K_is of every interval, in the order [y_0, K_01..K_0s, y_1, ..., y_N]
 * (FIRK/firk.jl:76-83).  code + BVP_FIRK_NESTED selects the nested form (nested_nlsolve =
 * true, FIRK/firk.jl:114-288): the unknowns are the node values, and every evaluation
 * solves each interval's stage equations from the fixed start K0 (u0 per stage, or the
 * mean of the guess, FIRK/utils.jl:41-83) by NewtonRaphson to bvp_opts.nest_abstol.
 * RadauIIa1, LobattoIIIb2 and LobattoIIIc2 do not adapt the
 * mesh (FIRK/algorithms.jl:411): they need adaptive = 0.  A STANDARD row inside the span
 * reads EvalSol of the FIRK cache (FIRK/interpolation.jl:338-441): the quintic of its
 * interval, whose midpoint pair the expanded form takes at (t - t_j) h in the stage
 * polynomial's variable, as the reference does, and the nested form at the midpoint from
 * stages solved from zero.  On the GPU: M = 1..4 and 6, DefectControl, FP64; nlsolve_stages
 * selects the stages of the polyalgorithm, the TrustRegion stage (bit 2) on the
 * expanded form. */
#define BVP_FIRK_NESTED     1000  /* added to a FIRK code: the nested form (below) */
#define BVP_RADAU_IIA(s)    (100 + (s))
#define BVP_LOBATTO_IIIA(s) (200 + (s))
#define BVP_LOBATTO_IIIB(s) (300 + (s))
#define BVP_LOBATTO_IIIC(s) (400 + (s))

/* ---- problem families: the right-hand side f(y, p, t) is compiled into the
 * device code, since a host closure cannot run on the GPU.  State dimensions:
 * M = 1..4 and 6 on the wave-per-instance kernels, M = 8, 16, 32 on the block
 * kernels (LINEAR_TANH).
 *   LINEAR_TANH : f = A y + c tanh(y);   params = [c, A(M*M, row-major)]
 *   PENDULUM    : f = [y2, -g sin(y1)];   params = [g]        (M = 2)
 *   BRATU       : f = [y2, -lambda e^y1]; params = [lambda]   (M = 2)
 *   SINGPERT    : f = [y2, (y1 - t y2)/eps]; params = [eps]   (M = 2)
 *   SCALAR_TU   : f = t y;                params = []         (M = 1)
 *   TURNING     : f = [y2, -t y2/eps - pi^2 cos(pi t) - pi t sin(pi t)/eps];
 *                 params = [eps]  (M = 2; the turning-point problem of the
 *                 reference's adaptivity test, test/misc/adaptivity_tests.jl:7-10)
 *   SWIRL       : f = [y2, (y1 y4 - y3 y2)/eps, y4, y5, y6, (-y3 y6 - y1 y2)/eps];
 *                 params = [eps]  (M = 6; the reference's Swirling Flow III test,
 *                 MIRKt/Core/mirk_basic_tests.jl:315-345) */
#define BVP_FAMILY_LINEAR_TANH 0
#define BVP_FAMILY_PENDULUM    1
#define BVP_FAMILY_BRATU       2
#define BVP_FAMILY_SINGPERT    3
#define BVP_FAMILY_SCALAR_TU   4
#define BVP_FAMILY_TURNING     5
#define BVP_FAMILY_SWIRL       6

/* ---- boundary-condition kinds
 *   TWO_POINT: TwoPointBVProblem.  Row k is y(bc_t[k])[bc_comp[k]] - bc_val[k]
 *              with bc_t[k] equal to t0 (a "bca" row) or t1 (a "bcb" row);
 *              residual layout [bca; Phi_1..Phi_N; bcb]  (MIRK/mirk.jl:509-521).
 *   STANDARD : BVProblem with bc!(resid, sol, p, t).  Row k is
 *              sol(bc_t[k])[bc_comp[k]] - bc_val[k], sol being the continuous
 *              MIRK extension (MIRK/interpolation.jl:214-242);
 *              residual layout [bc; Phi_1..Phi_N]  (MIRK/mirk.jl:471-482).
 *              Its Jacobian rows hold the stages frozen, as the reference's
 *              loss_bc does (MIRK/mirk.jl:565-572).                         */
#define BVP_BC_TWO_POINT 0
#define BVP_BC_STANDARD  1

/* ---- per-instance return codes (SciMLBase.ReturnCode subset) */
#define BVP_RET_SUCCESS  0
#define BVP_RET_FAILURE  1
#define BVP_RET_MAXITERS 2
#define BVP_RET_UNSTABLE 3   /* non-finite residual */

/* ---- linear-solve precision
 * M = 8, 16, 32 (block kernels): FP32_REFINE64 factorises in FP32 and refines each
 * linear solve refine_iters times against the FP64 residual.
 * M <= 4 (wave kernels): FP32_REFINE64 keeps the elimination in FP64 and stores the
 * chunk factors (the streamed back-substitution records) in FP32; the refinement is
 * Newton's own loop, whose residual and convergence test stay FP64 (refine_iters is
 * not used).  M = 6: FP64 records whatever the request (bvp_batch_precision reports
 * it). */
#define BVP_PREC_FP64          0  /* block LU in FP64 */
#define BVP_PREC_FP32_REFINE64 1  /* block LU in FP32 + FP64 iterative refinement */

/* ---- stages of the default nonlinear polyalgorithm (bvp_opts.nlsolve_stages;
 * Core/default_internal_solve.jl:31-45): NewtonRaphson, NewtonRaphson with the
 * BackTracking line search, TrustRegion (dogleg), tried in this order, each from
 * the same start */
#define BVP_STAGE_NEWTON       1
#define BVP_STAGE_BACKTRACKING 2
#define BVP_STAGE_TRUST_REGION 4
#define BVP_STAGES_DEFAULT     7

/* ---- error controllers (bvp_opts.controller; Core/calc_errors.jl:54-110,
 * MIRK/adaptivity.jl:23-243, 370-567) */
#define BVP_CONTROL_DEFECT     0  /* DefectControl (the default) */
#define BVP_CONTROL_GLOBAL     1  /* GlobalErrorControl(method) */
#define BVP_CONTROL_SEQUENTIAL 2  /* SequentialErrorControl: defect, then global error */
#define BVP_CONTROL_HYBRID     3  /* HybridErrorControl: DE * defect + GE * global error */
/* global-error methods (bvp_opts.global_method) */
#define BVP_GLOBAL_HO 0  /* HOErrorControl: MIRK(p+2) on the same mesh (MIRK4 only: no MIRK8) */
#define BVP_GLOBAL_RE 1  /* REErrorControl: Richardson extrapolation on the halved mesh */

/* ---- mesh-transfer mode after a redistribution (MIRK/mirk.jl:364-372) */
#define BVP_TRANSFER_REFERENCE 0  /* in place, as the reference writes y0 while reading it */
#define BVP_TRANSFER_CLEAN     1  /* interpolate from an untouched copy of the old solution */

/* ---- warm start */
#define BVP_WARM_NONE 0
#define BVP_WARM_MLP  1

/* ---- API error codes */
#define BVP_OK        0
#define BVP_EINVAL   -1
#define BVP_EHIP     -2
#define BVP_ENOMEM   -3
#define BVP_ENOTSUP  -4

typedef struct bvp_ctx bvp_ctx;

/* One problem instance.  Caller-owned, read-only during a call.
 * Mirrors BVProblem/TwoPointBVProblem (f, bc, u0, tspan, p) plus the mesh
 * that solve(...; dt) derives (Core/utils.jl:330-388, 694-704). */
typedef struct {
    int32_t family;          /* BVP_FAMILY_* */
    int32_t M;               /* state dimension */
    double t0, t1;           /* tspan */
    const double *params;    /* family parameters (see above) */
    int32_t n_params;
    int32_t bc_kind;         /* BVP_BC_* */
    int32_t n_bc;            /* number of BC rows; must equal M */
    const double *bc_t;      /* [n_bc] evaluation point of each row */
    const int32_t *bc_comp;  /* [n_bc] component (0-based) */
    const double *bc_val;    /* [n_bc] target value */
    int32_t N;               /* number of mesh intervals */
    const double *mesh;      /* [N+1] nodes, or NULL for range(t0, t1, N+1) */
    const double *guess;     /* [M*(N+1)] point-major initial guess, or NULL */
    const double *u0;        /* [M] constant guess used when guess == NULL
                                (Core/utils.jl:766-769) */
} bvp_instance;

/* Solver options.  Defaults (bvp_opts_default) follow the reference. */
typedef struct {
    int32_t method;               /* BVP_MIRK4 | BVP_MIRK6 */
    double  abstol;               /* 1e-6            MIRK/mirk.jl:50            */
    double  newton_abstol;        /* <=0 -> abstol   MIRK/mirk.jl:51            */
    int32_t adaptive;             /* 1               MIRK/mirk.jl:50            */
    double  defect_threshold;     /* 0.1             Core/calc_errors.jl:57     */
    int32_t max_num_subintervals; /* 3000            MIRK/algorithms.jl:60      */
    int32_t max_newton_iters;     /* 1000            (NonlinearSolve default)   */
    int32_t nlsolve_stages;       /* bit mask of the stages of the default nonlinear
                                     polyalgorithm run in order, each from the same
                                     start (Core/default_internal_solve.jl:31-45):
                                     bit0 NewtonRaphson, bit1 NewtonRaphson + BackTracking
                                     line search, bit2 TrustRegion (dogleg); default 7.
                                     Block kernels (M >= 8) run bits 0 and 1 only. */
    int32_t transfer_mode;        /* BVP_TRANSFER_*                              */
    int32_t precision;            /* BVP_PREC_*                                  */
    int32_t refine_iters;         /* FP64 refinement steps per linear solve (FP32 LU) */
    int32_t warm_start;           /* BVP_WARM_MLP: guess holds a learned guess;
                                     the solve falls back to u0 when its first
                                     Newton step does not contract (see below) */
    int32_t report_err;           /* 1: fill bvp_result.err_vs_fp64 (one more Newton
                                     correction per instance after the solve; the
                                     returned solution is unchanged).  No counterpart
                                     in the reference; default 0. */
    int32_t controller;           /* BVP_CONTROL_*; default DEFECT  MIRK/mirk.jl:51.
                                     The global-error controllers run on the M <= 4
                                     kernels with FP64 records (one kernel variant). */
    int32_t global_method;        /* BVP_GLOBAL_*; default HO  Core/calc_errors.jl:70 */
    double  hybrid_de, hybrid_ge; /* HybridErrorControl weights; 1, 1  Core/calc_errors.jl:100 */
    double  nest_abstol;          /* FIRK nested form: the inner solves stop at
                                     max|G| <= nest_abstol (nested_nlsolve_kwargs.abstol);
                                     <= 0: eps^(4/5), NonlinearSolve's default; default 0 */
} bvp_opts;

/* Per-instance result.  mesh/u are caller-allocated with room for
 * `capacity` nodes (at least max_num_subintervals + 1). */
typedef struct {
    int32_t retcode;             /* BVP_RET_* */
    int32_t N_final;             /* intervals of the returned mesh */
    int32_t newton_iters;        /* Newton iterations, summed over mesh passes */
    int32_t outer_iters;         /* mesh passes (MIRK/mirk.jl:299-307) */
    int32_t fallback_used;       /* bit0: the NewtonRaphson + BackTracking stage ran,
                                    bit1: warm-start fallback, bit2: FP64 fallback after an
                                    FP32 refinement stall, bit3: factor records did not fit
                                    (block kernels), bit4: the TrustRegion stage ran */
    int32_t capacity;            /* in: node capacity of mesh and u */
    double  defect_norm;         /* last defect estimate (MIRK/adaptivity.jl:410) */
    double  resid_norm;          /* max-norm of the final collocation residual */
    double  err_vs_fp64;         /* with opts.report_err: max_i |d_i| / max(1, |u_i|) for
                                    the FP64 Newton correction d = -J^-1 F at the returned
                                    (mesh, u): the distance, to first order, from u to the
                                    FP64 discrete solution on that mesh.  NaN otherwise.
                                    No counterpart in the reference. */
    double *mesh;                /* out: [N_final+1] */
    double *u;                   /* out: [M*(N_final+1)] point-major */
    double *stages;              /* out, FIRK methods: [N_final*s*M] the stage values of the
                                    final state (interval, then stage, then component), or
                                    NULL; not written for MIRK */
} bvp_result;

void bvp_opts_default(bvp_opts *opts);

int  bvp_ctx_create(int device, bvp_ctx **out);
void bvp_ctx_destroy(bvp_ctx *ctx);
const char *bvp_last_error(const bvp_ctx *ctx);

/* Full solve of n instances in one persistent launch on the GPU.
 * Replaces SciMLBase.__solve -> __init + solve! for MIRK
 * (Core/BoundaryValueDiffEqCore.jl:48-54; MIRK/mirk.jl:49-265, 286-388).
 * All instances share opts and must share M. */
int bvp_mirk_solve_batch(bvp_ctx *ctx, const bvp_opts *opts, int32_t n,
                         const bvp_instance *inst, bvp_result *res);

/* bvp_mirk_solve_batch without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * packed inputs, working state, elimination records and result staging of the call
 * live in a static device array of 64 MiB (held by each device from library load
 * until the process ends); the call allocates no device memory and returns
 * BVP_ENOMEM with a message when the batch does not fit (about 100 instances at
 * max_num_subintervals = 3000 for M = 2; fewer intervals, more instances).  The
 * wave-kernel families only (M = 1..4, 6; BVP_ENOTSUP otherwise).  Results are the same bits as
 * bvp_mirk_solve_batch.  All calls on one device share the array: calls that
 * overlap on one device (other host threads, other contexts) race. */
int bvp_mirk_solve_batch_noalloc(bvp_ctx *ctx, const bvp_opts *opts, int32_t n,
                                 const bvp_instance *inst, bvp_result *res);

/* Collocation residual at (mesh, guess) of each instance, reference layout
 * (TWO_POINT: [bca; Phi; bcb], STANDARD: [bc; Phi]).  Slot: M*(N+1) doubles.
 * Replaces the loss functions MIRK/mirk.jl:471-534 with
 * Phi! MIRK/collocation.jl:44-102 and the BC residual Core/utils.jl:157-209. */
int bvp_mirk_residual(bvp_ctx *ctx, int32_t method, int32_t n,
                      const bvp_instance *inst, double *resid);

/* Jacobian blocks at (mesh, guess): S_i = dPhi_i/dy_i and R_i = dPhi_i/dy_{i+1}
 * (M*M row-major each; slot N*M*M doubles for S and for R), computed with
 * in-kernel forward-mode dual numbers; bc_node receives, per BC row, the mesh
 * node whose component bc_comp[k] carries the unit Jacobian entry (slot n_bc).
 * Replaces the sparse ForwardDiff Jacobian MIRK/mirk.jl:697-862, 927-1002 and
 * MIRK/sparse_jacobians.jl:16-36. */
int bvp_mirk_jacobian_blocks(bvp_ctx *ctx, int32_t method, int32_t n,
                             const bvp_instance *inst, double *S, double *R,
                             int32_t *bc_node);

/* One Newton correction: solves J(y) delta = -F(y) at (mesh, guess) with the
 * partitioned block elimination (slot M*(N+1) doubles, point-major).
 * `precision` is BVP_PREC_*; refine_iters applies to FP32.
 * Replaces the LinearSolve call on the AlmostBanded/Banded Jacobian handed
 * over at Core/internal_problems.jl:107-169. */
int bvp_mirk_newton_step(bvp_ctx *ctx, int32_t method, int32_t precision,
                         int32_t refine_iters, int32_t n,
                         const bvp_instance *inst, double *delta);

/* Defect estimate at (mesh, guess): per-interval defect vectors (slot N*M)
 * and defect norm (one per instance).  Includes the extra interpolation
 * stages.  Replaces error_estimate!(::DefectControl) MIRK/adaptivity.jl:370-461
 * with interp_setup! MIRK/interpolation.jl:300-372. */
int bvp_mirk_defect(bvp_ctx *ctx, int32_t method, int32_t n,
                    const bvp_instance *inst, double *errors,
                    double *defect_norm);

/* Mesh selection and solution transfer at (mesh, guess) after a converged
 * solve: defect -> mesh_selector! (halving or redistribution) -> interp_eval!
 * of the old continuous extension onto the new nodes.
 * Outputs per instance: new_N (1 int32), info (BVP_RET_SUCCESS/FAILURE),
 * new_mesh (slot `capacity` doubles), new_y (slot M*capacity doubles).
 * Replaces MIRK/adaptivity.jl:6-75, 250-304 and MIRK/mirk.jl:361-373. */
int bvp_mirk_mesh_select(bvp_ctx *ctx, const bvp_opts *opts, int32_t n,
                         const bvp_instance *inst, int32_t capacity,
                         int32_t *new_N, int32_t *info, double *new_mesh,
                         double *new_y);

/* The global-error estimate of GlobalErrorControl at (mesh, guess) of each
 * instance (error_estimate! with REErrorControl / HOErrorControl and
 * global_error + GE_subinterval!, MIRK/adaptivity.jl:511-556): opts.global_method
 * picks the method (RE: the same method on the halved mesh; HO: MIRK6 on the same
 * mesh, MIRK4 only), solved with opts' Newton tolerance, iteration cap and stages.
 * Outputs: errors [sum N][M] (per interval the end-node relative error vector with
 * the larger max-abs, instance order), norm [n] (RE: scaled by 2^p / (2^p - 1)).
 * M <= 4 families. */
int bvp_mirk_global_error(bvp_ctx *ctx, const bvp_opts *opts, int32_t n, const bvp_instance *inst,
                          double *errors, double *norm);

/* Dense output: the continuous MIRK extension of a solution and its first
 * derivative at arbitrary points -- sol(t) and sol(t, Val{1}) of
 * MIRKInterpolation (MIRK/interpolation.jl:17-206).  The solution of instance i
 * is its (mesh, guess): N intervals, the N+1 nodes and the point-major state, as
 * a solve returns them.  n_t[i] points of instance i are packed in t (instance
 * order); u receives M values per point (packed the same way), du (optional,
 * NULL to skip) the derivative.  A point lies in interval(mesh, t)
 * (Core/utils.jl:119-121: a node belongs to the interval on its left; points
 * outside [t0, t1] extrapolate the end intervals).  The wave-kernel families
 * (M = 1..4, 6). */
int bvp_mirk_interpolate(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst,
                         const int32_t *n_t, const double *t, double *u, double *du);

/* ---- FIRK components.  The state of instance i is its (mesh, guess) and its stage
 * values: `stages` holds N_i*s*M doubles per instance (interval, then stage, then
 * component; instances back to back), or is NULL for the stages the reference
 * starts from, K_ij = y_{i+1} (extend_y, FIRK/firk.jl:76-83).  M = 1..4, 6.  Slot sizes
 * below use n_x = M*(N*(s+1)+1), the length of the
 * expanded unknown vector.  With a nested code (+ BVP_FIRK_NESTED) `stages` is not
 * read (the inner solves give them, from K0, at the default inner tolerance), and
 * residual and correction have the node layout, n_x = M*(N+1): [bc; Phi] or
 * [bca; Phi; bcb] (FIRK/collocation.jl:140-428). */

/* Expanded residual in the reference's layout (loss FIRK/firk.jl:1404-1455 with Phi!
 * FIRK/collocation.jl:59-93): STANDARD [bc; per interval Phi_i, G_i1..G_is],
 * TWO_POINT [bca; per interval ...; bcb], with
 *   Phi_i = y_{i+1} - y_i - h_i sum_j b_j K_ij,
 *   G_ir  = f(t_i + c_r h_i, y_i + h_i sum_j a_rj K_ij) - K_ir.
 * Slot n_x doubles. */
int bvp_firk_residual(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst,
                      const double *stages, double *resid);

/* One Newton correction of the expanded system, J delta = -F, in the unknown order
 * (slot n_x doubles).  Replaces the LinearSolve call on the banded Jacobian of
 * FIRK/firk.jl:806-922 (Core/internal_problems.jl:107-169). */
int bvp_firk_newton_step(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst,
                         const double *stages, double *delta);

/* Defect estimate (defect_estimate! FIRK/adaptivity.jl:349-405): per-interval defect
 * vectors from the stage polynomial at tau* and 1 - tau* (slot N*M) and the norm (one
 * per instance). */
int bvp_firk_defect(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst,
                    const double *stages, double *errors, double *defect_norm);

/* Dense output sol(t), sol(t, Val{1}) of FIRKExpandInterpolation
 * (FIRK/interpolation.jl:229-316): per interval the quintic through y_i, y_{i+1},
 * f(y_i), f(y_{i+1}) and the stage polynomial's value and slope at the midpoint.
 * Points packed as for bvp_mirk_interpolate. */
int bvp_firk_interpolate(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst,
                         const double *stages, const int32_t *n_t, const double *t, double *u,
                         double *du);

/* ---- batch plans: upload once, solve many times (no counterpart in the
 * reference, whose batches are EnsembleProblem solves run one per thread,
 * MIRKt/Core/ensemble_tests.jl:7-41).  For callers that time or repeat the
 * device solve with the inputs already resident in device memory.
 *   create : validates, packs and uploads the instances (synchronous)
 *   run    : enqueues the persistent solve on the context's stream
 *   sync   : waits for it; *ms = device time of the solve kernel
 *   results: copies the per-instance results out (as bvp_mirk_solve_batch)
 *   counters: work done by the last run, summed over instances:
 *            [0] interval-Newton-iterations, [1] defect intervals,
 *            [2] transfer nodes (old + new), [3] convergence-check intervals
 *   info   : [0] persistent waves, [1] node capacity, [2] M, [3] scratch doubles per wave */
typedef struct bvp_batch bvp_batch;
int  bvp_batch_create(bvp_ctx *ctx, const bvp_opts *opts, int32_t n,
                      const bvp_instance *inst, bvp_batch **out);
int  bvp_batch_run(bvp_ctx *ctx, bvp_batch *batch);
int  bvp_batch_sync(bvp_ctx *ctx, bvp_batch *batch, float *ms);
int  bvp_batch_results(bvp_ctx *ctx, bvp_batch *batch, bvp_result *res);
int  bvp_batch_counters(bvp_ctx *ctx, bvp_batch *batch, long long *out4);
int  bvp_batch_info(bvp_batch *batch, int32_t *out4);
void bvp_batch_destroy(bvp_batch *batch);

/* ---- hipGraph replay of a plan (no counterpart in the reference).  With
 * enable != 0 (the default) the first run captures the plan's launches (queue
 * reset, learned guess if attached, persistent solve) into a graph and later
 * runs replay it; attaching or detaching a network re-captures. */
int  bvp_batch_set_graph(bvp_batch *batch, int32_t enable);
/* ---- autotuner (no counterpart in the reference).  Times the plan's solve
 * with each launch configuration it can take -- the persistent solve kernel's
 * variants (for M = 2: register allocation held to 1, 2 or 3 waves per SIMD,
 * trading spills for latency hiding) at their full residency, or, for the block
 * kernels, 1, 2, 4 ... waves per CU -- `repeats` timed runs each (median), keeps
 * the fastest for the plan and remembers it in the context for later plans of
 * the same kind (M, method, precision).  A wave-kernel plan (M <= 4) built with
 * BVP_PREC_FP32_REFINE64 also times its variants with FP64 records (listed
 * after the FP32 ones) and keeps FP64 where that is faster: on stiff families
 * the FP32-record corrections cost extra Newton steps.  cand_ms (optional,
 * room for 16) receives the median of every candidate, *n_cand their number;
 * a wave-kernel plan then also times the interval-per-lane mapping of its fastest
 * candidate (bvp_batch_chunking; three more candidates).
 * The runs overwrite the plan's results. */
int  bvp_batch_autotune(bvp_ctx *ctx, bvp_batch *batch, int32_t repeats, float *cand_ms, int32_t *n_cand);
/* The interval-per-wavefront mapping of a wave-kernel plan (no counterpart in the
 * reference): an instance's N intervals are split into P = clamp(N / min_chunk, 1,
 * 64) contiguous chunks, one per lane of its wavefront (chunk elimination, then a
 * merge tree across the lanes).  The autotuner times min_chunk = 2, 4, 8 against 1
 * for its fastest kernel variant and keeps the fastest; it can be read or set
 * (1 ... 8; block-kernel plans: 1).  Different mappings give the same solution to
 * rounding (the elimination order differs). */
int  bvp_batch_chunking(bvp_batch *batch, int32_t *min_chunk);
int  bvp_batch_set_chunking(bvp_batch *batch, int32_t min_chunk);
/* The record / factor precision the plan runs with (BVP_PREC_*): the requested
 * one, or BVP_PREC_FP64 where the autotuner chose FP64 records (no counterpart
 * in the reference). */
int  bvp_batch_precision(bvp_batch *batch, int32_t *precision);
/* The plan's current launch configuration: kernel variant and persistent waves;
 * and setting it by hand (blocks at most the waves the plan was sized for). */
int  bvp_batch_config(bvp_batch *batch, int32_t *variant, int32_t *blocks);
int  bvp_batch_set_config(bvp_ctx *ctx, bvp_batch *batch, int32_t variant, int32_t blocks);

/* Device time of the last synced run split into the learned-guess part (queue
 * reset + guess kernel; 0 without a network) and the persistent solve. */
int  bvp_batch_timing(bvp_batch *batch, float *guess_ms, float *solve_ms);

/* ---- learned warm start (no counterpart in the reference; it plugs in where
 * the reference takes the initial guess from the caller, as a constant vector,
 * a vector of arrays, a function of (p, t) or an earlier solution:
 * __initial_guess_on_mesh, Core/utils.jl:750-777, and the u0 handling of
 * MIRK/mirk.jl:83-95).
 *
 * The network is a 4 -> 256 -> 256 -> 256 -> 2 MLP with tanh hidden units,
 * BF16 weights and FP32 accumulation, that maps
 *     (p, a, b, t) -> (y1, y2)(t)
 * for M = 2 families, p being params[0] (or log10 params[0], feature kind
 * BVP_FEAT_LOG10_PARAM), a and b the values of boundary rows 0 and 1.  Inputs
 * are normalised as (x - in_shift) * in_scale and outputs mapped back as
 * out * out_scale + out_shift.  Parameter vector (FP32, BVP_MLP_NPARAM values,
 * matrices input-major W[i][o]): W1[4][256] b1[256] W2[256][256] b2[256]
 * W3[256][256] b3[256] W4[256][2] b4[2].
 *
 * The solve uses it as follows (opts.warm_start == BVP_WARM_MLP, or a plan with
 * a network attached): Newton starts from the learned guess on the first mesh;
 * when its first step does not reduce max|F| (or the guess gives a non-finite
 * residual or a singular Newton matrix), the solve restarts from the constant
 * guess u0 exactly as the plain method would, and sets bit1 of fallback_used. */
#define BVP_MLP_NPARAM 133378
#define BVP_FEAT_PARAM       0
#define BVP_FEAT_LOG10_PARAM 1

typedef struct bvp_mlp bvp_mlp;

typedef struct {
    int32_t  epochs;        /* passes over the samples */
    int32_t  batch;         /* points per step, a multiple of 64 */
    float    lr, lr_min;    /* peak and final learning rate (cosine decay) */
    float    weight_decay;  /* decoupled, on weight matrices */
    int32_t  warmup;        /* linear learning-rate warm-up steps */
    uint64_t seed;          /* sample order */
    int32_t  use_graph;     /* replay one captured training step */
    int32_t  first_epoch;   /* this call runs epochs [first_epoch, first_epoch + epochs) of a */
    int32_t  total_epochs;  /* schedule of total_epochs (<= 0: epochs); calls that split a
                               schedule this way give the same result as one call */
} bvp_train_opts;

/* Glorot-uniform weights from `seed`, zero biases, identity normalisation. */
int  bvp_mlp_create(bvp_ctx *ctx, uint64_t seed, bvp_mlp **out);
void bvp_mlp_destroy(bvp_mlp *mlp);
int  bvp_mlp_set_normalization(bvp_mlp *mlp, const float *in_shift, const float *in_scale,
                               const float *out_shift, const float *out_scale);
int  bvp_mlp_get_params(bvp_mlp *mlp, float *w);        /* FP32 master parameters */
int  bvp_mlp_set_params(bvp_mlp *mlp, const float *w);  /* also clears the optimiser state */
/* Loss L = sum_points 0.5 |out - y_normalised|^2 / n and its FP32 gradient with
 * respect to the parameter vector, over n points x[n][4] (raw features) with
 * targets y[n][2] (raw), in one forward/backward pass of the training kernels. */
int  bvp_mlp_gradient(bvp_mlp *mlp, int32_t n, const float *x, const float *y, float *grad, float *loss);
/* y[n][2] = network(x[n][4]) in raw units. */
int  bvp_mlp_predict(bvp_mlp *mlp, int32_t n, const float *x, float *y);
/* Adam training on n samples (n < 2^31); loss_per_step (optional) receives the
 * loss of every step of this call, *steps their number, *ms the device time. */
int  bvp_mlp_train(bvp_mlp *mlp, int64_t n, const float *x, const float *y, const bvp_train_opts *opts,
                   float *loss_per_step, int32_t *steps, float *ms);
/* Attach (or with mlp == NULL detach) a network to a plan: each run first writes
 * the learned guess of every node on the device, then solves with the
 * contraction-checked fallback described above.  anchors == 0 evaluates the
 * network at every node of the initial mesh; anchors = K >= 2 evaluates it at K
 * equally spaced points of [t0, t1] per instance and fills the nodes by cubic
 * Hermite interpolation with the slopes y' = f(t, y) of the problem at those
 * points (K network evaluations per instance instead of N + 1).  M = 2 plans
 * only; every instance must carry u0.  The plan keeps the pointer: detach or
 * destroy the plan before the network. */
int  bvp_batch_set_warm_start(bvp_batch *batch, const bvp_mlp *mlp, int32_t feature_kind, int32_t anchors);
/* The learned guess of the attached network for every node of the initial
 * meshes, point-major and packed in instance order (sum (N+1) * M doubles): the
 * initial guess a warm-started run starts Newton from. */
int  bvp_batch_learned_guess(bvp_ctx *ctx, bvp_batch *batch, double *y);

/* ---- measurement helper (no counterpart in the reference): a plain device
 * copy of nbytes, run `repeats` times; *ms = fastest device time.  It is the
 * known-size reference for scaling profiler HBM byte counters. */
int  bvp_copy_probe(bvp_ctx *ctx, long long nbytes, int repeats, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* BVP_MIRK_H */
