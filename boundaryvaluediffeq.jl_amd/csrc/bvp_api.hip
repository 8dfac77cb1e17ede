// bvp_api.hip -- C ABI of the MIRK collocation library (include/bvp_mirk.h).
//
// Host side of the drop-in boundary: validates and packs the caller's problem
// instances, owns device memory and the stream, launches the gfx950 kernels and
// copies results back.  No pointer passed in is retained after a call returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/bvp_mirk.h"
#include "bvp_host.h"
#include "bvp_internal.h"
#include "mirk_wave.h"

namespace {

void set_err(bvp_ctx *ctx, const char *fmt, ...) {
    if (!ctx) return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
}

#define HIP_TRY(ctx, expr)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            set_err((ctx), "%s failed: %s", #expr, hipGetErrorString(e_));                \
            return BVP_EHIP;                                                               \
        }                                                                                  \
    } while (0)

using bvphost::params_needed;
using bvphost::uniform_mesh;

bool any_kernel_supports(int M, int family) { return wave_supports(M, family) || blockm_supports(M, family); }

int validate(bvp_ctx *ctx, int32_t n, const bvp_instance *inst, int &M) {
    std::string err;
    const int rc = bvphost::validate_instances(n, inst, M, any_kernel_supports, err);
    if (rc) set_err(ctx, "%s", err.c_str());
    return rc;
}

template <typename T>
int dalloc(bvp_ctx *ctx, DevArena &arena, T *&p, size_t count) {
    p = nullptr;
    if (count == 0) count = 1;
    if (arena.bump) {
        const size_t bytes = (count * sizeof(T) + 255) & ~(size_t)255;
        if (arena.bump_used + bytes > arena.bump_cap) {
            set_err(ctx, "the call needs more than the %zu bytes of the no-allocation scratch (%zu more bytes "
                         "requested with %zu in use): solve fewer instances per call or lower "
                         "max_num_subintervals", arena.bump_cap, bytes, arena.bump_used);
            return BVP_ENOMEM;
        }
        p = (T *)(arena.bump + arena.bump_used);
        arena.bump_used += bytes;
        return 0;
    }
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        set_err(ctx, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T), hipGetErrorString(e));
        return BVP_ENOMEM;
    }
    arena.ptrs.push_back(q);
    p = (T *)q;
    return 0;
}

template <typename T>
int upload(bvp_ctx *ctx, DevArena &arena, T *&p, const std::vector<T> &v) {
    int rc = dalloc(ctx, arena, p, v.size());
    if (rc) return rc;
    if (!v.empty()) HIP_TRY(ctx, hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

}  // namespace

DevArena::~DevArena() {
    for (void *p : ptrs) (void)hipFree(p);
}

// ------------------------------------------------------------------ no-allocation scratch
// Device memory of bvp_mirk_solve_batch_noalloc: the packed inputs, working
// state, elimination records and result staging of one call are carved from this
// array, which each device holds from module load until the process ends.  Host
// code reaches it only through the address hipGetSymbolAddress returns for the
// context's device (bvp_ctx::noalloc); calls on one device share it.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// ------------------------------------------------------------------ batch build
int batch_build(bvp_ctx *ctx, const bvp_opts *o, int32_t n, const bvp_instance *inst, int cap_req, bvp_batch **out,
                char *bump, size_t bump_cap) {
    int M;
    int rc = validate(ctx, n, inst, M);
    if (rc) return rc;
    const bool blk = !wave_supports(M, inst[0].family);
    if (blk && bump) {
        set_err(ctx, "the no-allocation entry point serves the wave-kernel state dimensions (M = 1..4, 6)");
        return BVP_ENOTSUP;
    }
    if (blk) {
        for (int i = 0; i < n; ++i)
            if (!block_supports(M, inst[i].family)) {
                set_err(ctx, "instance %d: family %d with M=%d is not supported", i, inst[i].family, M);
                return BVP_ENOTSUP;
            }
    }
    {
        std::string err;
        if ((rc = bvphost::validate_opts(o, n, inst, M, blk, err))) {
            set_err(ctx, "%s", err.c_str());
            return rc;
        }
    }
    const bool firk = bvphost::is_firk(o->method);
    bvphost::FirkTableau ftab;
    if (firk) bvphost::firk_tableau(o->method, ftab);  // validate_opts accepted the code
    bool has_u0 = true;
    for (int i = 0; i < n; ++i) has_u0 &= inst[i].u0 != nullptr;
    int maxN0 = 0;
    for (int i = 0; i < n; ++i) maxN0 = std::max(maxN0, inst[i].N);
    int cap = std::max(cap_req, std::max(o->max_num_subintervals, maxN0) + 1);
    auto *B = new bvp_batch();
    B->arena.bump = bump;
    B->arena.bump_cap = bump_cap;
    B->n = n;
    B->M = M;
    B->method = o->method;
    B->cap = cap;
    B->opts = *o;
    B->block = blk;
    B->firk = firk;
    B->firk_s = firk ? ftab.s : 0;
    B->has_u0 = has_u0;
    // ---- pack the problem description
    bvphost::Packed pk;
    bvphost::pack_instances(n, inst, M, pk);
    B->N0 = pk.N0;
    B->init_off = pk.off;
    B->total_nodes = pk.nodes;
    std::vector<int> &fam = pk.fam, &poff = pk.poff, &pcnt = pk.pcnt, &bck = pk.bck, &N0 = pk.N0, &bcc = pk.bcc,
                     &order = pk.order;
    std::vector<double> &t0 = pk.t0, &t1 = pk.t1, &params = pk.params, &bct = pk.bct, &bcv = pk.bcv,
                        &imesh = pk.mesh, &iy = pk.y, &u0c = pk.u0c;
    std::vector<long long> &ioff = pk.off;
    const bool nested = firk && o->method >= BVP_FIRK_NESTED;
    if (nested)
        // the inner solves' start k0 (__K0_on_u0, FIRK/utils.jl:41-83): u0, else the mean of
        // the guess's nodes
        for (int i = 0; i < n; ++i) {
            if (inst[i].u0) continue;
            double *k0 = &u0c[(size_t)i * M];
            const double *yi = &iy[(size_t)ioff[i] * M];
            for (int k = 0; k <= N0[i]; ++k)
                for (int c = 0; c < M; ++c) k0[c] += yi[(size_t)k * M + c];
            for (int c = 0; c < M; ++c) k0[c] /= (N0[i] + 1);
        }

    BatchDev &D = B->dev;
    std::memset(&D, 0, sizeof D);
    D.n = n;
    D.M = M;
    D.cap = cap;
    DevArena &A = B->arena;
    int *dfam, *dpoff, *dpcnt, *dbck, *dbcc, *dN0, *dorder;
    double *dt0, *dt1, *dparams, *dbct, *dbcv, *dim, *diy, *du0c;
    long long *dioff;
    if ((rc = upload(ctx, A, dfam, fam)) || (rc = upload(ctx, A, dpoff, poff)) || (rc = upload(ctx, A, dpcnt, pcnt)) ||
        (rc = upload(ctx, A, dbck, bck)) || (rc = upload(ctx, A, dbcc, bcc)) || (rc = upload(ctx, A, dN0, N0)) ||
        (rc = upload(ctx, A, dorder, order)) || (rc = upload(ctx, A, dt0, t0)) || (rc = upload(ctx, A, dt1, t1)) ||
        (rc = upload(ctx, A, dparams, params)) || (rc = upload(ctx, A, dbct, bct)) || (rc = upload(ctx, A, dbcv, bcv)) ||
        (rc = upload(ctx, A, dim, imesh)) || (rc = upload(ctx, A, diy, iy)) || (rc = upload(ctx, A, dioff, ioff)) ||
        (rc = upload(ctx, A, du0c, u0c))) {
        delete B;
        return rc;
    }
    D.family = dfam; D.param_off = dpoff; D.param_cnt = dpcnt; D.bc_kind = dbck; D.bc_comp = dbcc;
    D.N0 = dN0; D.order = dorder; D.t0 = dt0; D.t1 = dt1; D.params = dparams; D.bc_t = dbct; D.bc_val = dbcv;
    D.init_mesh = dim; D.init_y = diy; D.init_off = dioff;
    D.u0c = du0c;
    D.warm = (o->warm_start == BVP_WARM_MLP) ? 1 : 0;
    // working state and results
    if ((rc = dalloc(ctx, A, D.mesh, (size_t)n * 2 * cap)) || (rc = dalloc(ctx, A, D.y, (size_t)n * 2 * cap * M)) ||
        (rc = dalloc(ctx, A, D.shat, (size_t)n * cap)) || (rc = dalloc(ctx, A, D.retcode, n)) ||
        (rc = dalloc(ctx, A, D.N_final, n)) || (rc = dalloc(ctx, A, D.newton_iters, n)) ||
        (rc = dalloc(ctx, A, D.outer_iters, n)) || (rc = dalloc(ctx, A, D.fallback, n)) ||
        (rc = dalloc(ctx, A, D.cur_buf, n)) || (rc = dalloc(ctx, A, D.defect_norm, n)) ||
        (rc = dalloc(ctx, A, D.resid_norm, n)) || (rc = dalloc(ctx, A, D.err, n)) ||
        (rc = dalloc(ctx, A, D.cnt, (size_t)n * 4)) ||
        (rc = dalloc(ctx, A, D.counter, 1))) {
        delete B;
        return rc;
    }
    // persistent slots: the kernel variant and waves per CU from an earlier tuning
    // of this kind of plan, else variant 0 at its occupancy; scratch is sized for
    // the most waves any variant can keep resident, so that the autotuner can try them
    // wave kernels: FP32 storage of the chunk elimination factors (the records are
    // smaller than the FP64 slots they are laid out in)
    // (M = 6: FP64 records only)
    const int f32 = (!blk && !firk && M <= 4 && o->precision == BVP_PREC_FP32_REFINE64) ? 1 : 0;
    int ufam = inst[0].family;  // the batch's single family (two-point BCs only), or -1
    for (int i = 0; i < n; ++i)
        if (inst[i].family != ufam || inst[i].bc_kind != BVP_BC_TWO_POINT) ufam = -1;
    B->family = ufam;
    // kernel kind: 0 wave, 1 block, 2 wave under a global-error controller, 3 FIRK
    const auto key = std::make_tuple(M, o->method, o->precision,
                                     blk ? 1 : (firk ? 3 : (o->controller != BVP_CONTROL_DEFECT ? 2 : 0)), ufam);
    int variant = 0, per_cu = 0, tuned_f32 = f32, tuned_minc = 1;
    if (ctx->tuned.count(key)) std::tie(variant, per_cu, tuned_f32, tuned_minc) = ctx->tuned[key];
    if (ctx->blocks_per_cu > 0) per_cu = ctx->blocks_per_cu;
    if (firk && (variant < 0 || variant >= kFirkVariants)) variant = 0;
    if (per_cu < 1)
        per_cu = blk ? block_blocks_per_cu(M, o->method, o->precision)
                     : (firk ? firk_blocks_per_cu(M, ftab.s, variant, nested)
                             : wave_blocks_per_cu(M, o->method, variant, ufam, f32));
    if (per_cu < 1) per_cu = 1;
    int slots = std::min(n, ctx->num_cus * per_cu);
    if (firk)
        for (int v = 0; v < kFirkVariants; ++v)
            slots = std::max(slots, std::min(n, ctx->num_cus * std::max(1, firk_blocks_per_cu(M, ftab.s, v, nested))));
    if (!blk && !firk)
        for (int v = 0; v < wave_variants(M); ++v)
            slots = std::max(slots, std::min(n, ctx->num_cus * std::max(1, wave_blocks_per_cu(M, o->method, v, ufam, f32))));
    if (slots < 1) slots = 1;
    B->variant = variant;
    size_t scratch_doubles;
    const bool ge = o->controller != BVP_CONTROL_DEFECT;
    if (!blk) {
        const int NF = wave_rel_fields(M);
        // records per lane: a chunk holds at most max(ceil(cap / 64), 2 kMaxMinChunk)
        // intervals; the Richardson estimate solves on the halved mesh (2 cap)
        const size_t rcap = (size_t)cap * (ge ? 2 : 1);
        const size_t lmax = std::max((rcap + 63) / 64, (size_t)(2 * kMaxMinChunk)) + 1;
        D.tree_base = lmax * NF * 64;
        D.scratch_per_slot = D.tree_base + (size_t)64 * NF;
        D.aux_per_slot = firk ? firk_aux_doubles(cap, M, ftab.s) : (size_t)cap * (3 * M + 1);
        scratch_doubles = D.scratch_per_slot * slots;
        D.rec_cap = cap;
    } else {
        // factor records of one instance live in its wave's slot: rec_cap intervals.
        const size_t esz = (o->precision == BVP_PREC_FP32_REFINE64) ? 4 : 8;
        const size_t rv = block_record_values(M);
        // budget: 80% of the device memory free at this point (the problem data is
        // already allocated)
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b == 0) free_b = (size_t)120 << 30;
        const size_t budget = free_b / 10 * 8;
        // room for every mesh the solve may reach (max_num_subintervals) when it fits,
        // else for meshes up to twice the largest initial one, else for that one
        int rec_cap = std::max(o->max_num_subintervals, maxN0);
        if ((size_t)rec_cap * rv * esz * slots > budget) rec_cap = std::min(rec_cap, 2 * maxN0);
        if ((size_t)rec_cap * rv * esz * slots > budget) rec_cap = maxN0;
        const size_t per_slot = (size_t)rec_cap * rv;  // in factor-type elements
        if (per_slot * esz * slots > budget) slots = std::max(1, (int)(budget / (per_slot * esz)));
        D.rec_cap = rec_cap;
        D.tree_base = 0;
        D.scratch_per_slot = per_slot;
        D.aux_per_slot = (size_t)cap * (3 * M + 1);
        scratch_doubles = (per_slot * esz * slots + 7) / 8;
    }
    B->max_slots = slots;
    B->blocks = std::min(slots, std::max(1, std::min(n, ctx->num_cus * per_cu)));
    D.refine_iters = o->refine_iters;
    D.f32rec = f32 && tuned_f32 && !ge;  // the autotuner may have found FP64 records faster for this kind;
                                         // the global-error kernel runs FP64 records
    D.controller = o->controller;
    D.global_method = o->global_method;
    D.hybrid_de = o->hybrid_de;
    D.hybrid_ge = o->hybrid_ge;
    D.ge_per_slot = ge ? wave_ge_doubles(cap, M) : 0;
    D.min_chunk = blk ? 1 : std::max(1, std::min(kMaxMinChunk, tuned_minc));
    if ((rc = dalloc(ctx, A, D.scratch, scratch_doubles)) ||
        (rc = dalloc(ctx, A, D.aux, D.aux_per_slot * slots)) ||
        (ge && (rc = dalloc(ctx, A, D.ge, D.ge_per_slot * slots)))) {
        delete B;
        return rc;
    }
    if (firk) {
        FirkTab tab{};
        std::memcpy(tab.A, ftab.A, sizeof tab.A);
        std::memcpy(tab.b, ftab.b, sizeof tab.b);
        std::memcpy(tab.c, ftab.c, sizeof tab.c);
        std::memcpy(tab.Q, ftab.Q, sizeof tab.Q);
        tab.tau = ftab.tau;
        tab.s = ftab.s;
        tab.order = ftab.order;
        FirkTab *dtab = nullptr;
        if ((rc = upload(ctx, A, dtab, std::vector<FirkTab>(1, tab))) ||
            (rc = dalloc(ctx, A, D.stK, (size_t)n * cap * ftab.s * M))) {
            delete B;
            return rc;
        }
        D.firk_tab = dtab;
        D.firk_nested = nested;
        D.firk_irows = (o->nlsolve_stages & BVP_STAGE_TRUST_REGION) != 0;  // the stage lives in those kernels
        for (int i = 0; i < n; ++i)
            if (inst[i].bc_kind == BVP_BC_STANDARD)
                for (int k = 0; k < M; ++k)
                    D.firk_irows |= inst[i].bc_t[k] != inst[i].t0 && inst[i].bc_t[k] != inst[i].t1;
        D.nest_tol = o->nest_abstol > 0 ? o->nest_abstol : 3.000213634488528e-13;  // eps^(4/5)
    }
    // options
    D.adaptive = o->adaptive;
    D.max_sub = o->max_num_subintervals;
    D.max_newton = o->max_newton_iters;
    D.stages = o->nlsolve_stages;
    D.report_err = o->report_err ? 1 : 0;
    D.transfer_mode = o->transfer_mode;
    D.abstol = o->abstol;
    D.newton_tol = o->newton_abstol > 0 ? o->newton_abstol : o->abstol;
    D.defect_threshold = o->defect_threshold;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        set_err(ctx, "upload failed: %s", hipGetErrorString(e));
        delete B;
        return BVP_EHIP;
    }
    *out = B;
    return 0;
}

// The M = 32 block factorisation in FP64 hung on the GPU (cause not identified);
// it stays switched off until it is understood.  Operations that do not factor
// (residual, Jacobian blocks, defect, mesh selection) are not affected.
static bool fp64_block32_blocked(bvp_ctx *ctx, const bvp_batch *B) {
    if (B->block && B->M == 32 && B->opts.precision == BVP_PREC_FP64) {
        ctx->err = "M=32 with FP64 block factors is switched off; use BVP_PREC_FP32_REFINE64";
        return true;
    }
    return false;
}

// The launches of one run: the learned guess of every node when a network is
// attached, then the queue reset (a fill kernel, so that it can be captured) and
// the persistent solve.
static int enqueue_guess(bvp_ctx *ctx, bvp_batch *B) {
    HIP_TRY(ctx, mlp::launch_batch_guess(B->mlp->net(), B->dev, B->warm_y, B->total_nodes, B->feature_kind,
                                         B->anchors, B->anc, ctx->num_cus, ctx->stream));
    return 0;
}

static int enqueue_solve(bvp_ctx *ctx, bvp_batch *B) {
    BatchDev D = B->dev;
    HIP_TRY(ctx, wave_reset_counter(D.counter, ctx->stream));
    if (B->mlp) {
        if (B->anchors > 0) {
            D.anc = B->anc;
            D.anc_K = B->anchors;
        } else {
            D.init_y = B->warm_y;
        }
        D.warm = 1;
    }
    if (B->firk)
        HIP_TRY(ctx, firk_launch_solve(D, B->firk_s, B->variant, B->blocks, ctx->stream));
    else if (B->block)
        HIP_TRY(ctx, block_launch_solve(D, B->method, B->opts.precision, B->blocks, ctx->stream));
    else if (D.controller != BVP_CONTROL_DEFECT)
        HIP_TRY(ctx, wave_launch_solve_ge(D, B->method, B->blocks, ctx->stream));
    else
        HIP_TRY(ctx, wave_launch_solve(D, B->method, B->variant, B->blocks, B->family, ctx->stream));
    return 0;
}

static int capture(bvp_ctx *ctx, bvp_batch *B, int (*fn)(bvp_ctx *, bvp_batch *), hipGraph_t &g,
                   hipGraphExec_t &ge) {
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = fn(ctx, B);
    hipGraph_t gg = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &gg);
    if (rc) {
        if (gg) (void)hipGraphDestroy(gg);
        return rc;
    }
    HIP_TRY(ctx, e);
    g = gg;
    HIP_TRY(ctx, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    return 0;
}

int batch_run(bvp_ctx *ctx, bvp_batch *B) {
    if (fp64_block32_blocked(ctx, B)) return BVP_ENOTSUP;
    if (B->use_graph) {
        if (B->graph_exec && B->mlp && B->mlp_gen != B->mlp->gen) B->drop_graph();  // normalisation changed
        if (!B->graph_exec) {
            int rc = capture(ctx, B, enqueue_solve, B->graph, B->graph_exec);
            if (!rc && B->mlp) rc = capture(ctx, B, enqueue_guess, B->guess_graph, B->guess_exec);
            if (rc) {
                B->drop_graph();
                return rc;
            }
            if (B->mlp) B->mlp_gen = B->mlp->gen;
        }
    }
    B->timed_split = B->mlp != nullptr;
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (B->mlp) {
        if (B->use_graph) HIP_TRY(ctx, hipGraphLaunch(B->guess_exec, ctx->stream));
        else if (int rc = enqueue_guess(ctx, B)) return rc;
        HIP_TRY(ctx, hipEventRecord(ctx->evm, ctx->stream));
    }
    if (B->use_graph) HIP_TRY(ctx, hipGraphLaunch(B->graph_exec, ctx->stream));
    else if (int rc = enqueue_solve(ctx, B)) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    return 0;
}

int batch_sync(bvp_ctx *ctx, bvp_batch *B, float *ms) {
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    HIP_TRY(ctx, hipGetLastError());
    float t = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    B->last_ms = t;
    B->last_guess_ms = 0.f;
    B->last_solve_ms = t;
    if (B->timed_split) {
        float g = 0, s = 0;
        HIP_TRY(ctx, hipEventElapsedTime(&g, ctx->ev0, ctx->evm));
        HIP_TRY(ctx, hipEventElapsedTime(&s, ctx->evm, ctx->ev1));
        B->last_guess_ms = g;
        B->last_solve_ms = s;
    }
    if (ms) *ms = t;
    return 0;
}

int batch_results(bvp_ctx *ctx, bvp_batch *B, bvp_result *res) {
    const int n = B->n, M = B->M;
    const BatchDev &D = B->dev;
    std::vector<int> ret(n), Nf(n), it(n), ou(n), fb(n);
    std::vector<double> dn(n), rn(n), er(n);
    HIP_TRY(ctx, hipMemcpy(ret.data(), D.retcode, n * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(Nf.data(), D.N_final, n * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(it.data(), D.newton_iters, n * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(ou.data(), D.outer_iters, n * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(fb.data(), D.fallback, n * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(dn.data(), D.defect_norm, n * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(rn.data(), D.resid_norm, n * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(er.data(), D.err, n * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<long long> off(n);
    long long tot = 0;
    int maxn = 0;
    for (int i = 0; i < n; ++i) {
        off[i] = tot;
        tot += Nf[i] + 1;
        maxn = std::max(maxn, Nf[i] + 1);
    }
    DevArena tmp;
    if (B->arena.bump) {  // the gather buffers follow the plan's own part of the region
        tmp.bump = B->arena.bump + B->arena.bump_used;
        tmp.bump_cap = B->arena.bump_cap - B->arena.bump_used;
    }
    long long *doff;
    double *dm, *dy;
    int rc;
    if ((rc = upload(ctx, tmp, doff, off)) || (rc = dalloc(ctx, tmp, dm, (size_t)tot)) ||
        (rc = dalloc(ctx, tmp, dy, (size_t)tot * M)))
        return rc;
    HIP_TRY(ctx, wave_launch_gather(D, doff, dm, dy, maxn, ctx->stream));
    std::vector<double> hm((size_t)tot), hy((size_t)tot * M);
    HIP_TRY(ctx, hipMemcpyAsync(hm.data(), dm, hm.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(hy.data(), dy, hy.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        bvp_result &r = res[i];
        r.retcode = ret[i];
        r.N_final = Nf[i];
        r.newton_iters = it[i];
        r.outer_iters = ou[i];
        r.fallback_used = fb[i];
        r.defect_norm = dn[i];
        r.resid_norm = rn[i];
        r.err_vs_fp64 = er[i];
        if (r.capacity < Nf[i] + 1) { bad = 1; continue; }
        if (r.mesh) std::memcpy(r.mesh, &hm[(size_t)off[i]], sizeof(double) * (Nf[i] + 1));
        if (r.u) std::memcpy(r.u, &hy[(size_t)off[i] * M], sizeof(double) * (size_t)M * (Nf[i] + 1));
    }
    if (bad) { set_err(ctx, "result capacity smaller than the final mesh"); return BVP_EINVAL; }
    bool want_stages = false;
    for (int i = 0; i < n; ++i) want_stages |= B->firk && res[i].stages != nullptr;
    if (want_stages) {
        // final stage values, instances back to back at interval offsets
        const int S = B->firk_s;
        std::vector<long long> ioff(n);
        long long nint = 0;
        for (int i = 0; i < n; ++i) {
            ioff[i] = nint;
            nint += Nf[i];
        }
        long long *dio;
        double *dk;
        if ((rc = upload(ctx, tmp, dio, ioff)) || (rc = dalloc(ctx, tmp, dk, (size_t)std::max(1LL, nint) * S * M)))
            return rc;
        HIP_TRY(ctx, firk_launch_gather(D, S, dio, dk, ctx->stream));
        std::vector<double> hk((size_t)nint * S * M);
        HIP_TRY(ctx, hipMemcpyAsync(hk.data(), dk, hk.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < n; ++i)
            if (res[i].stages)
                std::memcpy(res[i].stages, &hk[(size_t)ioff[i] * S * M], sizeof(double) * (size_t)Nf[i] * S * M);
    }
    return 0;
}

// ------------------------------------------------------------------ C ABI
extern "C" {

void bvp_opts_default(bvp_opts *o) {
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
}

int bvp_ctx_create(int device, bvp_ctx **out) {
    if (!out) return BVP_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BVP_EHIP;
    if (device < 0 || device >= ndev) return BVP_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return BVP_EHIP;
    auto *c = new bvp_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return BVP_EHIP; }
    c->num_cus = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return BVP_EHIP; }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->evm) != hipSuccess) {
        delete c;
        return BVP_EHIP;
    }
    const char *env = std::getenv("BVP_BLOCKS_PER_CU");
    c->blocks_per_cu = env ? std::atoi(env) : 0;
    // this device's copy of the no-allocation scratch (the device is current here)
    void *sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || !sp) {
        bvp_ctx_destroy(c);
        return BVP_EHIP;
    }
    c->noalloc = (char *)sp;
    *out = c;
    return 0;
}

void bvp_ctx_destroy(bvp_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evm) (void)hipEventDestroy(c->evm);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *bvp_last_error(const bvp_ctx *c) { return c ? c->err.c_str() : "no context"; }

int bvp_mirk_solve_batch(bvp_ctx *ctx, const bvp_opts *o, int32_t n, const bvp_instance *inst, bvp_result *res) {
    if (!ctx || !o || !res) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    bvp_batch *B = nullptr;
    int rc = batch_build(ctx, o, n, inst, 0, &B, nullptr, 0);
    if (rc) return rc;
    if ((rc = batch_run(ctx, B)) == 0 && (rc = batch_sync(ctx, B, nullptr)) == 0) rc = batch_results(ctx, B, res);
    delete B;
    return rc;
}

int bvp_mirk_solve_batch_noalloc(bvp_ctx *ctx, const bvp_opts *o, int32_t n, const bvp_instance *inst,
                                 bvp_result *res) {
    if (!ctx || !o || !res) return BVP_EINVAL;
    if (!ctx->noalloc) { set_err(ctx, "no-allocation scratch unavailable on this device"); return BVP_EHIP; }
    (void)hipSetDevice(ctx->device);
    bvp_batch *B = nullptr;
    int rc = batch_build(ctx, o, n, inst, 0, &B, ctx->noalloc, kNoAllocBytes);
    if (rc) return rc;
    B->use_graph = false;  // launched directly: instantiating a graph may allocate
    if ((rc = batch_run(ctx, B)) == 0 && (rc = batch_sync(ctx, B, nullptr)) == 0) rc = batch_results(ctx, B, res);
    delete B;
    return rc;
}

// ---- plan API (extension of the ABI for repeated timed runs; see bvp_internal.h)
int bvp_batch_create(bvp_ctx *ctx, const bvp_opts *o, int32_t n, const bvp_instance *inst, bvp_batch **out) {
    if (!ctx || !o || !out) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    return batch_build(ctx, o, n, inst, 0, out, nullptr, 0);
}
int bvp_batch_run(bvp_ctx *ctx, bvp_batch *B) { return (ctx && B) ? batch_run(ctx, B) : BVP_EINVAL; }
int bvp_batch_sync(bvp_ctx *ctx, bvp_batch *B, float *ms) { return (ctx && B) ? batch_sync(ctx, B, ms) : BVP_EINVAL; }
int bvp_batch_results(bvp_ctx *ctx, bvp_batch *B, bvp_result *res) {
    return (ctx && B && res) ? batch_results(ctx, B, res) : BVP_EINVAL;
}
int bvp_batch_counters(bvp_ctx *ctx, bvp_batch *B, long long *out4) {
    if (!ctx || !B || !out4) return BVP_EINVAL;
    std::vector<long long> c((size_t)B->n * 4);
    HIP_TRY(ctx, hipMemcpy(c.data(), B->dev.cnt, c.size() * sizeof(long long), hipMemcpyDeviceToHost));
    for (int k = 0; k < 4; ++k) out4[k] = 0;
    for (int i = 0; i < B->n; ++i)
        for (int k = 0; k < 4; ++k) out4[k] += c[(size_t)i * 4 + k];
    return 0;
}
int bvp_batch_info(bvp_batch *B, int32_t *out4) {
    if (!B || !out4) return BVP_EINVAL;
    out4[0] = B->blocks;
    out4[1] = B->cap;
    out4[2] = B->M;
    out4[3] = (int32_t)B->dev.scratch_per_slot;
    return 0;
}
void bvp_batch_destroy(bvp_batch *B) { delete B; }

// ---- component entry points
static int run_component(bvp_ctx *ctx, int method, int op, int32_t n, const bvp_instance *inst, int cap_req,
                         const bvp_opts *o_in, bvp_batch **outB, bool firk_entry = false) {
    if (bvphost::is_firk(method) != firk_entry) {
        set_err(ctx, firk_entry ? "method %d is not a FIRK method" : "method %d: FIRK methods have bvp_firk_* entry points",
                method);
        return BVP_EINVAL;
    }
    bvp_opts o;
    if (o_in) o = *o_in;
    else bvp_opts_default(&o);
    o.method = method;
    if (firk_entry) {
        // the component kernels use neither adaptivity nor the nonlinear solver's options;
        // the stage state is sized for the given meshes
        o.adaptive = 0;
        o.nlsolve_stages = BVP_STAGE_NEWTON | BVP_STAGE_BACKTRACKING;
        o.precision = BVP_PREC_FP64;
        int maxN = 1;
        for (int i = 0; i < n && inst; ++i) maxN = std::max(maxN, inst[i].N);
        o.max_num_subintervals = maxN;
    }
    int rc = batch_build(ctx, &o, n, inst, cap_req, outB, nullptr, 0);
    if (rc) return rc;
    // a component call computes in the precision it asks for, with the default chunk
    // mapping, whatever the autotuner chose for plans of this kind
    bvp_batch *B = *outB;
    if (!B->block) {
        B->dev.f32rec = (o.precision == BVP_PREC_FP32_REFINE64) ? 1 : 0;
        B->dev.min_chunk = 1;
    }
    return 0;
}

static int launch_op(bvp_ctx *ctx, bvp_batch *B, int op) {
    HIP_TRY(ctx, wave_reset_counter(B->dev.counter, ctx->stream));
    if (B->block)
        HIP_TRY(ctx, block_launch_component(B->dev, B->method, B->opts.precision, op, B->blocks, ctx->stream));
    else
        HIP_TRY(ctx, wave_launch_component(B->dev, B->method, op, B->blocks, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int bvp_mirk_residual(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst, double *resid) {
    if (!ctx || !resid) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, method, OP_RESIDUAL, n, inst, 0, nullptr, &B);
    if (rc) return rc;
    const size_t tot = (size_t)B->total_nodes * B->M;
    if ((rc = dalloc(ctx, B->arena, B->dev.out0, tot)) == 0 && (rc = launch_op(ctx, B, OP_RESIDUAL)) == 0) {
        hipError_t e = hipMemcpy(resid, B->dev.out0, tot * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "copy back: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
    }
    delete B;
    return rc;
}

int bvp_mirk_jacobian_blocks(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst, double *S, double *R,
                             int32_t *bc_node) {
    if (!ctx || !S || !R || !bc_node) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, method, OP_JACOBIAN, n, inst, 0, nullptr, &B);
    if (rc) return rc;
    const int M = B->M;
    const size_t nint = (size_t)(B->total_nodes - n);
    if ((rc = dalloc(ctx, B->arena, B->dev.out0, nint * M * M)) == 0 &&
        (rc = dalloc(ctx, B->arena, B->dev.out1, nint * M * M)) == 0 &&
        (rc = dalloc(ctx, B->arena, B->dev.iout, (size_t)n * M)) == 0 && (rc = launch_op(ctx, B, OP_JACOBIAN)) == 0) {
        hipError_t e = hipMemcpy(S, B->dev.out0, nint * M * M * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(R, B->dev.out1, nint * M * M * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(bc_node, B->dev.iout, (size_t)n * M * sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "copy back: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
    }
    delete B;
    return rc;
}

int bvp_mirk_newton_step(bvp_ctx *ctx, int32_t method, int32_t precision, int32_t refine_iters, int32_t n,
                         const bvp_instance *inst, double *delta) {
    if (!ctx || !delta) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    bvp_opts o;
    bvp_opts_default(&o);
    o.precision = precision;
    o.refine_iters = refine_iters;
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, method, OP_NEWTON_STEP, n, inst, 0, &o, &B);
    if (rc) return rc;
    if (fp64_block32_blocked(ctx, B)) { delete B; return BVP_ENOTSUP; }
    const size_t tot = (size_t)B->total_nodes * B->M;
    std::vector<int> ok(n);
    if ((rc = dalloc(ctx, B->arena, B->dev.out0, tot)) == 0 && (rc = dalloc(ctx, B->arena, B->dev.iout, n)) == 0 &&
        (rc = launch_op(ctx, B, OP_NEWTON_STEP)) == 0) {
        hipError_t e = hipMemcpy(delta, B->dev.out0, tot * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(ok.data(), B->dev.iout, n * sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "copy back: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
        for (int i = 0; i < n && rc == 0; ++i)
            if (ok[i]) { set_err(ctx, "instance %d: singular Newton system", i); rc = 1; }
    }
    delete B;
    return rc;
}

int bvp_mirk_defect(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst, double *errors,
                    double *defect_norm) {
    if (!ctx || !errors || !defect_norm) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, method, OP_DEFECT, n, inst, 0, nullptr, &B);
    if (rc) return rc;
    const size_t nint = (size_t)(B->total_nodes - n);
    if ((rc = dalloc(ctx, B->arena, B->dev.out0, nint * B->M)) == 0 && (rc = dalloc(ctx, B->arena, B->dev.out1, n)) == 0 &&
        (rc = launch_op(ctx, B, OP_DEFECT)) == 0) {
        hipError_t e = hipMemcpy(errors, B->dev.out0, nint * B->M * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(defect_norm, B->dev.out1, n * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "copy back: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
    }
    delete B;
    return rc;
}

int bvp_mirk_global_error(bvp_ctx *ctx, const bvp_opts *opts, int32_t n, const bvp_instance *inst, double *errors,
                          double *norm) {
    if (!ctx || !opts || !errors || !norm) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    bvp_opts o = *opts;
    if (o.controller == BVP_CONTROL_DEFECT) o.controller = BVP_CONTROL_GLOBAL;  // sizes the global-error buffers
    o.precision = BVP_PREC_FP64;
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, o.method, -1, n, inst, 0, &o, &B);
    if (rc) return rc;
    const size_t nint = (size_t)(B->total_nodes - n);
    if ((rc = dalloc(ctx, B->arena, B->dev.out0, nint * B->M)) == 0 && (rc = dalloc(ctx, B->arena, B->dev.out1, n)) == 0) {
        hipError_t e = wave_reset_counter(B->dev.counter, ctx->stream);
        if (e == hipSuccess) e = wave_launch_global_error(B->dev, B->method, B->blocks, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(errors, B->dev.out0, nint * B->M * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(norm, B->dev.out1, n * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "global error: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
    }
    delete B;
    return rc;
}

int bvp_mirk_interpolate(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst, const int32_t *n_t,
                         const double *t, double *u, double *du) {
    if (!ctx || !n_t || !u) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    long long nq = 0;
    for (int i = 0; i < n; ++i) {
        if (n_t[i] < 0) { set_err(ctx, "instance %d: negative point count", i); return BVP_EINVAL; }
        nq += n_t[i];
    }
    if (nq > 0 && !t) return BVP_EINVAL;
    if (n > 0 && !wave_supports(inst[0].M, inst[0].family)) {
        set_err(ctx, "dense output serves the wave-kernel families (M = 1..4, 6)");
        return BVP_ENOTSUP;
    }
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, method, OP_RESIDUAL, n, inst, 0, nullptr, &B);
    if (rc) return rc;
    if (B->block) {
        delete B;
        set_err(ctx, "dense output serves the wave-kernel families (M = 1..4, 6)");
        return BVP_ENOTSUP;
    }
    std::vector<int> qi((size_t)nq);
    for (int i = 0, q = 0; i < n; ++i)
        for (int k = 0; k < n_t[i]; ++k) qi[q++] = i;
    std::vector<double> tv(t, t + nq);
    int *dqi = nullptr;
    double *dt = nullptr, *du_ = nullptr, *ddu = nullptr;
    const int M = B->M;
    if (nq > 0 &&
        ((rc = upload(ctx, B->arena, dqi, qi)) || (rc = upload(ctx, B->arena, dt, tv)) ||
         (rc = dalloc(ctx, B->arena, du_, (size_t)nq * M)) || (du && (rc = dalloc(ctx, B->arena, ddu, (size_t)nq * M))))) {
        delete B;
        return rc;
    }
    if (nq > 0) {
        hipError_t e = wave_launch_interp(B->dev, method, dqi, dt, du_, ddu, nq, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(u, du_, (size_t)nq * M * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && du) e = hipMemcpy(du, ddu, (size_t)nq * M * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "interpolate: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
    }
    delete B;
    return rc;
}

// ---- FIRK components
// The stage values of the component kernels: the caller's, or extend_y's
// (K_ij = y_{i+1} from the instance's guess or u0), packed at interval offsets.
static int firk_upload_stages(bvp_ctx *ctx, bvp_batch *B, int32_t n, const bvp_instance *inst, const double *stages) {
    if (B->dev.firk_nested) return 0;  // the nested form solves its stages from k0
    const int M = B->M, S = B->firk_s;
    const long long nint = B->total_nodes - n;
    std::vector<double> K((size_t)std::max(1LL, nint) * S * M, 0.0);
    if (stages) {
        std::memcpy(K.data(), stages, sizeof(double) * (size_t)nint * S * M);
    } else {
        long long o = 0;
        for (int i = 0; i < n; ++i) {
            const bvp_instance &P = inst[i];
            for (int k = 0; k < P.N; ++k)
                for (int r = 0; r < S; ++r)
                    for (int c = 0; c < M; ++c)
                        K[((size_t)(o + k) * S + r) * M + c] = P.guess ? P.guess[(size_t)(k + 1) * M + c] : P.u0[c];
            o += P.N;
        }
    }
    double *dk = nullptr;
    int rc = upload(ctx, B->arena, dk, K);
    if (rc) return rc;
    B->dev.init_K = dk;
    return 0;
}

static int firk_component(bvp_ctx *ctx, int32_t method, int op, int32_t n, const bvp_instance *inst,
                          const double *stages, size_t out0_count, double *out0, double *out1) {
    (void)hipSetDevice(ctx->device);
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, method, op, n, inst, 0, nullptr, &B, true);
    if (rc) return rc;
    if ((rc = firk_upload_stages(ctx, B, n, inst, stages)) == 0 &&
        (rc = dalloc(ctx, B->arena, B->dev.out0, std::max<size_t>(1, out0_count))) == 0 &&
        (!out1 || (rc = dalloc(ctx, B->arena, B->dev.out1, n)) == 0)) {
        hipError_t e = wave_reset_counter(B->dev.counter, ctx->stream);
        if (e == hipSuccess) e = firk_launch_component(B->dev, B->firk_s, op, B->blocks, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(out0, B->dev.out0, out0_count * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && out1) e = hipMemcpy(out1, B->dev.out1, n * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "FIRK component: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
    }
    delete B;
    return rc;
}

// doubles of the unknown vectors of a batch: sum_i M (N_i (s + 1) + 1) (expanded) or
// sum_i M (N_i + 1) (nested)
static size_t firk_nx_total(int method, int32_t n, const bvp_instance *inst) {
    size_t t = 0;
    const int S = method >= BVP_FIRK_NESTED ? 0 : method % 100;
    for (int i = 0; i < n; ++i) t += (size_t)inst[i].M * ((size_t)inst[i].N * (S + 1) + 1);
    return t;
}

int bvp_firk_residual(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst, const double *stages,
                      double *resid) {
    if (!ctx || !resid || !inst || n <= 0) return BVP_EINVAL;
    return firk_component(ctx, method, OP_RESIDUAL, n, inst, stages, firk_nx_total(method, n, inst), resid, nullptr);
}

int bvp_firk_newton_step(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst, const double *stages,
                         double *delta) {
    if (!ctx || !delta || !inst || n <= 0) return BVP_EINVAL;
    return firk_component(ctx, method, OP_NEWTON_STEP, n, inst, stages, firk_nx_total(method, n, inst), delta,
                          nullptr);
}

int bvp_firk_defect(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst, const double *stages,
                    double *errors, double *defect_norm) {
    if (!ctx || !errors || !defect_norm || !inst || n <= 0) return BVP_EINVAL;
    size_t nint = 0;
    for (int i = 0; i < n; ++i) nint += (size_t)inst[i].N * inst[i].M;
    return firk_component(ctx, method, OP_DEFECT, n, inst, stages, nint, errors, defect_norm);
}

int bvp_firk_interpolate(bvp_ctx *ctx, int32_t method, int32_t n, const bvp_instance *inst, const double *stages,
                         const int32_t *n_t, const double *t, double *u, double *du) {
    if (!ctx || !n_t || !u || !inst || n <= 0) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    long long nq = 0;
    for (int i = 0; i < n; ++i) {
        if (n_t[i] < 0) { set_err(ctx, "instance %d: negative point count", i); return BVP_EINVAL; }
        nq += n_t[i];
    }
    if (nq > 0 && !t) return BVP_EINVAL;
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, method, OP_RESIDUAL, n, inst, 0, nullptr, &B, true);
    if (rc) return rc;
    if ((rc = firk_upload_stages(ctx, B, n, inst, stages))) {
        delete B;
        return rc;
    }
    std::vector<int> qi((size_t)nq);
    for (int i = 0, q = 0; i < n; ++i)
        for (int k = 0; k < n_t[i]; ++k) qi[q++] = i;
    std::vector<double> tv(t, t + nq);
    int *dqi = nullptr;
    double *dt = nullptr, *du_ = nullptr, *ddu = nullptr;
    const int M = B->M;
    if (nq > 0 &&
        ((rc = upload(ctx, B->arena, dqi, qi)) || (rc = upload(ctx, B->arena, dt, tv)) ||
         (rc = dalloc(ctx, B->arena, du_, (size_t)nq * M)) || (du && (rc = dalloc(ctx, B->arena, ddu, (size_t)nq * M))))) {
        delete B;
        return rc;
    }
    if (nq > 0) {
        hipError_t e = firk_launch_interp(B->dev, B->firk_s, dqi, dt, du_, ddu, nq, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(u, du_, (size_t)nq * M * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && du) e = hipMemcpy(du, ddu, (size_t)nq * M * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "FIRK interpolate: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
    }
    delete B;
    return rc;
}

int bvp_mirk_mesh_select(bvp_ctx *ctx, const bvp_opts *opts, int32_t n, const bvp_instance *inst, int32_t capacity,
                         int32_t *new_N, int32_t *info, double *new_mesh, double *new_y) {
    if (!ctx || !opts || !new_N || !info || !new_mesh || !new_y) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    bvp_opts o = *opts;
    if (capacity < o.max_num_subintervals + 1) {
        set_err(ctx, "capacity must be at least max_num_subintervals + 1");
        return BVP_EINVAL;
    }
    bvp_batch *B = nullptr;
    int rc = run_component(ctx, o.method, OP_MESH_SELECT, n, inst, capacity, &o, &B);
    if (rc) return rc;
    const int M = B->M, cap = B->cap;
    if ((rc = dalloc(ctx, B->arena, B->dev.out0, (size_t)n * cap)) == 0 &&
        (rc = dalloc(ctx, B->arena, B->dev.out1, (size_t)n * cap * M)) == 0 &&
        (rc = dalloc(ctx, B->arena, B->dev.iout, (size_t)n * 2)) == 0 && (rc = launch_op(ctx, B, OP_MESH_SELECT)) == 0) {
        std::vector<int> io((size_t)n * 2);
        std::vector<double> hm((size_t)n * cap), hy((size_t)n * cap * M);
        hipError_t e = hipMemcpy(io.data(), B->dev.iout, io.size() * sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(hm.data(), B->dev.out0, hm.size() * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(hy.data(), B->dev.out1, hy.size() * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_err(ctx, "copy back: %s", hipGetErrorString(e)); rc = BVP_EHIP; }
        for (int i = 0; i < n && rc == 0; ++i) {
            new_N[i] = io[(size_t)i * 2];
            info[i] = io[(size_t)i * 2 + 1];
            if (info[i] == BVP_RET_SUCCESS) {
                std::memcpy(new_mesh + (size_t)i * capacity, &hm[(size_t)i * cap], sizeof(double) * (new_N[i] + 1));
                std::memcpy(new_y + (size_t)i * capacity * M, &hy[(size_t)i * cap * M],
                            sizeof(double) * (size_t)M * (new_N[i] + 1));
            }
        }
    }
    delete B;
    return rc;
}

// Streams `nbytes` (rounded down to doubles) from one device buffer to another
// with a plain copy kernel and reports its device time; the reference point for
// converting HBM byte counters of a profiler into bytes.
int bvp_copy_probe(bvp_ctx *ctx, long long nbytes, int repeats, float *ms) {
    if (!ctx || nbytes < 8 || repeats < 1) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    const size_t n = (size_t)nbytes / 8;
    DevArena A;
    double *src, *dst;
    int rc;
    if ((rc = dalloc(ctx, A, src, n)) || (rc = dalloc(ctx, A, dst, n))) return rc;
    HIP_TRY(ctx, hipMemsetAsync(src, 0, n * 8, ctx->stream));
    float best = 1e30f;
    for (int r = 0; r < repeats; ++r) {
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIP_TRY(ctx, launch_copy_probe(src, dst, n, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
        float t = 0;
        HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        best = std::min(best, t);
    }
    if (ms) *ms = best;
    return 0;
}

int bvp_batch_learned_guess(bvp_ctx *ctx, bvp_batch *B, double *y) {
    if (!ctx || !B || !y) return BVP_EINVAL;
    if (!B->mlp) { set_err(ctx, "no network attached to the plan"); return BVP_EINVAL; }
    (void)hipSetDevice(ctx->device);
    if (int rc = enqueue_guess(ctx, B)) return rc;
    if (B->anchors > 0) HIP_TRY(ctx, mlp::launch_anchor_fill(B->dev, B->anc, B->anchors, B->warm_y, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(y, B->warm_y, sizeof(double) * (size_t)B->total_nodes * B->M,
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int bvp_batch_autotune(bvp_ctx *ctx, bvp_batch *B, int32_t repeats, float *cand_ms, int32_t *n_cand) {
    if (!ctx || !B || repeats < 1) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    struct Cand { int variant, blocks, f32, minc; };
    std::vector<Cand> cands;
    if (B->firk) {
        // the two register variants, each at 1, 2, 4, ... waves per CU up to its occupancy
        for (int v = 0; v < kFirkVariants; ++v) {
            const int occ = std::max(1, firk_blocks_per_cu(B->M, B->firk_s, v, B->dev.firk_nested));
            for (int pc = 1; pc <= occ; pc *= 2) cands.push_back({v, std::min(B->max_slots, ctx->num_cus * pc), 0, 1});
            if (cands.back().blocks != std::min(B->max_slots, ctx->num_cus * occ))
                cands.push_back({v, std::min(B->max_slots, ctx->num_cus * occ), 0, 1});
        }
    } else if (B->block) {
        const int occ = std::max(1, block_blocks_per_cu(B->M, B->method, B->opts.precision));
        for (int pc = 1; pc <= occ; pc *= 2) cands.push_back({0, std::min(B->max_slots, ctx->num_cus * pc), 0, 1});
        if (cands.back().blocks != std::min(B->max_slots, ctx->num_cus * occ))
            cands.push_back({0, std::min(B->max_slots, ctx->num_cus * occ), 0, 1});
    } else {
        // FP32_REFINE64 on the wave kernels: FP32 records first, then FP64 records (always
        // at least as accurate; faster where the inexact corrections cost Newton steps)
        // (the global-error kernel: one variant, FP64 records)
        const bool ge = B->dev.controller != BVP_CONTROL_DEFECT;
        const bool mixed = B->opts.precision == BVP_PREC_FP32_REFINE64 && !ge && B->M <= 4;
        for (int f = mixed ? 1 : 0; f >= 0; --f)
            for (int v = 0; v < (ge ? 1 : wave_variants(B->M)); ++v)
                cands.push_back({v, std::min(B->max_slots,
                                             ctx->num_cus * std::max(1, wave_blocks_per_cu(B->M, B->method, v, B->family, f))),
                                 f, 1});
    }
    const int keep_v = B->variant, keep_b = B->blocks, keep_f = B->dev.f32rec, keep_m = B->dev.min_chunk;
    int best = -1;
    float best_ms = 1e30f;
    auto time_cand = [&](const Cand &cd, float &med) -> int {
        B->variant = cd.variant;
        B->blocks = std::max(1, cd.blocks);
        if (!B->block) {
            B->dev.f32rec = cd.f32;
            B->dev.min_chunk = cd.minc;
        }
        B->drop_graph();
        std::vector<float> t;
        for (int r = 0; r <= repeats; ++r) {  // first run untimed
            int rc = batch_run(ctx, B);
            if (!rc) rc = batch_sync(ctx, B, nullptr);
            if (rc) return rc;
            if (r > 0) t.push_back(B->last_solve_ms);
        }
        std::sort(t.begin(), t.end());
        med = t[t.size() / 2];
        return 0;
    };
    // phase 1: kernel variant and record precision; phase 2 (wave kernels): the
    // interval-per-lane mapping of the fastest one (at least 2, 4, 8 intervals per lane)
    for (int phase = 0; phase < ((B->block || B->firk) ? 1 : 2); ++phase) {
        if (phase == 1) {
            const Cand b0 = cands[best];
            for (int m = 2; m <= kMaxMinChunk; m *= 2) cands.push_back({b0.variant, b0.blocks, b0.f32, m});
        }
        for (size_t c = (phase == 0 ? 0 : cands.size() - 3); c < cands.size(); ++c) {
            float med = 0.f;
            const int rc = time_cand(cands[c], med);
            if (rc) {
                B->variant = keep_v;
                B->blocks = keep_b;
                B->dev.f32rec = keep_f;
                B->dev.min_chunk = keep_m;
                B->drop_graph();
                return rc;
            }
            if (cand_ms && (int)c < 16) cand_ms[c] = med;
            if (med < best_ms) {
                best_ms = med;
                best = (int)c;
            }
        }
    }
    if (n_cand) *n_cand = (int32_t)std::min<size_t>(cands.size(), 16);
    B->variant = cands[best].variant;
    B->blocks = std::max(1, cands[best].blocks);
    if (!B->block) {
        B->dev.f32rec = cands[best].f32;
        B->dev.min_chunk = cands[best].minc;
    }
    B->drop_graph();
    const int per_cu = (B->blocks + ctx->num_cus - 1) / ctx->num_cus;
    ctx->tuned[std::make_tuple(B->M, B->method, B->opts.precision,
                               B->block ? 1 : (B->firk ? 3 : (B->dev.controller != BVP_CONTROL_DEFECT ? 2 : 0)),
                               B->family)] =
        std::make_tuple(B->variant, per_cu, (int)B->dev.f32rec, (int)B->dev.min_chunk);
    return 0;
}

int bvp_batch_chunking(bvp_batch *B, int32_t *min_chunk) {
    if (!B || !min_chunk) return BVP_EINVAL;
    *min_chunk = B->dev.min_chunk;
    return 0;
}

int bvp_batch_set_chunking(bvp_batch *B, int32_t min_chunk) {
    if (!B || min_chunk < 1 || min_chunk > kMaxMinChunk || ((B->block || B->firk) && min_chunk != 1))
        return BVP_EINVAL;
    B->dev.min_chunk = min_chunk;
    B->drop_graph();
    return 0;
}

int bvp_batch_precision(bvp_batch *B, int32_t *precision) {
    if (!B || !precision) return BVP_EINVAL;
    *precision = B->block ? B->opts.precision : (B->dev.f32rec ? BVP_PREC_FP32_REFINE64 : BVP_PREC_FP64);
    return 0;
}

int bvp_batch_set_config(bvp_ctx *ctx, bvp_batch *B, int32_t variant, int32_t blocks) {
    if (!ctx || !B) return BVP_EINVAL;
    const int nv = B->block ? 1 : (B->firk ? kFirkVariants : wave_variants(B->M));
    if (variant < 0 || variant >= nv || blocks < 1 || blocks > B->max_slots) {
        set_err(ctx, "variant must be in [0, %d) and blocks in [1, %d]", nv, B->max_slots);
        return BVP_EINVAL;
    }
    B->variant = variant;
    B->blocks = blocks;
    B->drop_graph();
    return 0;
}

int bvp_batch_config(bvp_batch *B, int32_t *variant, int32_t *blocks) {
    if (!B || !variant || !blocks) return BVP_EINVAL;
    *variant = B->variant;
    *blocks = B->blocks;
    return 0;
}

int bvp_batch_timing(bvp_batch *B, float *guess_ms, float *solve_ms) {
    if (!B || !guess_ms || !solve_ms) return BVP_EINVAL;
    *guess_ms = B->last_guess_ms;
    *solve_ms = B->last_solve_ms;
    return 0;
}

int bvp_batch_set_graph(bvp_batch *B, int32_t enable) {
    if (!B) return BVP_EINVAL;
    B->drop_graph();
    B->use_graph = enable != 0;
    return 0;
}

// ---- learned warm start
static uint64_t mix64(uint64_t &s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unit01(uint64_t &s) { return (double)(mix64(s) >> 11) * (1.0 / 9007199254740992.0); }

void bvp_mlp_destroy(bvp_mlp *m) {
    if (!m) return;
    if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
}

static int mlp_clear_state(bvp_mlp *m) {
    bvp_ctx *ctx = m->ctx;
    HIP_TRY(ctx, hipMemsetAsync(m->nb.m, 0, sizeof(float) * mlp::NPARAM, ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(m->nb.v, 0, sizeof(float) * mlp::NPARAM, ctx->stream));
    HIP_TRY(ctx, mlp::launch_quantize(m->nb, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int bvp_mlp_set_params(bvp_mlp *m, const float *w) {
    if (!m || !w) return BVP_EINVAL;
    bvp_ctx *ctx = m->ctx;
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipMemcpyAsync(m->nb.w, w, sizeof(float) * mlp::NPARAM, hipMemcpyHostToDevice, ctx->stream));
    return mlp_clear_state(m);
}

int bvp_mlp_create(bvp_ctx *ctx, uint64_t seed, bvp_mlp **out) {
    if (!ctx || !out) return BVP_EINVAL;
    (void)hipSetDevice(ctx->device);
    *out = nullptr;
    auto *m = new bvp_mlp();
    m->ctx = ctx;
    int rc;
    mlp::NetBufs &nb = m->nb;
    if ((rc = dalloc(ctx, m->arena, nb.w, mlp::NPARAM)) || (rc = dalloc(ctx, m->arena, nb.m, mlp::NPARAM)) ||
        (rc = dalloc(ctx, m->arena, nb.v, mlp::NPARAM)) || (rc = dalloc(ctx, m->arena, nb.q, mlp::NPARAM)) ||
        (rc = dalloc(ctx, m->arena, nb.w2f, (size_t)mlp::H * mlp::H)) ||
        (rc = dalloc(ctx, m->arena, nb.w3f, (size_t)mlp::H * mlp::H)) ||
        (rc = dalloc(ctx, m->arena, nb.w2b, (size_t)mlp::H * mlp::H)) ||
        (rc = dalloc(ctx, m->arena, nb.w3b, (size_t)mlp::H * mlp::H))) {
        delete m;
        return rc;
    }
    // Glorot-uniform matrices, zero biases
    std::vector<float> w(mlp::NPARAM, 0.f);
    uint64_t s = seed;
    auto fill = [&](int off, int fin, int fout) {
        const double r = std::sqrt(6.0 / (fin + fout));
        for (int e = 0; e < fin * fout; ++e) w[off + e] = (float)((2.0 * unit01(s) - 1.0) * r);
    };
    fill(mlp::OW1, mlp::NIN, mlp::H);
    fill(mlp::OW2, mlp::H, mlp::H);
    fill(mlp::OW3, mlp::H, mlp::H);
    fill(mlp::OW4, mlp::H, mlp::NOUT);
    if ((rc = bvp_mlp_set_params(m, w.data()))) {
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

int bvp_mlp_set_normalization(bvp_mlp *m, const float *in_shift, const float *in_scale, const float *out_shift,
                              const float *out_scale) {
    if (!m || !in_shift || !in_scale || !out_shift || !out_scale) return BVP_EINVAL;
    for (int i = 0; i < mlp::NIN; ++i) {
        if (!std::isfinite(in_shift[i]) || !std::isfinite(in_scale[i]) || in_scale[i] == 0.f) {
            set_err(m->ctx, "bad input normalisation");
            return BVP_EINVAL;
        }
        m->in_shift[i] = in_shift[i];
        m->in_scale[i] = in_scale[i];
    }
    for (int o = 0; o < mlp::NOUT; ++o) {
        if (!std::isfinite(out_shift[o]) || !std::isfinite(out_scale[o]) || out_scale[o] == 0.f) {
            set_err(m->ctx, "bad output normalisation");
            return BVP_EINVAL;
        }
        m->out_shift[o] = out_shift[o];
        m->out_scale[o] = out_scale[o];
    }
    m->gen++;
    return 0;
}

int bvp_mlp_get_params(bvp_mlp *m, float *w) {
    if (!m || !w) return BVP_EINVAL;
    bvp_ctx *ctx = m->ctx;
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(w, m->nb.w, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost));
    return 0;
}

int bvp_mlp_gradient(bvp_mlp *m, int32_t n, const float *x, const float *y, float *grad, float *loss) {
    if (!m || n < 1 || !x || !y || !grad) return BVP_EINVAL;
    bvp_ctx *ctx = m->ctx;
    (void)hipSetDevice(ctx->device);
    DevArena A;
    float *dx, *dy, *dpart, *dg;
    const int tiles = (n + mlp::T - 1) / mlp::T;
    int rc;
    if ((rc = dalloc(ctx, A, dx, (size_t)n * mlp::NIN)) || (rc = dalloc(ctx, A, dy, (size_t)n * mlp::NOUT)) ||
        (rc = dalloc(ctx, A, dpart, (size_t)tiles * mlp::NSLOT)) || (rc = dalloc(ctx, A, dg, mlp::NSLOT)))
        return rc;
    HIP_TRY(ctx, hipMemcpyAsync(dx, x, sizeof(float) * n * mlp::NIN, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dy, y, sizeof(float) * n * mlp::NOUT, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, mlp::launch_grad(m->net(), dx, dy, n, dpart, dg, ctx->stream));
    std::vector<float> h(mlp::NSLOT);
    HIP_TRY(ctx, hipMemcpyAsync(h.data(), dg, sizeof(float) * mlp::NSLOT, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    std::memcpy(grad, h.data(), sizeof(float) * mlp::NPARAM);
    if (loss) *loss = h[mlp::NPARAM];
    return 0;
}

int bvp_mlp_predict(bvp_mlp *m, int32_t n, const float *x, float *y) {
    if (!m || n < 1 || !x || !y) return BVP_EINVAL;
    bvp_ctx *ctx = m->ctx;
    (void)hipSetDevice(ctx->device);
    DevArena A;
    float *dx, *dy;
    int rc;
    if ((rc = dalloc(ctx, A, dx, (size_t)n * mlp::NIN)) || (rc = dalloc(ctx, A, dy, (size_t)n * mlp::NOUT))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(dx, x, sizeof(float) * n * mlp::NIN, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, mlp::launch_predict(m->net(), dx, n, dy, ctx->num_cus, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(y, dy, sizeof(float) * n * mlp::NOUT, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

static long long gcd_ll(long long a, long long b) {
    while (b) { const long long t = a % b; a = b; b = t; }
    return a;
}

int bvp_mlp_train(bvp_mlp *m, int64_t n, const float *x, const float *y, const bvp_train_opts *o,
                  float *loss_per_step, int32_t *steps, float *ms) {
    if (!m || !x || !y || !o) return BVP_EINVAL;
    bvp_ctx *ctx = m->ctx;
    (void)hipSetDevice(ctx->device);
    if (o->batch < mlp::T || o->batch % mlp::T) { set_err(ctx, "batch must be a positive multiple of 64"); return BVP_EINVAL; }
    if (n < o->batch || n >= (1ll << 31)) { set_err(ctx, "need batch <= n < 2^31 samples"); return BVP_EINVAL; }
    if (o->epochs < 1 || !(o->lr > 0)) { set_err(ctx, "need epochs >= 1 and lr > 0"); return BVP_EINVAL; }
    const int all_epochs = o->total_epochs > 0 ? o->total_epochs : o->epochs;
    if (o->first_epoch < 0 || o->first_epoch + o->epochs > all_epochs) {
        set_err(ctx, "need 0 <= first_epoch and first_epoch + epochs <= total_epochs");
        return BVP_EINVAL;
    }
    const int spe = (int)(n / o->batch);
    const long long total = (long long)spe * o->epochs;          // steps of this call
    const long long sched = (long long)spe * all_epochs;         // steps of the whole schedule
    const long long step0 = (long long)spe * o->first_epoch;
    if (sched >= (1ll << 31)) { set_err(ctx, "too many steps"); return BVP_EINVAL; }
    // per-epoch sample order j -> (a j + c) mod n with gcd(a, n) = 1, drawn from a
    // stream of its own per epoch (the same for any split of the schedule)
    std::vector<unsigned long long> perm(2 * (size_t)all_epochs, 1ull);
    for (int e = o->first_epoch; e < o->first_epoch + o->epochs; ++e) {
        uint64_t s = o->seed ^ 0x6A09E667F3BCC909ull ^ ((uint64_t)(e + 1) * 0xD1B54A32D192ED03ull);
        long long a = 1;
        if (n > 2)
            do { a = 1 + (long long)(unit01(s) * (double)(n - 1)); } while (a >= n || gcd_ll(a, n) != 1);
        perm[2 * e] = (unsigned long long)a;
        perm[2 * e + 1] = (unsigned long long)(unit01(s) * (double)n) % (unsigned long long)n;
    }
    DevArena A;
    float *dx, *dy, *dpart, *dloss;
    unsigned long long *dperm;
    int *dstep;
    const int tiles = o->batch / mlp::T;
    int rc;
    if ((rc = dalloc(ctx, A, dx, (size_t)n * mlp::NIN)) || (rc = dalloc(ctx, A, dy, (size_t)n * mlp::NOUT)) ||
        (rc = dalloc(ctx, A, dpart, (size_t)tiles * mlp::NSLOT)) || (rc = dalloc(ctx, A, dloss, (size_t)total)) ||
        (rc = upload(ctx, A, dperm, perm)) || (rc = dalloc(ctx, A, dstep, 1)))
        return rc;
    HIP_TRY(ctx, hipMemcpyAsync(dx, x, sizeof(float) * n * mlp::NIN, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dy, y, sizeof(float) * n * mlp::NOUT, hipMemcpyHostToDevice, ctx->stream));
    {
        const int st0 = (int)step0;
        HIP_TRY(ctx, hipMemcpyAsync(dstep, &st0, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    mlp::TrainDev td{};
    td.x = dx;
    td.y = dy;
    td.n = n;
    td.batch_tiles = tiles;
    td.steps_per_epoch = spe;
    td.total_steps = (int)sched;
    td.step0 = (int)step0;
    td.perm = dperm;
    td.step = dstep;
    td.partial = dpart;
    td.loss = dloss;
    td.lr = o->lr;
    td.lr_min = o->lr_min;
    td.beta1 = 0.9f;
    td.beta2 = 0.999f;
    td.eps = 1e-8f;
    td.weight_decay = o->weight_decay;
    td.warmup = o->warmup;
    const mlp::Net net = m->net();
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (o->use_graph) {
        HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t e1 = mlp::launch_train_step(m->nb, net, td, ctx->stream);
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            set_err(ctx, "capture of the training step failed: %s",
                    hipGetErrorString(e1 != hipSuccess ? e1 : e2));
            return BVP_EHIP;
        }
        if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            set_err(ctx, "graph instantiation failed");
            return BVP_EHIP;
        }
    }
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    for (long long k = 0; k < total && e == hipSuccess; ++k)
        e = o->use_graph ? hipGraphLaunch(ge, ctx->stream) : mlp::launch_train_step(m->nb, net, td, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    HIP_TRY(ctx, e);
    if (loss_per_step)
        HIP_TRY(ctx, hipMemcpy(loss_per_step, dloss, sizeof(float) * total, hipMemcpyDeviceToHost));
    if (steps) *steps = (int32_t)total;
    if (ms) *ms = t;
    return 0;
}

int bvp_batch_set_warm_start(bvp_batch *B, const bvp_mlp *m, int32_t kind, int32_t anchors) {
    if (!B) return BVP_EINVAL;
    if (!m) {
        B->mlp = nullptr;
        B->drop_graph();
        return 0;
    }
    bvp_ctx *ctx = m->ctx;
    if (B->block || B->M != 2) { set_err(ctx, "the learned warm start serves M = 2 plans"); return BVP_ENOTSUP; }
    if (!B->has_u0) { set_err(ctx, "warm start: every instance needs u0 (the fallback guess)"); return BVP_EINVAL; }
    if (kind != BVP_FEAT_PARAM && kind != BVP_FEAT_LOG10_PARAM) { set_err(ctx, "bad feature kind"); return BVP_EINVAL; }
    if (anchors == 1 || anchors < 0 || anchors > 4096) {
        set_err(ctx, "anchors must be 0 (every node) or 2..4096");
        return BVP_EINVAL;
    }
    if (!B->warm_y) {
        const int rc = dalloc(ctx, B->arena, B->warm_y, (size_t)B->total_nodes * B->M);
        if (rc) return rc;
    }
    if (anchors > B->anc_cap) {  // grows only; the arena frees with the plan
        const int rc = dalloc(ctx, B->arena, B->anc, (size_t)B->n * anchors * 4);
        if (rc) return rc;
        B->anc_cap = anchors;
    }
    B->anchors = anchors;
    B->mlp = m;
    B->feature_kind = kind;
    B->drop_graph();
    return 0;
}

}  // extern "C"
