// bvp_internal.h -- host-side state behind the opaque handles of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/bvp_mirk.h"
#include "mirk_wave.h"
#include "mlp.h"

struct bvp_ctx {
    int device = 0;
    int num_cus = 0;
    std::string arch;
    hipStream_t stream = nullptr;  // default priority
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evm = nullptr;      // between the learned guess and the solve of a plan run
    int blocks_per_cu = 0;         // 0: from the occupancy calculator
    std::string err;
    // device address of the no-allocation entry point's scratch array on this
    // context's device (hipGetSymbolAddress at creation)
    char *noalloc = nullptr;
    // autotuner choices, (M, method, precision, kernel kind (0 wave, 1 block, 2 wave
    // under a global-error controller), family or -1) -> (variant,
    // waves per CU, FP32 records, minimum intervals per lane); later plans of the same
    // kind start from them
    std::map<std::tuple<int, int, int, int, int>, std::tuple<int, int, int, int>> tuned;
};

// device allocations owned by one batch; freed together.  In bump mode (bump set)
// allocations are carved from a region the arena does not own and nothing is freed.
struct DevArena {
    std::vector<void *> ptrs;
    char *bump = nullptr;
    size_t bump_cap = 0, bump_used = 0;
    DevArena() = default;
    DevArena(const DevArena &) = delete;
    DevArena &operator=(const DevArena &) = delete;
    ~DevArena();
};

struct bvp_mlp {
    bvp_ctx *ctx = nullptr;
    mlp::NetBufs nb{};
    float in_shift[mlp::NIN] = {0, 0, 0, 0}, in_scale[mlp::NIN] = {1, 1, 1, 1};
    float out_shift[mlp::NOUT] = {0, 0}, out_scale[mlp::NOUT] = {1, 1};
    int gen = 0;  // bumped when the normalisation changes (plans re-capture)
    DevArena arena;
    mlp::Net net() const {
        mlp::Net n = nb.view();
        for (int i = 0; i < mlp::NIN; ++i) { n.in_shift[i] = in_shift[i]; n.in_scale[i] = in_scale[i]; }
        for (int o = 0; o < mlp::NOUT; ++o) { n.out_shift[o] = out_shift[o]; n.out_scale[o] = out_scale[o]; }
        return n;
    }
};

struct bvp_batch {
    int n = 0, M = 0, method = 4, cap = 0, blocks = 0;
    int variant = 0;     // solve-kernel variant (wave_variants)
    int family = -1;     // the batch's single family when all its BCs are two-point, else -1
    int max_slots = 0;   // persistent waves the per-wave scratch was sized for
    bool block = false;  // served by the block kernels (M = 8, 16, 32)
    bool firk = false;   // a FIRK method (firk_wave_impl.h), firk_s stages
    int firk_s = 0;
    bool has_u0 = false; // every instance carries u0 (needed by a warm start)
    bvp_opts opts{};
    std::vector<int> N0;
    std::vector<long long> init_off;
    long long total_nodes = 0;
    float last_ms = 0.f;
    float last_guess_ms = 0.f, last_solve_ms = 0.f;
    bool timed_split = false;     // the last run recorded ctx->evm
    BatchDev dev{};
    DevArena arena;
    // learned warm start attached to the plan
    const bvp_mlp *mlp = nullptr;
    int feature_kind = 0;
    double *warm_y = nullptr;  // [total_nodes][M]
    int anchors = 0;           // 0: network at every node; else anchors per instance + Hermite
    double *anc = nullptr;     // [n][anchors][4]: (y1, y2, f1, f2) at the anchors
    int anc_cap = 0;
    int mlp_gen = 0;
    // graph replay of run(): the solve (queue reset + persistent kernel) and the
    // learned guess are captured as two graphs so that an event between them
    // splits the device time
    bool use_graph = true;
    hipGraph_t graph = nullptr, guess_graph = nullptr;
    hipGraphExec_t graph_exec = nullptr, guess_exec = nullptr;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (guess_exec) (void)hipGraphExecDestroy(guess_exec);
        if (guess_graph) (void)hipGraphDestroy(guess_graph);
        graph_exec = guess_exec = nullptr;
        graph = guess_graph = nullptr;
    }
    ~bvp_batch() { drop_graph(); }
};

int wave_blocks_per_cu(int M, int method, int variant, int fam, int f32);
bool blockm_supports(int M, int family);
