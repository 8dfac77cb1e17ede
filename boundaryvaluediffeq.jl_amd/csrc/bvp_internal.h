// bvp_internal.h -- host-side state behind the opaque handles of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/bvp_mirk.h"
#include "mirk_wave.h"

struct bvp_ctx {
    int device = 0;
    int num_cus = 0;
    std::string arch;
    hipStream_t stream = nullptr;  // default priority
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int blocks_per_cu = 0;         // 0: from the occupancy calculator
    std::string err;
};

// device allocations owned by one batch; freed together
struct DevArena {
    std::vector<void *> ptrs;
    DevArena() = default;
    DevArena(const DevArena &) = delete;
    DevArena &operator=(const DevArena &) = delete;
    ~DevArena();
};

struct bvp_batch {
    int n = 0, M = 0, method = 4, cap = 0, blocks = 0;
    bool block = false;  // served by the block kernels (M = 8, 16, 32)
    bvp_opts opts{};
    std::vector<int> N0;
    std::vector<long long> init_off;
    long long total_nodes = 0;
    float last_ms = 0.f;
    BatchDev dev{};
    DevArena arena;
};

int wave_blocks_per_cu(int M, int method);
bool blockm_supports(int M, int family);
