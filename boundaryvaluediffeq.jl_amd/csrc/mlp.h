// mlp.h -- the warm-start network: a 4 -> 256 -> 256 -> 256 -> 2 tanh MLP with
// BF16 weights and FP32 accumulation, trained and evaluated by the kernels of
// mlp.hip.  It maps (problem parameter, left boundary value, right boundary
// value, t) to an approximation of the solution (y1, y2)(t), which becomes the
// initial guess of the collocation solve (the guess that the reference takes
// from the user as a vector, function or previous solution, Core/utils.jl:750-777).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mirk_wave.h"

namespace mlp {

constexpr int NIN = 4, H = 256, NOUT = 2;
constexpr int T = 64;        // points per tile (one workgroup of 4 waves)
constexpr int HP = H + 8;    // LDS row stride of the activation tiles (bf16 elements)

// offsets in the FP32 parameter vector; matrices are input-major: W[i][o]
constexpr int OW1 = 0;
constexpr int OB1 = OW1 + NIN * H;
constexpr int OW2 = OB1 + H;
constexpr int OB2 = OW2 + H * H;
constexpr int OW3 = OB2 + H;
constexpr int OB3 = OW3 + H * H;
constexpr int OW4 = OB3 + H;
constexpr int OB4 = OW4 + H * NOUT;
constexpr int NPARAM = OB4 + NOUT;
// per-tile partial: gradient, loss at [NPARAM], padded to a multiple of 4 floats
// so that every slot (and its 16-byte gradient stores) stays 16-byte aligned
constexpr int NSLOT = (NPARAM + 1 + 3) & ~3;

// Device view of a network.
struct Net {
    const float *q;                 // every parameter rounded to BF16, held as FP32 (scalar paths)
    const uint16_t *w2f, *w3f;      // BF16 hidden weights, output-major  [o][i]  (forward)
    const uint16_t *w2b, *w3b;      // BF16 hidden weights, input-major   [i][o]  (backward)
    float in_shift[NIN], in_scale[NIN];
    float out_shift[NOUT], out_scale[NOUT];
};

// Device buffers of a network with its optimiser state.
struct NetBufs {
    float *w;                       // FP32 master parameters [NPARAM]
    float *m, *v;                   // Adam moments
    float *q;
    uint16_t *w2f, *w3f, *w2b, *w3b;
    Net view() const;
};

// Training schedule read by the kernels from device memory, so that one captured
// step replays unchanged.
struct TrainDev {
    const float *x;                 // samples [n][NIN] (raw features)
    const float *y;                 // targets [n][NOUT] (raw)
    long long n;
    int batch_tiles;                // tiles per step
    int steps_per_epoch;
    int total_steps;
    const unsigned long long *perm; // per epoch: (a, c) of the index permutation j -> (a j + c) mod n
    int *step;                      // device step counter
    float *partial;                 // [batch_tiles][NSLOT]
    float *loss;                    // [steps of this call], indexed by step - step0
    int step0;                      // first step of this call
    float lr, lr_min, beta1, beta2, eps, weight_decay;
    int warmup;
};

// feature kinds of the batch warm start
enum { FEAT_PARAM = 0, FEAT_LOG10_PARAM = 1 };

hipError_t launch_quantize(const NetBufs &nb, hipStream_t s);
hipError_t launch_train_step(const NetBufs &nb, const Net &net, const TrainDev &td, hipStream_t s);
// one forward/backward over n points in order, FP32 gradient (sum over tiles) -> grad[NPARAM], loss -> grad[NPARAM]
hipError_t launch_grad(const Net &net, const float *x, const float *y, int n, float *partial, float *grad,
                       hipStream_t s);
hipError_t launch_predict(const Net &net, const float *x, int n, float *out, int num_cus, hipStream_t s);
// The learned guess of a batch (M = 2): the network at every node of the initial
// meshes into init_y[node][2] (anchors == 0), or at `anchors` equally spaced
// points per instance into anc[n][anchors][4] = (y1, y2, f1, f2), f = f(t, y)
// being the problem's slope there; a solve then interpolates the nodes from the
// anchors (anchors.h).  Persistent launch over num_cus workgroups.
hipError_t launch_batch_guess(const Net &net, const BatchDev &B, double *init_y, long long total_nodes,
                              int feature_kind, int anchors, double *anc, int num_cus, hipStream_t s);
// materialises the anchor interpolation at every node into init_y[node][2]
hipError_t launch_anchor_fill(const BatchDev &B, const double *anc, int anchors, double *init_y, hipStream_t s);

}  // namespace mlp
