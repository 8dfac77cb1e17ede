// mlp.hip -- training and inference kernels of the warm-start network (mlp.h).
//
// One workgroup of 4 waves handles a tile of 64 points and keeps the whole
// tile on chip: inputs, the three BF16 activation layers and one BF16 gradient
// layer live in LDS (about 137 KB), the hidden 256 x 256 products run on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation; each wave
// owns 64 output columns of the 64-row tile), the 4-wide input layer and the
// 2-wide output layer run on the vector ALUs.  Training keeps no activations in
// HBM: the tile goes forward and backward in the same workgroup and writes its
// FP32 parameter gradient to a per-tile slot; a second kernel sums the slots,
// applies Adam to the FP32 master parameters and re-rounds them to BF16 in the
// two layouts the products read (output-major for the forward pass, input-major
// for the backward pass).  The step counter lives in device memory so that a
// captured step replays unchanged.
//
// Rounding points (the CPU oracle of the gradient check rounds at the same
// places): parameters to BF16 (round to nearest even), every activation to
// BF16 after tanh, every hidden-layer gradient dL/dz to BF16; all sums in FP32.
#include "mlp.h"

#include <math.h>

#include "anchors.h"
#include "mirk_device.h"

namespace mlp {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (gfx950: one v_cvt_pk_bf16_f32)
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// tanh z = 1 - 2 / (1 + e^{2z}) with the hardware exponential (absolute error
// ~1e-7, far below the BF16 rounding the activations then get)
__device__ inline float tanh_f(float z) { return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + __expf(2.f * z)); }

struct __align__(16) Tile {
    uint16_t h[3][T][HP];   // activations of the hidden layers (BF16)
    uint16_t dz[T][HP];     // dL/dz of the layer being back-propagated (BF16)
    float x[T][NIN];        // normalised inputs
    float out[T][NOUT];     // network output (normalised units), then dL/dout
    float tgt[T][NOUT];
    float lossv[T];
    int valid[T];
};

// acc[rt][ct] (rows 16rt.., columns 64*wave + 16ct..) = A (LDS, T x H) times the
// matrix whose column c is row c of G (global BF16, H x H, row-major).
__device__ inline void gemm_tile(const uint16_t *A, const uint16_t *G, int wave, int lane, f4 acc[4][4]) {
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < H; k0 += 32) {
        bf8 a[4], b[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) a[rt] = *(const bf8 *)(A + (size_t)(16 * rt + i) * HP + k0 + 8 * g);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) b[ct] = *(const bf8 *)(G + (size_t)(64 * wave + 16 * ct + i) * H + k0 + 8 * g);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
    }
}

typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// eight BF16 values at p[0], p[stride], ..., p[7 stride], packed as one MFMA operand
__device__ inline bf8 gather8(const uint16_t *p, int stride) {
    u4 r;
    r.x = (unsigned)p[0] | ((unsigned)p[stride] << 16);
    r.y = (unsigned)p[2 * stride] | ((unsigned)p[3 * stride] << 16);
    r.z = (unsigned)p[4 * stride] | ((unsigned)p[5 * stride] << 16);
    r.w = (unsigned)p[6 * stride] | ((unsigned)p[7 * stride] << 16);
    return __builtin_bit_cast(bf8, r);
}

// dW[k][n] = sum_p Hin[p][k] dz[p][n] for this tile, written output-major to
// part[n*H + k] (four consecutive k per lane: one 16-byte store); wave w owns
// rows k in [64w, 64w+64).
__device__ inline void wgrad_tile(const uint16_t *Hin, const uint16_t *dz, float *part, int wave, int lane) {
    const int i = lane & 15, g = lane >> 4;
    bf8 a[4][2];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) a[rt][ks] = gather8(Hin + (size_t)(32 * ks + 8 * g) * HP + 64 * wave + 16 * rt + i, HP);
    for (int np = 0; np < 4; ++np) {
        f4 acc[4][4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bf8 b[4];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) b[ct] = gather8(dz + (size_t)(32 * ks + 8 * g) * HP + 64 * np + 16 * ct + i, HP);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rt][ks], b[ct], acc[rt][ct], 0, 0, 0);
        }
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            f4 *dst = (f4 *)(part + (size_t)(64 * np + 16 * ct + i) * H + 64 * wave + 4 * g);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) dst[4 * rt] = acc[rt][ct];
        }
    }
}

// Index in the parameter vector of slot entry j (the hidden matrices are
// stored output-major in the per-tile slots, input-major in the parameters).
__device__ inline int slot_param(int j) {
    if (j >= OW2 && j < OW2 + H * H) return OW2 + ((j - OW2) % H) * H + (j - OW2) / H;
    if (j >= OW3 && j < OW3 + H * H) return OW3 + ((j - OW3) % H) * H + (j - OW3) / H;
    return j;
}

// Forward pass of the tile (S.x filled): hidden activations into S.h, output into S.out.
__device__ void forward_tile(Tile &S, const Net &net, int tid) {
    const int wave = tid >> 6, lane = tid & 63;
    {  // layer 1 (4 -> 256): thread tid owns column n = tid
        const int n = tid;
        float w[NIN];
#pragma unroll
        for (int i = 0; i < NIN; ++i) w[i] = net.q[OW1 + i * H + n];
        const float b = net.q[OB1 + n];
#pragma unroll 8
        for (int p = 0; p < T; ++p) {
            float z = b;
#pragma unroll
            for (int i = 0; i < NIN; ++i) z += S.x[p][i] * w[i];
            S.h[0][p][n] = f2bf(tanh_f(z));
        }
    }
    __syncthreads();
    const int i = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (int l = 0; l < 2; ++l) {  // layers 2 and 3 (256 -> 256) on the matrix cores
        f4 acc[4][4];
        gemm_tile(&S.h[l][0][0], l == 0 ? net.w2f : net.w3f, wave, lane, acc);
        const float *bias = net.q + (l == 0 ? OB2 : OB3);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const int n = 64 * wave + 16 * ct + i;
                const float bn = bias[n];
#pragma unroll
                for (int j = 0; j < 4; ++j) S.h[l + 1][16 * rt + 4 * g + j][n] = f2bf(tanh_f(acc[rt][ct][j] + bn));
            }
        __syncthreads();
    }
    {  // output layer (256 -> 2): four threads per point, 64 inputs each
        const int p = tid >> 2, part = tid & 3;
        float o0 = 0.f, o1 = 0.f;
#pragma unroll 16
        for (int k = 64 * part; k < 64 * part + 64; ++k) {
            const float hk = bf2f(S.h[2][p][k]);
            o0 += hk * net.q[OW4 + 2 * k];
            o1 += hk * net.q[OW4 + 2 * k + 1];
        }
        o0 += __shfl_xor(o0, 1);
        o1 += __shfl_xor(o1, 1);
        o0 += __shfl_xor(o0, 2);
        o1 += __shfl_xor(o1, 2);
        if (part == 0) {
            S.out[p][0] = o0 + net.q[OB4];
            S.out[p][1] = o1 + net.q[OB4 + 1];
        }
    }
    __syncthreads();
}

// Backward pass of the tile after forward_tile; inv_b = 1 / (points in the step).
// Writes the tile's gradient and loss into part[NSLOT].
__device__ void backward_tile(Tile &S, const Net &net, int tid, float inv_b, float *part) {
    const int wave = tid >> 6, lane = tid & 63;
    const int i = lane & 15, g = lane >> 4;
    // dL/dout and the loss (L = sum 0.5 (out - y)^2 / B)
    if (tid < T) {
        const int p = tid;
        float l = 0.f;
#pragma unroll
        for (int o = 0; o < NOUT; ++o) {
            const float d = S.valid[p] ? (S.out[p][o] - S.tgt[p][o]) : 0.f;
            l += 0.5f * d * d;
            S.out[p][o] = d * inv_b;
        }
        S.lossv[p] = l * inv_b;
    }
    __syncthreads();
    {  // output layer: dW4, db4, and dz3 = (dout W4^T) * (1 - h3^2); thread tid owns hidden unit k
        const int k = tid;
        const float w0 = net.q[OW4 + 2 * k], w1 = net.q[OW4 + 2 * k + 1];
        float g0 = 0.f, g1 = 0.f;
        for (int p = 0; p < T; ++p) {
            const float hk = bf2f(S.h[2][p][k]);
            const float d0 = S.out[p][0], d1 = S.out[p][1];
            g0 += hk * d0;
            g1 += hk * d1;
            S.dz[p][k] = f2bf((d0 * w0 + d1 * w1) * (1.f - hk * hk));
        }
        part[OW4 + 2 * k] = g0;
        part[OW4 + 2 * k + 1] = g1;
        if (tid < NOUT) {
            float s = 0.f;
            for (int p = 0; p < T; ++p) s += S.out[p][tid];
            part[OB4 + tid] = s;
        }
        if (tid == 0) {
            float s = 0.f;
            for (int p = 0; p < T; ++p) s += S.lossv[p];
            part[NPARAM] = s;
        }
    }
    __syncthreads();
#pragma unroll 1
    for (int l = 1; l >= 0; --l) {  // hidden layer l + 2 (weights W3 for l = 1, W2 for l = 0)
        const uint16_t *hin = &S.h[l][0][0];
        wgrad_tile(hin, &S.dz[0][0], part + (l == 1 ? OW3 : OW2), wave, lane);
        {
            const int n = tid;
            float s = 0.f;
            for (int p = 0; p < T; ++p) s += bf2f(S.dz[p][n]);
            part[(l == 1 ? OB3 : OB2) + n] = s;
        }
        f4 acc[4][4];
        gemm_tile(&S.dz[0][0], l == 1 ? net.w3b : net.w2b, wave, lane, acc);
        __syncthreads();
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const int k = 64 * wave + 16 * ct + i;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int p = 16 * rt + 4 * g + j;
                    const float hk = bf2f(S.h[l][p][k]);
                    S.dz[p][k] = f2bf(acc[rt][ct][j] * (1.f - hk * hk));
                }
            }
        __syncthreads();
    }
    {  // input layer: dW1, db1
        const int n = tid;
        float gw[NIN] = {0.f, 0.f, 0.f, 0.f}, gb = 0.f;
        for (int p = 0; p < T; ++p) {
            const float d = bf2f(S.dz[p][n]);
            gb += d;
#pragma unroll
            for (int q = 0; q < NIN; ++q) gw[q] += S.x[p][q] * d;
        }
#pragma unroll
        for (int q = 0; q < NIN; ++q) part[OW1 + q * H + n] = gw[q];
        part[OB1 + n] = gb;
    }
}

__device__ inline void load_point(Tile &S, const Net &net, int p, bool valid, const float *xr, const float *yr) {
#pragma unroll
    for (int q = 0; q < NIN; ++q) S.x[p][q] = valid ? (xr[q] - net.in_shift[q]) * net.in_scale[q] : 0.f;
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
        S.tgt[p][o] = (valid && yr) ? (yr[o] - net.out_shift[o]) / net.out_scale[o] : 0.f;
    S.valid[p] = valid ? 1 : 0;
}

// ------------------------------------------------------------------ kernels
__global__ __launch_bounds__(256) void k_train_tile(Net net, TrainDev td) {
    __shared__ Tile S;
    const int tid = threadIdx.x;
    const int step = td.step[0];
    const int epoch = step / td.steps_per_epoch, sin = step % td.steps_per_epoch;
    const unsigned long long a = td.perm[2 * epoch], c = td.perm[2 * epoch + 1];
    const long long bsz = (long long)td.batch_tiles * T;
    if (tid < T) {
        const unsigned long long gidx = (unsigned long long)sin * bsz + (unsigned long long)blockIdx.x * T + tid;
        const long long j = (long long)((a * gidx + c) % (unsigned long long)td.n);
        load_point(S, net, tid, true, td.x + j * NIN, td.y + j * NOUT);
    }
    __syncthreads();
    forward_tile(S, net, tid);
    backward_tile(S, net, tid, 1.0f / (float)bsz, td.partial + (size_t)blockIdx.x * NSLOT);
}

__global__ __launch_bounds__(256) void k_grad_tile(Net net, const float *x, const float *y, int n, float *partial) {
    __shared__ Tile S;
    const int tid = threadIdx.x;
    if (tid < T) {
        const long long j = (long long)blockIdx.x * T + tid;
        const bool v = j < n;
        load_point(S, net, tid, v, x + (v ? j : 0) * NIN, y + (v ? j : 0) * NOUT);
    }
    __syncthreads();
    forward_tile(S, net, tid);
    backward_tile(S, net, tid, 1.0f / (float)n, partial + (size_t)blockIdx.x * NSLOT);
}

// sum of the per-tile slots, in parameter order: grad[0..NPARAM) and the loss in grad[NPARAM]
// Sum of value j over the per-tile gradient slots, in a fixed order (the same bits
// on every run): eight independent partial sums, so that eight loads of a thread
// are in flight at a time, combined pairwise.
__device__ inline float sum_slots(const float *partial, int tiles, int j) {
    float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int t = 0;
    for (; t + 8 <= tiles; t += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partial[(size_t)(t + u) * NSLOT + j];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += v[u];
    }
    for (; t < tiles; ++t) a[0] += partial[(size_t)t * NSLOT + j];
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

__global__ void k_sum_slots(const float *partial, int tiles, float *grad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > NPARAM) return;
    grad[j == NPARAM ? j : slot_param(j)] = sum_slots(partial, tiles, j);
}

__device__ inline void requant(const NetBufs &nb, int j, float w) {
    const uint16_t b = f2bf(w);
    nb.q[j] = bf2f(b);
    if (j >= OW2 && j < OW2 + H * H) {
        const int k = (j - OW2) / H, o = (j - OW2) % H;
        nb.w2b[k * H + o] = b;
        nb.w2f[o * H + k] = b;
    } else if (j >= OW3 && j < OW3 + H * H) {
        const int k = (j - OW3) / H, o = (j - OW3) % H;
        nb.w3b[k * H + o] = b;
        nb.w3f[o * H + k] = b;
    }
}

__global__ void k_quantize(NetBufs nb) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < NPARAM) requant(nb, j, nb.w[j]);
}

// Adam on the summed tile gradients (decoupled weight decay on matrices only),
// learning rate: linear warm-up then cosine decay to lr_min.
__global__ void k_adam(NetBufs nb, TrainDev td) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > NPARAM) return;
    const int step = td.step[0];
    const float g = sum_slots(td.partial, td.batch_tiles, j);
    if (j == NPARAM) {
        td.loss[step - td.step0] = g;
        return;
    }
    const int pj = slot_param(j);
    float lr;
    if (step < td.warmup) lr = td.lr * (float)(step + 1) / (float)td.warmup;
    else {
        const float f = (float)(step - td.warmup) / (float)max(1, td.total_steps - td.warmup);
        lr = td.lr_min + 0.5f * (td.lr - td.lr_min) * (1.f + cosf(3.14159265358979f * fminf(f, 1.f)));
    }
    const float t = (float)(step + 1);
    const float m = nb.m[pj] = td.beta1 * nb.m[pj] + (1.f - td.beta1) * g;
    const float v = nb.v[pj] = td.beta2 * nb.v[pj] + (1.f - td.beta2) * g * g;
    const float mh = m / (1.f - powf(td.beta1, t)), vh = v / (1.f - powf(td.beta2, t));
    const bool matrix = (pj < OB1) || (pj >= OW2 && pj < OB2) || (pj >= OW3 && pj < OB3) || (pj >= OW4 && pj < OB4);
    float w = nb.w[pj];
    if (matrix) w -= lr * td.weight_decay * w;
    w -= lr * mh / (sqrtf(vh) + td.eps);
    nb.w[pj] = w;
    requant(nb, pj, w);
}

__global__ void k_advance(int *step) {
    if (threadIdx.x == 0) step[0] += 1;
}

// last instance whose first node is <= gidx
__device__ inline int instance_of_node(const BatchDev &B, long long gidx) {
    int lo = 0, hi = B.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.init_off[mid] <= gidx) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ inference
// Persistent inference: each workgroup keeps its waves' column slices of both
// hidden weight matrices in registers (2 x 4 x 8 operands of 8 BF16 per lane)
// and streams 64-point tiles through two LDS activation buffers, so that the
// weights are read from memory once per workgroup instead of once per tile.
// what a source keeps of a point between reading its features and storing its
// output (the anchor source: the instance's parameters, time and family)
constexpr int kStash = 8;
// the features of kAhead tiles are read together, by all 256 threads
constexpr int kAhead = 4;
struct InferTile {
    uint16_t h[2][T][HP];
    float w4[H][NOUT];
    float x[kAhead][T][NIN];
    float out[T][NOUT];
    double stash[kAhead][T][kStash];
    int valid[kAhead][T];
};

__device__ inline void hidden_reg(const uint16_t *A, uint16_t *Hout, const bf8 (&w)[4][8], const float (&bias)[4],
                                  int wave, int lane) {
    const int i = lane & 15, g = lane >> 4;
    f4 acc[4][4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        bf8 a[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) a[rt] = *(const bf8 *)(A + (size_t)(16 * rt + i) * HP + 32 * ks + 8 * g);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rt], w[ct][ks], acc[rt][ct], 0, 0, 0);
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                Hout[(size_t)(16 * rt + 4 * g + j) * HP + 64 * wave + 16 * ct + i] =
                    f2bf(tanh_f(acc[rt][ct][j] + bias[ct]));
}

template <class Src>
__global__ __launch_bounds__(256) void k_infer(Net net, Src src, long long npts) {
    __shared__ InferTile S;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, i = lane & 15, g = lane >> 4;
    bf8 w2[4][8], w3[4][8];
    float b2[4], b3[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int n = 64 * wave + 16 * ct + i;
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            w2[ct][ks] = *(const bf8 *)(net.w2f + (size_t)n * H + 32 * ks + 8 * g);
            w3[ct][ks] = *(const bf8 *)(net.w3f + (size_t)n * H + 32 * ks + 8 * g);
        }
        b2[ct] = net.q[OB2 + n];
        b3[ct] = net.q[OB3 + n];
    }
    float w1[NIN];
#pragma unroll
    for (int q = 0; q < NIN; ++q) w1[q] = net.q[OW1 + q * H + tid];
    const float b1 = net.q[OB1 + tid];
    S.w4[tid][0] = net.q[OW4 + 2 * tid];
    S.w4[tid][1] = net.q[OW4 + 2 * tid + 1];
    const float b40 = net.q[OB4], b41 = net.q[OB4 + 1];
    const long long ntiles = (npts + T - 1) / T;
    static_assert(kAhead * T == 256, "one feature point per thread");
    for (long long tile0 = blockIdx.x; tile0 < ntiles; tile0 += (long long)kAhead * gridDim.x) {
        {  // features of this workgroup's next kAhead tiles (tile0 + qa * gridDim.x)
            const int qa = tid / T, pa = tid % T;
            const long long tile = tile0 + (long long)qa * gridDim.x;
            const long long gi = tile * T + pa;
            const bool v = tile < ntiles && gi < npts;
            float xr[NIN] = {0.f, 0.f, 0.f, 0.f};
            if (v) src.features(gi, xr, &S.stash[qa][pa][0]);
#pragma unroll
            for (int q = 0; q < NIN; ++q) S.x[qa][pa][q] = v ? (xr[q] - net.in_shift[q]) * net.in_scale[q] : 0.f;
            S.valid[qa][pa] = v;
        }
        __syncthreads();
        for (int qa = 0; qa < kAhead; ++qa) {
        const long long tile = tile0 + (long long)qa * gridDim.x;
        if (tile >= ntiles) break;
#pragma unroll 8
        for (int p = 0; p < T; ++p) {
            float z = b1;
#pragma unroll
            for (int q = 0; q < NIN; ++q) z += S.x[qa][p][q] * w1[q];
            S.h[0][p][tid] = f2bf(tanh_f(z));
        }
        __syncthreads();
        hidden_reg(&S.h[0][0][0], &S.h[1][0][0], w2, b2, wave, lane);
        __syncthreads();
        hidden_reg(&S.h[1][0][0], &S.h[0][0][0], w3, b3, wave, lane);
        __syncthreads();
        {
            const int p = tid >> 2, part = tid & 3;
            float o0 = 0.f, o1 = 0.f;
#pragma unroll 16
            for (int k = 64 * part; k < 64 * part + 64; ++k) {
                const float hk = bf2f(S.h[0][p][k]);
                o0 += hk * S.w4[k][0];
                o1 += hk * S.w4[k][1];
            }
            o0 += __shfl_xor(o0, 1);
            o1 += __shfl_xor(o1, 1);
            o0 += __shfl_xor(o0, 2);
            o1 += __shfl_xor(o1, 2);
            if (part == 0) {
                S.out[p][0] = (o0 + b40) * net.out_scale[0] + net.out_shift[0];
                S.out[p][1] = (o1 + b41) * net.out_scale[1] + net.out_shift[1];
            }
        }
        __syncthreads();
        if (tid < T && S.valid[qa][tid]) src.store(tile * T + tid, S.out[tid], &S.stash[qa][tid][0]);
        }
        __syncthreads();  // the next feature round overwrites x, stash and valid
    }
}

struct PredictSrc {
    const float *x;
    float *y;
    __device__ void features(long long gi, float *xr, double *) const {
#pragma unroll
        for (int q = 0; q < NIN; ++q) xr[q] = x[gi * NIN + q];
    }
    __device__ void store(long long gi, const float *o, const double *) const {
        y[gi * NOUT] = o[0];
        y[gi * NOUT + 1] = o[1];
    }
};

__device__ inline void batch_features(const BatchDev &B, int b, double t, int kind, float *xr) {
    const double p0 = B.params[B.param_off[b]];
    xr[0] = (float)(kind == FEAT_LOG10_PARAM ? log10(p0) : p0);
    xr[1] = (float)B.bc_val[(size_t)b * B.M];
    xr[2] = (float)B.bc_val[(size_t)b * B.M + 1];
    xr[3] = (float)t;
}

// every node of the initial meshes of a batch -> out[node][2]
struct NodeSrc {
    BatchDev B;
    double *out;
    int kind;
    __device__ void features(long long gi, float *xr, double *) const {
        batch_features(B, instance_of_node(B, gi), B.init_mesh[gi], kind, xr);
    }
    __device__ void store(long long gi, const float *o, const double *) const {
        out[gi * 2] = (double)o[0];
        out[gi * 2 + 1] = (double)o[1];
    }
};

// K anchors per instance -> anc[b][j] = (y1, y2, f1, f2) with f = f(t_j, y)
struct AnchorSrc {
    BatchDev B;
    double *anc;
    int K;
    int kind;
    // the stash keeps the instance's parameters (5), the anchor time and the family
    __device__ void features(long long gi, float *xr, double *st) const {
        const int b = (int)(gi / K);
        const double t = anchor_t(B, b, (int)(gi % K), K);
        batch_features(B, b, t, kind, xr);
        const int po = B.param_off[b], pn = B.param_cnt[b];
#pragma unroll
        for (int k = 0; k < bvpdev::Params<2>::MAXP; ++k) st[k] = k < pn ? B.params[po + k] : 0.0;
        st[5] = t;
        st[6] = (double)B.family[b];
    }
    __device__ void store(long long gi, const float *o, const double *st) const {
        bvpdev::Params<2> P;
#pragma unroll
        for (int k = 0; k < bvpdev::Params<2>::MAXP; ++k) P.p[k] = st[k];
        const double y[2] = {(double)o[0], (double)o[1]};
        double f[2];
        bvpdev::f_eval<2, double>((int)st[6], P, st[5], y, f);
        anc[gi * 4] = y[0];
        anc[gi * 4 + 1] = y[1];
        anc[gi * 4 + 2] = f[0];
        anc[gi * 4 + 3] = f[1];
    }
};

// the guess at every node from the anchors (materialised for inspection; a solve
// interpolates on the fly)
__global__ void k_anchor_fill(BatchDev B, const double *anc, int K, double *init_y) {
    const int b = blockIdx.x;
    const long long first = B.init_off[b];
    const int nodes = B.N0[b] + 1;
    for (int k = threadIdx.x; k < nodes; k += blockDim.x)
        anchor_hermite(anc, K, b, B.t0[b], B.t1[b], B.init_mesh[first + k], init_y + (first + k) * 2);
}

// ------------------------------------------------------------------ host side
Net NetBufs::view() const {
    Net n{};
    n.q = q;
    n.w2f = w2f;
    n.w3f = w3f;
    n.w2b = w2b;
    n.w3b = w3b;
    return n;
}

hipError_t launch_quantize(const NetBufs &nb, hipStream_t s) {
    k_quantize<<<(NPARAM + 255) / 256, 256, 0, s>>>(nb);
    return hipGetLastError();
}

hipError_t launch_train_step(const NetBufs &nb, const Net &net, const TrainDev &td, hipStream_t s) {
    k_train_tile<<<td.batch_tiles, 256, 0, s>>>(net, td);
    k_adam<<<(NSLOT + 255) / 256, 256, 0, s>>>(nb, td);
    k_advance<<<1, 64, 0, s>>>(td.step);
    return hipGetLastError();
}

hipError_t launch_grad(const Net &net, const float *x, const float *y, int n, float *partial, float *grad,
                       hipStream_t s) {
    const int tiles = (n + T - 1) / T;
    k_grad_tile<<<tiles, 256, 0, s>>>(net, x, y, n, partial);
    k_sum_slots<<<(NSLOT + 255) / 256, 256, 0, s>>>(partial, tiles, grad);
    return hipGetLastError();
}

static unsigned infer_grid(long long npts, int num_cus) {
    const long long tiles = (npts + T - 1) / T;
    return (unsigned)(tiles < num_cus ? (tiles > 0 ? tiles : 1) : num_cus);
}

hipError_t launch_predict(const Net &net, const float *x, int n, float *out, int num_cus, hipStream_t s) {
    k_infer<PredictSrc><<<infer_grid(n, num_cus), 256, 0, s>>>(net, PredictSrc{x, out}, n);
    return hipGetLastError();
}

hipError_t launch_batch_guess(const Net &net, const BatchDev &B, double *init_y, long long total_nodes,
                              int feature_kind, int anchors, double *anc, int num_cus, hipStream_t s) {
    if (anchors == 0) {
        k_infer<NodeSrc><<<infer_grid(total_nodes, num_cus), 256, 0, s>>>(net, NodeSrc{B, init_y, feature_kind},
                                                                          total_nodes);
        return hipGetLastError();
    }
    const long long pts = (long long)B.n * anchors;
    k_infer<AnchorSrc><<<infer_grid(pts, num_cus), 256, 0, s>>>(net, AnchorSrc{B, anc, anchors, feature_kind}, pts);
    return hipGetLastError();
}

hipError_t launch_anchor_fill(const BatchDev &B, const double *anc, int anchors, double *init_y, hipStream_t s) {
    k_anchor_fill<<<B.n, 256, 0, s>>>(B, anc, anchors, init_y);
    return hipGetLastError();
}

}  // namespace mlp
