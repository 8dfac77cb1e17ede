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

namespace mlp {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

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
        for (int p = 0; p < T; ++p) {
            float z = b;
#pragma unroll
            for (int i = 0; i < NIN; ++i) z += S.x[p][i] * w[i];
            S.h[0][p][n] = f2bf(tanhf(z));
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
                for (int j = 0; j < 4; ++j) S.h[l + 1][16 * rt + 4 * g + j][n] = f2bf(tanhf(acc[rt][ct][j] + bn));
            }
        __syncthreads();
    }
    {  // output layer (256 -> 2): four threads per point, 64 inputs each
        const int p = tid >> 2, part = tid & 3;
        float o0 = 0.f, o1 = 0.f;
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
__global__ void k_sum_slots(const float *partial, int tiles, float *grad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > NPARAM) return;
    float s = 0.f;
    for (int t = 0; t < tiles; ++t) s += partial[(size_t)t * NSLOT + j];
    grad[j == NPARAM ? j : slot_param(j)] = s;
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
    float g = 0.f;
    for (int t = 0; t < td.batch_tiles; ++t) g += td.partial[(size_t)t * NSLOT + j];
    if (j == NPARAM) {
        td.loss[step] = g;
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

__global__ __launch_bounds__(256) void k_predict(Net net, const float *x, int n, float *out) {
    __shared__ Tile S;
    const int tid = threadIdx.x;
    if (tid < T) {
        const long long j = (long long)blockIdx.x * T + tid;
        const bool v = j < n;
        load_point(S, net, tid, v, x + (v ? j : 0) * NIN, nullptr);
    }
    __syncthreads();
    forward_tile(S, net, tid);
    if (tid < T) {
        const long long j = (long long)blockIdx.x * T + tid;
        if (j < n)
#pragma unroll
            for (int o = 0; o < NOUT; ++o) out[j * NOUT + o] = S.out[tid][o] * net.out_scale[o] + net.out_shift[o];
    }
}

// Warm-start guess for every node of a batch (M = 2): features (p0 or log10 p0,
// boundary value of row 0, boundary value of row 1, t).
__global__ __launch_bounds__(256) void k_batch_guess(Net net, BatchDev B, double *init_y, long long total, int kind) {
    __shared__ Tile S;
    const int tid = threadIdx.x;
    if (tid < T) {
        const long long gidx = (long long)blockIdx.x * T + tid;
        const bool v = gidx < total;
        float xr[NIN] = {0.f, 0.f, 0.f, 0.f};
        if (v) {
            int lo = 0, hi = B.n - 1;  // last instance whose first node is <= gidx
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (B.init_off[mid] <= gidx) lo = mid;
                else hi = mid - 1;
            }
            const double p0 = B.params[B.param_off[lo]];
            xr[0] = (float)(kind == FEAT_LOG10_PARAM ? log10(p0) : p0);
            xr[1] = (float)B.bc_val[(size_t)lo * B.M];
            xr[2] = (float)B.bc_val[(size_t)lo * B.M + 1];
            xr[3] = (float)B.init_mesh[gidx];
        }
        load_point(S, net, tid, v, xr, nullptr);
    }
    __syncthreads();
    forward_tile(S, net, tid);
    if (tid < T) {
        const long long gidx = (long long)blockIdx.x * T + tid;
        if (gidx < total)
#pragma unroll
            for (int o = 0; o < NOUT; ++o)
                init_y[gidx * NOUT + o] = (double)(S.out[tid][o] * net.out_scale[o] + net.out_shift[o]);
    }
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

hipError_t launch_predict(const Net &net, const float *x, int n, float *out, hipStream_t s) {
    k_predict<<<(n + T - 1) / T, 256, 0, s>>>(net, x, n, out);
    return hipGetLastError();
}

hipError_t launch_batch_guess(const Net &net, const BatchDev &B, double *init_y, long long total_nodes,
                              int feature_kind, hipStream_t s) {
    const long long tiles = (total_nodes + T - 1) / T;
    k_batch_guess<<<(unsigned)tiles, 256, 0, s>>>(net, B, init_y, total_nodes, feature_kind);
    return hipGetLastError();
}

}  // namespace mlp
