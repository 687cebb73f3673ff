// Node-scoring network for learned pivot selection: 8 -> 128 -> 128 -> 1,
// ReLU hidden layers, BF16 operands on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16), FP32 accumulation, FP32 master weights and
// Adam moments.
//
// Layout.  Parameters are one FP32 vector (offsets ZG_MLP_*):
//   W1[f][o] (8 x 128), b1[128], W2[i][o] (128 x 128), b2[128], w3[128], b3.
// Every layer is computed transposed (neurons x nodes): a wavefront owns 32
// nodes, the MFMA's N dimension (lane & 31 is the node), and the neurons are
// the M dimension.  An MFMA result register i of lane l holds neuron
// 32t + 8(i/4) + 4(l/32) + (i%4) of M-tile t, so a layer's output is already
// the next layer's B operand: the K order of each 16-wide K step is permuted
// to the order the registers hold (slot j of lane half h in K step k is
// neuron 16k + 8(j/4) + 4h + (j%4)), and the weight fragments (A operands)
// are packed in that order once per weight update.
//
// Training (zg_mlp_gradient) takes the mean squared error over the batch.
// Per 128-node chunk a block (4 wavefronts) runs forward and backward per
// wavefront, then stages relu(z1), dz2 and dz1 (BF16) in LDS so that the
// weight gradients dW2 = H1 dZ2^T and dW1 = X dZ1^T (contractions over
// nodes) also run on MFMA; b1, b2, w3 and b3 gradients are FP32 sums
// (a halving butterfly over the 32 lanes of a half-wave).  Each block writes
// its partial gradient; a second kernel sums the blocks in a fixed order.
#include "common.h"

#include <cmath>
#include <cstring>

typedef __bf16 zg_bf16x8 __attribute__((ext_vector_type(8)));
typedef float zg_f32x16 __attribute__((ext_vector_type(16)));

#define ZG_MLP_H 128
#define ZG_MLP_W1 0
#define ZG_MLP_B1 (ZG_MLP_W1 + ZG_NFEAT * ZG_MLP_H)
#define ZG_MLP_W2 (ZG_MLP_B1 + ZG_MLP_H)
#define ZG_MLP_B2 (ZG_MLP_W2 + ZG_MLP_H * ZG_MLP_H)
#define ZG_MLP_W3 (ZG_MLP_B2 + ZG_MLP_H)
#define ZG_MLP_B3 (ZG_MLP_W3 + ZG_MLP_H)
#define ZG_MLP_P (ZG_MLP_B3 + 1)
#define ZG_MLP_BLOCKS (ZG_MLP_MAX_BATCH / 128)  // training blocks: one 128-node chunk and partial each
static_assert(ZG_MLP_P == ZG_MLP_PARAMS, "parameter layout");
#define ZG_MLP_LDS_STRIDE 136  // BF16 row stride of the staged [neuron][node] tiles

struct zg_mlp {
    int device = 0;
    uint64_t step = 0;
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    float *part = nullptr;      // [ZG_MLP_BLOCKS][ZG_MLP_P + 1] (last entry: loss)
    double *loss = nullptr;     // [1]
    uint4 *pw1 = nullptr;       // [4][64] layer-1 A fragments
    uint4 *pw2f = nullptr;      // [4][8][64] layer-2 forward A fragments (W2^T)
    uint4 *pw2b = nullptr;      // [4][8][64] layer-2 backward A fragments (W2)
};

// Hides a pointer's value from the optimiser so that loads through it stay
// inside the loop (hoisting every weight fragment out of the node loop would
// need several hundred registers).
template <typename T>
__device__ __forceinline__ const T *zg_opaque(const T *p)
{
    asm volatile("" : "+s"(p));
    return p;
}

union ZgFrag {
    uint4 u;
    uint16_t h[8];
};

// FP32 -> BF16, round to nearest even.
__device__ __forceinline__ uint16_t zg_f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ zg_bf16x8 zg_as_bf16x8(uint4 u) { return __builtin_bit_cast(zg_bf16x8, u); }

__device__ __forceinline__ zg_f32x16 zg_mfma(uint4 a, uint4 b, zg_f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(zg_as_bf16x8(a), zg_as_bf16x8(b), c, 0, 0, 0);
}

// 1.0 if x is non-zero (either sign of zero gives 0.0), without a compare:
// per-element compares would each hold a 64-bit lane mask in scalar
// registers.
__device__ __forceinline__ float zg_nonzero(float x) { return (float)min(__float_as_uint(x) & 0x7FFFFFFFu, 1u); }

// Neuron of result register i (of M-tile t) in lane half h.
__device__ __forceinline__ int zg_neuron(int t, int i, int h) { return 32 * t + 8 * (i >> 2) + 4 * h + (i & 3); }

// Input neuron of slot j of lane half h in K step k.
__device__ __forceinline__ int zg_kslot(int k, int h, int j) { return 16 * k + 8 * (j >> 2) + 4 * h + (j & 3); }

// Packs the BF16 A fragments from the FP32 parameters.
__global__ void mlp_pack_kernel(const float *p, uint4 *pw1, uint4 *pw2f, uint4 *pw2b)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = id & 63, h = lane >> 5, r = lane & 31;
    ZgFrag f;
    if (id < 4 * 64) {
        const int t = id >> 6, o = 32 * t + r;
        for (int j = 0; j < 8; j++) f.h[j] = h == 0 ? zg_f2bf(p[ZG_MLP_W1 + j * ZG_MLP_H + o]) : 0;
        pw1[id] = f.u;
        return;
    }
    const int id2 = id - 4 * 64;
    if (id2 >= 2 * 4 * 8 * 64) return;
    const bool back = id2 >= 4 * 8 * 64;
    const int q = back ? id2 - 4 * 8 * 64 : id2;
    const int t = q >> 9, k = (q >> 6) & 7, m = 32 * t + r;
    for (int j = 0; j < 8; j++) {
        const int kk = zg_kslot(k, h, j);
        // forward A = W2^T [out m][in kk]; backward A = W2 [in m][out kk]
        f.h[j] = zg_f2bf(back ? p[ZG_MLP_W2 + m * ZG_MLP_H + kk] : p[ZG_MLP_W2 + kk * ZG_MLP_H + m]);
    }
    (back ? pw2b : pw2f)[q] = f.u;
}

struct MlpArgs {
    const uint4 *pw1, *pw2f, *pw2b;
    const float *p;
    const float *x, *target;
    uint64_t n;
    float scale;   // d loss / d y = scale * (y - target)
    float *y;
    float *part;
};

// Forward of the 32-node tile starting at n0 (one wavefront): hb are the
// BF16 relu(z1) fragments (K step k of layer 2), z2 the layer-2
// pre-activations (bias included); returns y for node lane & 31 (both
// halves).
__device__ __forceinline__ float mlp_fwd_tile(const MlpArgs &A, const uint4 *pw1, const uint4 *pw2f, const float *b1,
                                              const float *b2, const float *w3, float b3, uint64_t n0, int lane,
                                              uint4 (&hb)[8], zg_f32x16 (&z2)[4])
{
    const int h = lane >> 5;
    const uint64_t node = n0 + (lane & 31);
    ZgFrag bx;
    bx.u = make_uint4(0, 0, 0, 0);
    if (h == 0 && node < A.n) {
        const float4 a = *(const float4 *)(A.x + node * ZG_NFEAT);
        const float4 b = *(const float4 *)(A.x + node * ZG_NFEAT + 4);
        bx.h[0] = zg_f2bf(a.x); bx.h[1] = zg_f2bf(a.y); bx.h[2] = zg_f2bf(a.z); bx.h[3] = zg_f2bf(a.w);
        bx.h[4] = zg_f2bf(b.x); bx.h[5] = zg_f2bf(b.y); bx.h[6] = zg_f2bf(b.z); bx.h[7] = zg_f2bf(b.w);
    }
    const zg_f32x16 zero = {};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const zg_f32x16 z = zg_mfma(pw1[t * 64 + lane], bx.u, zero);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            ZgFrag f;
#pragma unroll
            for (int j = 0; j < 8; j++) f.h[j] = zg_f2bf(fmaxf(z[8 * c + j] + b1[zg_neuron(t, 8 * c + j, h)], 0.f));
            hb[2 * t + c] = f.u;
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) z2[t] = zero;
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
        for (int t = 0; t < 4; t++) z2[t] = zg_mfma(pw2f[(t * 8 + k) * 64 + lane], hb[k], z2[t]);
    float y = 0.f;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            z2[t][i] += b2[zg_neuron(t, i, h)];
            y += w3[zg_neuron(t, i, h)] * fmaxf(z2[t][i], 0.f);
        }
    y += __shfl_xor(y, 32, 64);
    return y + b3;
}

// Inference: y[node] for n nodes; each block stages the weight fragments in
// LDS and its wavefronts stride over 32-node tiles.
__global__ __launch_bounds__(256) void mlp_forward_kernel(MlpArgs A)
{
    __shared__ uint4 s_w1[4 * 64];
    __shared__ uint4 s_w2[4 * 8 * 64];
    __shared__ float s_b1[ZG_MLP_H], s_b2[ZG_MLP_H], s_w3[ZG_MLP_H];
    const int tid = threadIdx.x;
    for (int i = tid; i < 4 * 64; i += 256) s_w1[i] = A.pw1[i];
    for (int i = tid; i < 4 * 8 * 64; i += 256) s_w2[i] = A.pw2f[i];
    if (tid < ZG_MLP_H) {
        s_b1[tid] = A.p[ZG_MLP_B1 + tid];
        s_b2[tid] = A.p[ZG_MLP_B2 + tid];
        s_w3[tid] = A.p[ZG_MLP_W3 + tid];
    }
    __syncthreads();
    const float b3 = A.p[ZG_MLP_B3];
    const int lane = tid & 63;
    const uint64_t tiles = (A.n + 31) / 32;
    const uint64_t waves = (uint64_t)gridDim.x * 4;
    for (uint64_t tile = (uint64_t)blockIdx.x * 4 + (tid >> 6); tile < tiles; tile += waves) {
        uint4 hb[8];
        zg_f32x16 z2[4];
        const float y = mlp_fwd_tile(A, zg_opaque(s_w1), zg_opaque(s_w2), s_b1, s_b2, s_w3, b3, tile * 32, lane, hb, z2);
        const uint64_t node = tile * 32 + (lane & 31);
        if (lane < 32 && node < A.n) A.y[node] = y;
    }
}

// x from lane (lane ^ O) of the same half-wave (ds_swizzle bit mode: and
// mask 0x1F, xor mask O).
template <int O>
__device__ __forceinline__ float zg_xor32(float x)
{
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), 0x1F | (O << 10)));
}

template <int O>
__device__ __forceinline__ void zg_halve(float *a, int c)
{
    const bool up = (c & O) != 0;
#pragma unroll
    for (int k = 0; k < O; k++) {
        const float send = up ? a[k] : a[k + O];
        const float keep = up ? a[k + O] : a[k];
        a[k] = keep + zg_xor32<O>(send);
    }
}

// Sums a[0..31] over the 32 lanes of each half-wave by recursive halving
// (31 exchanges): afterwards lane c of a half holds the sum of slot c.
__device__ __forceinline__ float zg_half_reduce32(float (&a)[32], int c)
{
    zg_halve<16>(a, c);
    zg_halve<8>(a, c);
    zg_halve<4>(a, c);
    zg_halve<2>(a, c);
    zg_halve<1>(a, c);
    return a[0];
}

// Slot r of the 64 result registers of a lane: M-tile r/16, register r%16.
__device__ __forceinline__ int zg_slot_neuron(int r, int h) { return zg_neuron(r >> 4, r & 15, h); }

__global__ __launch_bounds__(256, 1) void mlp_grad_kernel(MlpArgs A)
{
    __shared__ uint16_t sH1[ZG_MLP_H * ZG_MLP_LDS_STRIDE];  // relu(z1) [neuron][node]
    __shared__ uint16_t sD2[ZG_MLP_H * ZG_MLP_LDS_STRIDE];  // dz2 [neuron][node]
    __shared__ uint16_t sD1[ZG_MLP_H * ZG_MLP_LDS_STRIDE];  // dz1 [neuron][node]
    __shared__ uint16_t sX[ZG_NFEAT * ZG_MLP_LDS_STRIDE];   // x [feature][node]
    __shared__ float s_b1[ZG_MLP_H], s_b2[ZG_MLP_H], s_w3[ZG_MLP_H];
    __shared__ float s_acc[3][ZG_MLP_H];                    // b1, b2, w3 gradients
    __shared__ float s_red[2][4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = lane >> 5, col = lane & 31;
    if (tid < ZG_MLP_H) {
        s_b1[tid] = A.p[ZG_MLP_B1 + tid];
        s_b2[tid] = A.p[ZG_MLP_B2 + tid];
        s_w3[tid] = A.p[ZG_MLP_W3 + tid];
        s_acc[0][tid] = s_acc[1][tid] = s_acc[2][tid] = 0.f;
    }
    __syncthreads();
    const float b3 = A.p[ZG_MLP_B3];
    const zg_f32x16 zero = {};
    zg_f32x16 gw2[4] = {zero, zero, zero, zero};
    zg_f32x16 gw1 = zero;
    float g_b1[2] = {0.f, 0.f}, g_b2[2] = {0.f, 0.f}, g_w3[2] = {0.f, 0.f};
    float g_b3 = 0.f, g_loss = 0.f;
    {
        const uint64_t c0 = (uint64_t)blockIdx.x * 128;
        const uint64_t n0 = c0 + 32 * w;
        const uint64_t node = n0 + col;
        const bool valid = node < A.n;
        const int nl = 32 * w + col;
        const uint16_t vmask = valid ? 1 : 0;
        uint4 hb[8];
        zg_f32x16 z2[4];
        const float y = mlp_fwd_tile(A, zg_opaque(A.pw1), zg_opaque(A.pw2f), s_b1, s_b2, s_w3, b3, n0, lane, hb, z2);
        const uint4 *pw2b = zg_opaque(A.pw2b);
        const float err = valid ? y - A.target[node] : 0.f;
        const float dy = A.scale * err;
        if (h == 0) {
            g_loss += err * err;
            g_b3 += dy;
        }
        // w3 gradient terms dy * relu(z2), slots 0..31 then 32..63
#pragma unroll
        for (int q = 0; q < 2; q++) {
            float a[32];
#pragma unroll
            for (int r = 0; r < 32; r++) a[r] = dy * fmaxf(z2[2 * q + (r >> 4)][r & 15], 0.f);
            g_w3[q] += zg_half_reduce32(a, col);
        }
        // dz2 = dy * w3 * relu'(z2) in place; b2 gradient; BF16 fragments
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 16; i++) z2[t][i] = dy * s_w3[zg_neuron(t, i, h)] * zg_nonzero(fmaxf(z2[t][i], 0.f));
#pragma unroll
        for (int q = 0; q < 2; q++) {
            float a[32];
#pragma unroll
            for (int r = 0; r < 32; r++) a[r] = z2[2 * q + (r >> 4)][r & 15];
            g_b2[q] += zg_half_reduce32(a, col);
        }
        uint4 db[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            ZgFrag f;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                f.h[j] = zg_f2bf(z2[k >> 1][8 * (k & 1) + j]);
                const int m = zg_neuron(k >> 1, 8 * (k & 1) + j, h);
                sD2[m * ZG_MLP_LDS_STRIDE + nl] = f.h[j];
            }
            db[k] = f.u;
        }
        // dh1 = W2 dz2; dz1 = dh1 * relu'(z1) (z1 > 0 exactly where relu(z1) is non-zero)
        zg_f32x16 d1[4] = {zero, zero, zero, zero};
#pragma unroll
        for (int k = 0; k < 8; k++)
#pragma unroll
            for (int t = 0; t < 4; t++) d1[t] = zg_mfma(pw2b[(t * 8 + k) * 64 + lane], db[k], d1[t]);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            ZgFrag f;
            f.u = hb[k];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int t = k >> 1, i = 8 * (k & 1) + j;
                const int m = zg_neuron(t, i, h);
                d1[t][i] *= (float)min((uint32_t)f.h[j], 1u);
                sH1[m * ZG_MLP_LDS_STRIDE + nl] = f.h[j] * vmask;
                sD1[m * ZG_MLP_LDS_STRIDE + nl] = zg_f2bf(d1[t][i]);
            }
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            float a[32];
#pragma unroll
            for (int r = 0; r < 32; r++) a[r] = d1[2 * q + (r >> 4)][r & 15];
            g_b1[q] += zg_half_reduce32(a, col);
        }
        if (h == 0) {
#pragma unroll
            for (int f = 0; f < ZG_NFEAT; f++)
                sX[f * ZG_MLP_LDS_STRIDE + nl] = valid ? zg_f2bf(A.x[node * ZG_NFEAT + f]) : 0;
        }
        __syncthreads();
        // dW2[in][out] (M-tile w of inputs) and dW1^T[neuron][feature]
        // (M-tile w of neurons), K = the chunk's 128 nodes
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int kn = 16 * k + 8 * h;
            const uint4 a2 = *(const uint4 *)&sH1[(32 * w + col) * ZG_MLP_LDS_STRIDE + kn];
#pragma unroll
            for (int to = 0; to < 4; to++) {
                const uint4 b2 = *(const uint4 *)&sD2[(32 * to + col) * ZG_MLP_LDS_STRIDE + kn];
                gw2[to] = zg_mfma(a2, b2, gw2[to]);
            }
            const uint4 a1 = *(const uint4 *)&sD1[(32 * w + col) * ZG_MLP_LDS_STRIDE + kn];
            const uint4 b1 = col < ZG_NFEAT ? *(const uint4 *)&sX[col * ZG_MLP_LDS_STRIDE + kn] : make_uint4(0, 0, 0, 0);
            gw1 = zg_mfma(a1, b1, gw1);
        }
        __syncthreads();
    }
    // small-vector gradients: lane c of half h holds slots c and 32 + c
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int m = zg_slot_neuron(32 * q + col, h);
        atomicAdd(&s_acc[0][m], g_b1[q]);
        atomicAdd(&s_acc[1][m], g_b2[q]);
        atomicAdd(&s_acc[2][m], g_w3[q]);
    }
    g_loss = zg_wave_sum((double)g_loss);
    g_b3 = zg_wave_sum((double)g_b3);
    if (lane == 0) {
        s_red[0][w] = g_loss;
        s_red[1][w] = g_b3;
    }
    __syncthreads();
    float *out = A.part + (size_t)blockIdx.x * (ZG_MLP_P + 1);
#pragma unroll
    for (int to = 0; to < 4; to++)
#pragma unroll
        for (int i = 0; i < 16; i++)
            out[ZG_MLP_W2 + zg_neuron(w, i, h) * ZG_MLP_H + 32 * to + col] = gw2[to][i];
    if (col < ZG_NFEAT)
#pragma unroll
        for (int i = 0; i < 16; i++) out[ZG_MLP_W1 + col * ZG_MLP_H + zg_neuron(w, i, h)] = gw1[i];
    if (tid < ZG_MLP_H) {
        out[ZG_MLP_B1 + tid] = s_acc[0][tid];
        out[ZG_MLP_B2 + tid] = s_acc[1][tid];
        out[ZG_MLP_W3 + tid] = s_acc[2][tid];
    }
    if (tid == 0) {
        out[ZG_MLP_B3] = s_red[1][0] + s_red[1][1] + s_red[1][2] + s_red[1][3];
        out[ZG_MLP_P] = s_red[0][0] + s_red[0][1] + s_red[0][2] + s_red[0][3];
    }
}

// grad[p] = sum over blocks (fixed order); loss = mean squared error.
__global__ void mlp_reduce_kernel(const float *part, uint32_t blocks, uint64_t n, float *grad, double *loss)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > ZG_MLP_P) return;
    double s = 0.0;
    for (uint32_t b = 0; b < blocks; b++) s += part[(size_t)b * (ZG_MLP_P + 1) + p];
    if (p < ZG_MLP_P) grad[p] = (float)s;
    else *loss = n ? s / (double)n : 0.0;
}

__global__ void mlp_adam_kernel(float *p, float *m, float *v, const float *g, float lr, float b1, float b2, float eps,
                                float c1, float c2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ZG_MLP_P) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static int mlp_pack(zg_mlp *m, hipStream_t s)
{
    const int threads = 4 * 64 + 2 * 4 * 8 * 64;
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((threads + 255) / 256), dim3(256), 0, s, m->params, m->pw1, m->pw2f,
                       m->pw2b);
    return zg_hip_code(hipGetLastError());
}

static MlpArgs mlp_args(const zg_mlp *m)
{
    MlpArgs A;
    memset(&A, 0, sizeof(A));
    A.pw1 = m->pw1;
    A.pw2f = m->pw2f;
    A.pw2b = m->pw2b;
    A.p = m->params;
    A.part = m->part;
    return A;
}

// SplitMix64 stream for the initial weights.
static inline uint64_t mlp_splitmix(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" void zg_mlp_destroy(zg_mlp *m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    zg_dfree(m->params); zg_dfree(m->m); zg_dfree(m->v); zg_dfree(m->grad); zg_dfree(m->part); zg_dfree(m->loss);
    zg_dfree(m->pw1); zg_dfree(m->pw2f); zg_dfree(m->pw2b);
    delete m;
}

extern "C" int zg_mlp_create(uint64_t seed, zg_mlp **out)
{
    if (!out) return ZG_ERR_INVALID_ARG;
    *out = nullptr;
    int dev = 0;
    ZG_TRY(hipGetDevice(&dev));
    zg_mlp *m = new (std::nothrow) zg_mlp;
    if (!m) return ZG_ERR_OOM;
    m->device = dev;
    int rc = zg_dmalloc(&m->params, ZG_MLP_P);
    if (rc == ZG_OK) rc = zg_dmalloc(&m->m, ZG_MLP_P);
    if (rc == ZG_OK) rc = zg_dmalloc(&m->v, ZG_MLP_P);
    if (rc == ZG_OK) rc = zg_dmalloc(&m->grad, ZG_MLP_P);
    if (rc == ZG_OK) rc = zg_dmalloc(&m->part, (size_t)ZG_MLP_BLOCKS * (ZG_MLP_P + 1));
    if (rc == ZG_OK) rc = zg_dmalloc(&m->loss, 1);
    if (rc == ZG_OK) rc = zg_dmalloc(&m->pw1, 4 * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&m->pw2f, 4 * 8 * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&m->pw2b, 4 * 8 * 64);
    if (rc != ZG_OK) {
        zg_mlp_destroy(m);
        return rc;
    }
    // He-uniform hidden layers, uniform(+-sqrt(3/128)) output, zero biases
    std::vector<float> p(ZG_MLP_P, 0.f);
    uint64_t st = seed;
    auto uni = [&](float bound) {
        const double u = (double)(mlp_splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
        return (float)((2.0 * u - 1.0) * bound);
    };
    for (int i = 0; i < ZG_NFEAT * ZG_MLP_H; i++) p[ZG_MLP_W1 + i] = uni(sqrtf(6.f / ZG_NFEAT));
    for (int i = 0; i < ZG_MLP_H * ZG_MLP_H; i++) p[ZG_MLP_W2 + i] = uni(sqrtf(6.f / ZG_MLP_H));
    for (int i = 0; i < ZG_MLP_H; i++) p[ZG_MLP_W3 + i] = uni(sqrtf(3.f / ZG_MLP_H));
    rc = zg_hip_code(hipMemcpy(m->params, p.data(), sizeof(float) * ZG_MLP_P, hipMemcpyHostToDevice));
    if (rc == ZG_OK) rc = zg_hip_code(hipMemset(m->m, 0, sizeof(float) * ZG_MLP_P));
    if (rc == ZG_OK) rc = zg_hip_code(hipMemset(m->v, 0, sizeof(float) * ZG_MLP_P));
    if (rc == ZG_OK) rc = mlp_pack(m, nullptr);
    if (rc == ZG_OK) rc = zg_hip_code(hipDeviceSynchronize());
    if (rc != ZG_OK) {
        zg_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return ZG_OK;
}

extern "C" int zg_mlp_get_params(zg_mlp *m, float *out)
{
    if (!m) return ZG_ERR_INVALID_HANDLE;
    if (!out) return ZG_ERR_INVALID_ARG;
    ZG_TRY(hipSetDevice(m->device));
    ZG_TRY(hipMemcpy(out, m->params, sizeof(float) * ZG_MLP_P, hipMemcpyDeviceToHost));
    return ZG_OK;
}

extern "C" int zg_mlp_set_params(zg_mlp *m, const float *in)
{
    if (!m) return ZG_ERR_INVALID_HANDLE;
    if (!in) return ZG_ERR_INVALID_ARG;
    ZG_TRY(hipSetDevice(m->device));
    ZG_TRY(hipMemcpy(m->params, in, sizeof(float) * ZG_MLP_P, hipMemcpyHostToDevice));
    ZG_TRY(hipMemset(m->m, 0, sizeof(float) * ZG_MLP_P));
    ZG_TRY(hipMemset(m->v, 0, sizeof(float) * ZG_MLP_P));
    m->step = 0;
    ZG_TRY_RC(mlp_pack(m, nullptr));
    ZG_TRY(hipDeviceSynchronize());
    return ZG_OK;
}

extern "C" int zg_mlp_forward(zg_mlp *m, const float *d_x, uint64_t n, float *d_y, void *stream)
{
    if (!m) return ZG_ERR_INVALID_HANDLE;
    if (n && (!d_x || !d_y)) return ZG_ERR_INVALID_ARG;
    if (n == 0) return ZG_OK;
    ZG_TRY(hipSetDevice(m->device));
    hipStream_t s = (hipStream_t)stream;
    MlpArgs A = mlp_args(m);
    A.x = d_x;
    A.n = n;
    A.y = d_y;
    const uint64_t tiles = (n + 31) / 32;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((tiles + 3) / 4, 2048);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, s, A);
    ZG_TRY(hipGetLastError());
    ZG_TRY(hipStreamSynchronize(s));
    return ZG_OK;
}

extern "C" int zg_mlp_gradient(zg_mlp *m, const float *d_x, const float *d_target, uint64_t n, float *grad_out,
                               double *loss_out, void *stream)
{
    if (!m) return ZG_ERR_INVALID_HANDLE;
    if (n == 0 || !d_x || !d_target) return ZG_ERR_INVALID_ARG;
    ZG_TRY(hipSetDevice(m->device));
    hipStream_t s = (hipStream_t)stream;
    MlpArgs A = mlp_args(m);
    A.x = d_x;
    A.target = d_target;
    A.n = n;
    A.scale = 2.0f / (float)n;
    if (n > (uint64_t)ZG_MLP_BLOCKS * 128) return ZG_ERR_INVALID_ARG;  // batch limit
    const uint32_t blocks = (uint32_t)((n + 127) / 128);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(256), 0, s, A);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((ZG_MLP_P + 1 + 255) / 256), dim3(256), 0, s, m->part, blocks, n,
                       m->grad, m->loss);
    ZG_TRY(hipGetLastError());
    if (grad_out) ZG_TRY(hipMemcpyAsync(grad_out, m->grad, sizeof(float) * ZG_MLP_P, hipMemcpyDeviceToHost, s));
    if (loss_out) ZG_TRY(hipMemcpyAsync(loss_out, m->loss, sizeof(double), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    return ZG_OK;
}

extern "C" int zg_mlp_adam(zg_mlp *m, float lr, float beta1, float beta2, float eps, void *stream)
{
    if (!m) return ZG_ERR_INVALID_HANDLE;
    if (!(lr >= 0.f) || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
        return ZG_ERR_INVALID_ARG;
    ZG_TRY(hipSetDevice(m->device));
    hipStream_t s = (hipStream_t)stream;
    m->step++;
    const float c1 = 1.f - powf(beta1, (float)m->step), c2 = 1.f - powf(beta2, (float)m->step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((ZG_MLP_P + 255) / 256), dim3(256), 0, s, m->params, m->m, m->v, m->grad,
                       lr, beta1, beta2, eps, c1, c2);
    ZG_TRY(hipGetLastError());
    ZG_TRY_RC(mlp_pack(m, s));
    ZG_TRY(hipStreamSynchronize(s));
    return ZG_OK;
}
