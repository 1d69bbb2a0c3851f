// Two-bounce GI paths and the learned GI radiance cache (include/vf_gicache.h).
//
// The network (frequency encoding -> 5 x 64 ReLU -> RGB) runs on the matrix cores with
// v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation.  Lane layout of that instruction
// (lane l = 0..63, g = l / 16, c = l % 16):  A (16x32): A[c][8g + i];  B (32x16): B[8g + i][c];
// D (16x16): D[4g + i][c].  Every layer is computed transposed, D = W^T (out x in) . X^T (in x samples),
// so that a lane's B operand is 8 consecutive features of one sample: one 16-byte LDS read from a
// [sample][feature] activation tile.  Weights live in "fragment" order (each lane's 8 BF16 of one
// 16x32 operand block are contiguous), written by the optimizer kernel after every update, so a weight
// operand is one coalesced 16-byte global load per lane.
//
// A training wavefront owns 64 samples: the forward pass keeps every layer's BF16 activations in LDS,
// the backward pass overwrites each activation tile in place with its gradient once the weight gradient
// that needs it has been accumulated.  Weight gradients are MFMA products with the samples as the K
// dimension; each training wavefront writes its own partial, and the optimizer kernel sums the partials
// in a fixed order (deterministic), then applies Adam to the FP32 master weights.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/vf_gicache.h"
#include "vf_ctx.h"
#include "vf_gi.h"

using vf::blocks_for;
using vf::launch_check;

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kP = VF_GICACHE_PARAMS;
constexpr int kPP = (kP + 3) & ~3;      // row stride of the gradient partials (float4 aligned)
constexpr int kSlices = 16;             // first-level slices of the partial reduction
constexpr int kIn = VF_GICACHE_INPUTS;   // 78 features padded to 96
// Layer shapes: (96, 64), (64, 64) x 4, (64, 3).
__host__ __device__ constexpr int in_dim(int l) { return l == 0 ? 96 : 64; }
__host__ __device__ constexpr int out_dim(int l) { return l == 5 ? 3 : 64; }
__host__ __device__ constexpr int w_off(int l) { return l == 0 ? 0 : 96 * 64 + 64 + (l - 1) * (64 * 64 + 64); }
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + in_dim(l) * out_dim(l); }
static_assert(b_off(5) + 3 == kP, "parameter count");
// The training kernel's gradient partials keep each layer's weights in the order its MFMA tiles hold them:
// weight (row, out) of layer l < 5 at w_off(l) + ((row / 16 * 4 + out / 16) * 64 + lane) * 4 + row % 4 with
// lane = (row % 16 / 4) * 16 + out % 16; of layer 5 at w_off(5) + ((row / 4) * 3 + out) * 4 + row % 4.
// Biases keep their places.
__host__ __device__ constexpr int part_index(int p) {
    if (p >= w_off(5)) {
        if (p >= b_off(5)) return p;
        const int q = p - w_off(5), row = q / 3, o = q % 3;
        return w_off(5) + ((row >> 2) * 3 + o) * 4 + (row & 3);
    }
    const int l = p < w_off(1) ? 0 : 1 + (p - w_off(1)) / (64 * 64 + 64);
    if (p >= b_off(l)) return p;
    const int q = p - w_off(l), row = q >> 6, o = q & 63;
    return w_off(l) + (((row >> 4) * 4 + (o >> 4)) * 64 + ((row & 15) >> 2) * 16 + (o & 15)) * 4 + (row & 3);
}
constexpr bool part_index_is_permutation() {
    bool seen[kP] = {};
    for (int p = 0; p < kP; ++p) {
        const int q = part_index(p);
        if (q < 0 || q >= kP || seen[q]) return false;
        seen[q] = true;
    }
    return true;
}
static_assert(part_index_is_permutation(), "the partials' tile order is a permutation of the parameters");
static_assert(w_off(1) % 4 == 0 && w_off(2) % 4 == 0 && w_off(3) % 4 == 0 && w_off(4) % 4 == 0 && w_off(5) % 4 == 0,
              "16-byte tile stores");
// Forward fragments of layer l: (out padded to 16)/16 row tiles x in/32 K blocks; backward fragments
// (layers 1..5): in/16 row tiles x (out padded to 32)/32 K blocks.  1 KB each.
__host__ __device__ constexpr int fwd_mt(int l) { return l == 5 ? 1 : 4; }
__host__ __device__ constexpr int fwd_kb(int l) { return in_dim(l) / 32; }
__host__ __device__ constexpr int fwd_base(int l) { return l == 0 ? 0 : 12 + (l - 1) * 8; }
__host__ __device__ constexpr int bwd_kb(int l) { return l == 5 ? 1 : 2; }
__host__ __device__ constexpr int bwd_base(int l) { return (l - 1) * 8; }
constexpr int kFwdFrags = 46, kBwdFrags = 36;
// The inference kernel keeps activations in registers: the output tile of a layer (lane g, c holds
// features 16 mt + 4 g + i of sample c) is the next layer's B operand when K slot j of lane group g in
// K block kb stands for input feature perm(kb, g, j) = (2 kb + j / 4) * 16 + 4 g + j % 4.  Its
// fragments of layers 1..5 are packed in that K order (layer 0 reads the encoding in natural order).
__host__ __device__ constexpr int perm_in(int kb, int g, int j) { return (2 * kb + (j >> 2)) * 16 + 4 * g + (j & 3); }
static_assert(fwd_base(5) + fwd_mt(5) * fwd_kb(5) == kFwdFrags, "forward fragments");
static_assert(bwd_base(5) + 4 * bwd_kb(5) == kBwdFrags, "backward fragments");

// LDS row strides (BF16 elements) of the activation tiles: rows stay 16-byte aligned.
constexpr int kS0 = 104;   // encoded input, 96 features
constexpr int kSH = 72;    // hidden, 64 features
constexpr int kSG = 40;    // output gradient, 3 used of 32
// BF16 elements of a training workgroup's tiles: the encoded input, five hidden activations, the output
// gradient and two hidden gradients.
constexpr int kTrainLds = 64 * kS0 + 5 * 64 * kSH + 64 * kSG + 2 * 64 * kSH;
constexpr int kBiasLds = 5 * 64 + 4;

// sin(pi u), cos(pi u): u - rint(u) is exact, then Taylor polynomials on [-pi/2, pi/2] (terms to x^11
// and x^12; truncation < 6e-8) and the sign of the half period.
__device__ __forceinline__ void sincospi(float u, float &s, float &c) {
    const float n = rintf(u);
    const float x = (u - n) * 3.14159265358979f;
    const float x2 = x * x;
    float ps = x2 * -2.50521084e-08f + 2.75573192e-06f;
    ps = ps * x2 + -1.98412698e-04f;
    ps = ps * x2 + 8.33333333e-03f;
    ps = ps * x2 + -1.66666667e-01f;
    s = x + x * x2 * ps;
    float pc = x2 * 2.08767570e-09f + -2.75573192e-07f;
    pc = pc * x2 + 2.48015873e-05f;
    pc = pc * x2 + -1.38888889e-03f;
    pc = pc * x2 + 4.16666667e-02f;
    pc = pc * x2 + -0.5f;
    c = 1.0f + x2 * pc;
    if ((int32_t)n & 1) { s = -s; c = -c; }
}

// Feature f of a record: f = 24 a + 2 k (+1): sin (cos) of pi * pos[a] * 2^(k - shift); 72..74 normal,
// 75..77 view, 78..95 zero.  Group q (0..11) = features 8q..8q+7 = four consecutive bands of one
// coordinate: the first from sincospi, the next three by the double-angle formulas (sin 2t = 2 sin t
// cos t, cos 2t = (cos t - sin t)(cos t + sin t)), which at most doubles the error per band (< 2e-6).
__device__ __forceinline__ void encode_group(const float4 &r0, const float4 &r1, const float4 &r2, int shift, int q,
                                             float f[8]) {
    if (q < 9) {
        const int a = q / 3, k0 = (q % 3) * 4;
        const float x = a == 0 ? r0.x : (a == 1 ? r0.y : r0.z);
        float s, c;
        sincospi(ldexpf(x, k0 - shift), s, c);
        f[0] = s;
        f[1] = c;
#pragma unroll
        for (int j = 1; j < 4; ++j) {
            const float s2 = 2.0f * s * c, c2 = (c - s) * (c + s);
            s = s2;
            c = c2;
            f[2 * j] = s;
            f[2 * j + 1] = c;
        }
    } else if (q == 9) {
        f[0] = r1.x; f[1] = r1.y; f[2] = r1.z;
        f[3] = r2.x; f[4] = r2.y; f[5] = r2.z;
        f[6] = 0.0f; f[7] = 0.0f;
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) f[j] = 0.0f;
    }
}

__device__ __forceinline__ bf16x8 to_bf16x8(const float f[8]) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (bf16)f[j];
    return v;
}

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// A training workgroup is four wavefronts sharing one 64-sample chunk at a time.  Wave w owns row tile w
// (features 16 w .. 16 w + 15) of every 64-wide product of the forward and backward passes, and the
// weight-gradient tiles (k, w) of every layer (row tile k of the layer's inputs, column tile w of its
// outputs; tile (w, 0) of the 3-output layer).
constexpr int kTrainWaves = 4;

// What a training wavefront reads of the weights, loaded into registers once per launch (not at the start
// of each of the eleven products): its forward fragments (layers 0..4 and the output layer), its backward
// fragments (layers 5..1) and its biases.
struct TrainWeights {
    bf16x8 f0[3];      // layer 0, K blocks 0..2
    bf16x8 fh[4][2];   // layers 1..4
    bf16x8 fo[2];      // output layer
    bf16x8 b5[1];      // backward of layer 5 (one K block)
    bf16x8 bh[4][2];   // backward of layers 1..4
    float bz[5][4];    // hidden biases 16 w + 4 g + i
    float bo[3];       // output biases
};

__device__ __forceinline__ void load_train_weights(const bf16x8 *ff, const bf16x8 *bf, const float *w, int lane, int wv,
                                                   TrainWeights &T) {
    const int g = lane >> 4;
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) T.f0[kb] = ff[(fwd_base(0) + wv * 3 + kb) * 64 + lane];
#pragma unroll
    for (int l = 1; l < 5; ++l)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            T.fh[l - 1][kb] = ff[(fwd_base(l) + wv * 2 + kb) * 64 + lane];
            T.bh[l - 1][kb] = bf[(bwd_base(l) + wv * 2 + kb) * 64 + lane];
        }
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) T.fo[kb] = ff[(fwd_base(5) + kb) * 64 + lane];
    T.b5[0] = bf[(bwd_base(5) + wv) * 64 + lane];
#pragma unroll
    for (int l = 0; l < 5; ++l)
#pragma unroll
        for (int i = 0; i < 4; ++i) T.bz[l][i] = w[b_off(l) + wv * 16 + 4 * g + i];
#pragma unroll
    for (int o = 0; o < 3; ++o) T.bo[o] = w[b_off(5) + o];
}

// Row tile mt of hidden layer l (KB = in/32): act_out = bf16(relu(W^T . act_in + b)) for the chunk's 64
// samples.
template <int KB>
__device__ __forceinline__ void hidden_tile(const bf16 *act_in, int s_in, bf16 *act_out, const bf16x8 (&a)[KB],
                                            const float (&bz)[4], int lane, int mt) {
    const int col = lane & 15, g = lane >> 4;
    f32x4 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(act_in + (nt * 16 + col) * s_in + kb * 32 + 8 * g);
            acc[nt] = mfma(a[kb], b, acc[nt]);
        }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        bf16x4 o;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = acc[nt][i] + bz[i];
            o[i] = (bf16)(v > 0.0f ? v : 0.0f);
        }
        *reinterpret_cast<bf16x4 *>(act_out + (nt * 16 + col) * kSH + mt * 16 + 4 * g) = o;
    }
}

// Output layer (64 -> 3, rows padded to 16) for samples 16 nt .. 16 nt + 15: y[i] = output 4g + i of
// sample 16 nt + col (only g = 0, i < 3 are outputs).
__device__ __forceinline__ void output_tile(const bf16 *act5, const TrainWeights &T, int lane, int nt, f32x4 &y) {
    const int col = lane & 15, g = lane >> 4;
    const bf16 *row = act5 + (nt * 16 + col) * kSH + 8 * g;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma(T.fo[0], *reinterpret_cast<const bf16x8 *>(row), acc);
    acc = mfma(T.fo[1], *reinterpret_cast<const bf16x8 *>(row + 32), acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = acc[i] + ((g == 0 && i < 3) ? T.bo[i < 3 ? i : 0] : 0.0f);
}

// The forward pass of the chunk's 64 encoded samples; hidden activations stay in act[1..5]; wave w
// computes the outputs of samples 16 w .. 16 w + 15.
__device__ __forceinline__ void forward(bf16 *act0, bf16 *const act[6], const TrainWeights &T, int lane, int wv,
                                        f32x4 &y) {
    hidden_tile<3>(act0, kS0, act[1], T.f0, T.bz[0], lane, wv);
    __syncthreads();
#pragma unroll
    for (int l = 1; l < 5; ++l) {
        hidden_tile<2>(act[l], kSH, act[l + 1], T.fh[l - 1], T.bz[l], lane, wv);
        __syncthreads();
    }
    output_tile(act[5], T, lane, wv, y);
}

// Row tile mt of the gradient into the previous layer: g_out := bf16((W_l . G^T)^T * (act[l] > 0)), where G
// (the gradient at layer l's output, KB = out/32 blocks) is read from g_in and act is layer l's input.
template <int KB>
__device__ __forceinline__ void backward_tile(const bf16 *g_in, int s_g, const bf16 *act, bf16 *g_out,
                                              const bf16x8 (&a)[KB], int lane, int mt) {
    const int col = lane & 15, g = lane >> 4;
    f32x4 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(g_in + (nt * 16 + col) * s_g + kb * 32 + 8 * g);
            acc[nt] = mfma(a[kb], b, acc[nt]);
        }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int off = (nt * 16 + col) * kSH + mt * 16 + 4 * g;
        const bf16x4 m = *reinterpret_cast<const bf16x4 *>(act + off);
        bf16x4 o;
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = (float)m[i] > 0.0f ? (bf16)acc[nt][i] : (bf16)0.0f;
        *reinterpret_cast<bf16x4 *>(g_out + off) = o;
    }
}

// Weight-gradient tile (mt, nt) over the chunk's 64 samples, added to acc: dW[in][out] += sum_s act[s][in]
// G[s][out] (MFMA with the samples as K).
__device__ __forceinline__ void wgrad_tile(const bf16 *act, int s_act, const bf16 *gr, int s_g, int mt, int nt,
                                           int lane, f32x4 &acc) {
    const int col = lane & 15, g = lane >> 4;
    bf16x8 a[2], b[2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            a[kb][i] = act[(kb * 32 + 8 * g + i) * s_act + mt * 16 + col];
            b[kb][i] = gr[(kb * 32 + 8 * g + i) * s_g + nt * 16 + col];
        }
    const f32x4 t = mfma(a[0], b[0], acc);
    acc = mfma(a[1], b[1], t);
}

// Bias gradient of column tile nt over the chunk, sum_s G[s][16 nt + col], added to db (set by the first
// chunk): an MFMA product with an all-ones A operand, whose every row is the column sum; lane (g, col)
// gets output 16 nt + col.  (Its B operand is the one the weight-gradient tiles of column tile nt read.)
__device__ __forceinline__ void bias_tile(const bf16 *gr, int s_g, int nt, int lane, bool first, float &db) {
    const int col = lane & 15, g = lane >> 4;
    bf16x8 one, b[2];
#pragma unroll
    for (int i = 0; i < 8; ++i) one[i] = (bf16)1.0f;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int i = 0; i < 8; ++i) b[kb][i] = gr[(kb * 32 + 8 * g + i) * s_g + nt * 16 + col];
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma(one, b[0], acc);
    acc = mfma(one, b[1], acc);
    db = first ? acc[0] : db + acc[0];
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct TrainArgs {
    const float4 *samples;   // n records, 4 x float4 each
    long long n;
    int32_t n_chunks;        // ceil(n / 64)
    int32_t shift;
    float inv_n, loss_eps;
    const bf16x8 *ffrag, *bfrag;
    const float *w;          // FP32 master weights (biases are read from here)
    float *part;             // [gridDim.x][kP]
    float *loss_part;        // [gridDim.x]
    int32_t *valid_part;     // [gridDim.x]
};

// The training record of sample s of chunk ch (zeros past the end).
__device__ __forceinline__ void load_record(const TrainArgs &A, int ch, int s, float4 &r0, float4 &r1, float4 &r2,
                                            float4 &r3) {
    const long long sidx = (long long)ch * 64 + s;
    r0 = make_float4(0.f, 0.f, 0.f, 0.f);
    r1 = r0; r2 = r0; r3 = r0;
    if (ch < A.n_chunks && sidx < A.n) {
        const float4 *rec = A.samples + sidx * 4;
        r0 = rec[0]; r1 = rec[1]; r2 = rec[2]; r3 = rec[3];
    }
}

// Forward + backward of the relative L2 loss over 64-sample chunks, strided over the grid.  The weight
// gradients of the workgroup's chunks are summed, chunk by chunk, in registers (each wave its own tiles,
// bias sums by one wave per layer) and written to the workgroup's partial once, at the end.
__global__ void __launch_bounds__(256) k_mlp_train(TrainArgs A) {
    __shared__ __attribute__((aligned(16))) bf16 lds[kTrainLds];
    __shared__ float4 tgt[64];    // per sample: target RGB, valid
    __shared__ float2 lterm[64];  // per sample: its loss term, valid
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    static_assert(kTrainWaves == 4, "one row tile of each 64-wide layer per wave");
    // encoding: four lanes per sample, each three of its twelve feature groups.  The first chunk's records
    // are requested before the weights, so the encoding waits for them only; each later chunk's while the
    // one before it is trained.
    const int es = wv * 16 + (lane >> 2), pq = lane & 3;
    float4 r0, r1, r2, r3;
    load_record(A, blockIdx.x, es, r0, r1, r2, r3);
    TrainWeights T;
    load_train_weights(A.ffrag, A.bfrag, A.w, lane, wv, T);
    bf16 *act0 = lds;
    bf16 *act[6];
    act[0] = act0;
#pragma unroll
    for (int l = 1; l < 6; ++l) act[l] = lds + 64 * kS0 + (l - 1) * 64 * kSH;
    bf16 *g6 = lds + 64 * kS0 + 5 * 64 * kSH;
    bf16 *const gb[2] = {g6 + 64 * kSG, g6 + 64 * kSG + 64 * kSH};   // hidden gradients, alternating
    for (int j = tid; j < 64 * kSG; j += 256) g6[j] = (bf16)0.0f;
    // gradient accumulators: layer 0 tiles (k, wv) k < 6, layers 1..4 tiles (k, wv) k < 4, layer 5 tile
    // (wv, 0); bias sums (layer l's by wave l & 3; layer 5's by wave 3)
    f32x4 g0[6], gh[4][4], g5;
    float db[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) g0[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int k = 0; k < 4; ++k) gh[l][k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    g5 = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < 6; ++l) db[l] = 0.0f;
    float loss = 0.0f, nvalid = 0.0f;
    bool first = true;
    for (int ch = blockIdx.x; ch < A.n_chunks; ch += gridDim.x) {
        {
            const bool present = (long long)ch * 64 + es < A.n;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float f[8];
                encode_group(r0, r1, r2, A.shift, pq + 4 * k, f);
                *reinterpret_cast<bf16x8 *>(act0 + es * kS0 + 8 * (pq + 4 * k)) = to_bf16x8(f);
            }
            if (pq == 0) tgt[es] = make_float4(r3.x, r3.y, r3.z, (present && __float_as_uint(r0.w) != 0u) ? 1.0f : 0.0f);
        }
        // the next chunk's records, in flight while this one is trained
        load_record(A, ch + gridDim.x, es, r0, r1, r2, r3);
        __syncthreads();
        f32x4 y;
        forward(act0, act, T, lane, wv, y);
        // loss gradient: lanes 0..15 of wave w hold outputs 0..2 of samples 16 w + lane
        if (lane < 16) {
            const int s = wv * 16 + lane;
            const float4 t = tgt[s];
            const float y0 = y[0], y1 = y[1], y2 = y[2];
            const float mean = ((y0 + y1) + y2) / 3.0f;
            const float den = mean * mean + A.loss_eps;
            const float d0 = y0 - t.x, d1 = y1 - t.y, d2 = y2 - t.z;
            bf16x4 o = {(bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f};
            lterm[s] = make_float2(t.w != 0.0f ? ((d0 * d0 + d1 * d1) + d2 * d2) / den / 3.0f : 0.0f, t.w);
            if (t.w != 0.0f) {
                const float k = (2.0f / 3.0f) * A.inv_n / den;
                o[0] = (bf16)(d0 * k);
                o[1] = (bf16)(d1 * k);
                o[2] = (bf16)(d2 * k);
            }
            *reinterpret_cast<bf16x4 *>(g6 + s * kSG) = o;
        }
        __syncthreads();
        // the chunk's loss terms into wave 0's lane sums, samples lane, 16 + lane, 32 + lane, 48 + lane in turn
        if (wv == 0 && lane < 16) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const float2 lt = lterm[nt * 16 + lane];
                if (lt.y != 0.0f) {
                    nvalid += 1.0f;
                    loss += lt.x;
                }
            }
        }
        // Layer 5 (64 -> 3), then 4..1, then 0: one phase per layer computes its weight and bias gradients
        // and the gradient into the layer below, which goes to the other of the two gradient tiles.
        wgrad_tile(act[5], kSH, g6, kSG, wv, 0, lane, g5);
        if (wv == 0) bias_tile(g6, kSG, 0, lane, first, db[5]);
        backward_tile<1>(g6, kSG, act[5], gb[0], T.b5, lane, wv);
        __syncthreads();
#pragma unroll
        for (int l = 4; l >= 1; --l) {
            const bf16 *gl = gb[(4 - l) & 1];
#pragma unroll
            for (int k = 0; k < 4; ++k) wgrad_tile(act[l], kSH, gl, kSH, k, wv, lane, gh[l - 1][k]);
            bias_tile(gl, kSH, wv, lane, first, db[l]);
            backward_tile<2>(gl, kSH, act[l], gb[(5 - l) & 1], T.bh[l - 1], lane, wv);
            __syncthreads();
        }
        // gb[0]: the gradient at layer 0's output
#pragma unroll
        for (int k = 0; k < 6; ++k) wgrad_tile(act0, kS0, gb[0], kSH, k, wv, lane, g0[k]);
        bias_tile(gb[0], kSH, wv, lane, first, db[0]);
        __syncthreads();
        first = false;
    }
    // The workgroup's partial (zero for a workgroup without a chunk): each gradient tile as the wave holds
    // it, one 16-byte store per lane (the partial's tile order, part_index).
    {
        float *pw = A.part + (size_t)blockIdx.x * kPP;
        const int col = lane & 15, g = lane >> 4;
#pragma unroll
        for (int k = 0; k < 6; ++k) *reinterpret_cast<f32x4 *>(pw + w_off(0) + ((k * 4 + wv) * 64 + lane) * 4) = g0[k];
#pragma unroll
        for (int l = 1; l < 5; ++l)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                *reinterpret_cast<f32x4 *>(pw + w_off(l) + ((k * 4 + wv) * 64 + lane) * 4) = gh[l - 1][k];
        if (col < 3) *reinterpret_cast<f32x4 *>(pw + w_off(5) + ((wv * 4 + g) * 3 + col) * 4) = g5;
        if (g == 0) {
#pragma unroll
            for (int l = 0; l < 5; ++l) pw[b_off(l) + wv * 16 + col] = db[l];
            if (wv == 0 && col < 3) pw[b_off(5) + col] = db[5];
        }
    }
    if (wv == 0) {
        loss = wave_sum(loss);
        nvalid = wave_sum(nvalid);
        if (lane == 0) {
            A.loss_part[blockIdx.x] = loss * A.inv_n;
            A.valid_part[blockIdx.x] = (int32_t)nvalid;
        }
    }
}

// First level of the deterministic gradient reduction: slice y of the partial rows, four parameters
// per thread (float4), summed in row order into part2[y].
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *part, int32_t nparts, float *part2) {
    const int p4 = blockIdx.x * 256 + threadIdx.x;
    if (p4 * 4 >= kPP) return;
    const int r0 = (int)((long long)nparts * blockIdx.y / kSlices), r1 = (int)((long long)nparts * (blockIdx.y + 1) / kSlices);
    const float4 *src = reinterpret_cast<const float4 *>(part) + p4;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    // rows in order, eight loads in flight at a time
    for (int r = r0; r < r1; r += 8) {
        float4 v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = r + j < r1 ? src[(size_t)(r + j) * (kPP / 4)] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (r + j < r1) {
                s.x += v[j].x; s.y += v[j].y; s.z += v[j].z; s.w += v[j].w;
            }
    }
    reinterpret_cast<float4 *>(part2 + (size_t)blockIdx.y * kPP)[p4] = s;
}

struct AdamArgs {
    float *w, *m, *v;
    float *we;             // the inference weights: an exponential mean of w (w itself with ema 0)
    float ema;
    const float *part;     // [kSlices][kPP]: the first-level sums (mode 1, 2 without grad_in)
    int32_t nparts;
    const float *grad_in;  // mode 2: the gradient instead of the partials
    float *grad_out;       // mode 1: the reduced gradient
    int32_t mode;          // 0: fragments only; 1: reduce only; 2: reduce (or grad_in) + Adam + fragments
    float lr, b1, b2, c1, c2, bc1, bc2, eps;
    bf16 *ffrag, *bfrag, *ifrag;
};

// wv: the training weight (forward and backward fragments); wi: the inference weight.
__device__ __forceinline__ void write_frags(const AdamArgs &A, int p, float wv, float wi) {
    int l = 0;
    while (l < 5 && p >= w_off(l + 1)) ++l;
    if (p >= b_off(l)) return;   // a bias: no fragment
    const int q = p - w_off(l), nout = out_dim(l);
    const int in = q / nout, out = q % nout;
    const bf16 bv = (bf16)wv, bi = (bf16)wi;
    {
        const int mt = out / 16, kb = in / 32, ln = (out & 15) + 16 * ((in & 31) >> 3), i = in & 7;
        A.ffrag[((size_t)(fwd_base(l) + mt * fwd_kb(l) + kb) * 64 + ln) * 8 + i] = bv;
        if (l == 0) {
            A.ifrag[((size_t)(fwd_base(l) + mt * fwd_kb(l) + kb) * 64 + ln) * 8 + i] = bi;
        } else {   // in = perm_in(kb, g, j)
            const int t = in >> 4, r = in & 15;
            const int kb2 = t >> 1, g2 = r >> 2, j = ((t & 1) << 2) | (r & 3);
            A.ifrag[((size_t)(fwd_base(l) + mt * 2 + kb2) * 64 + (out & 15) + 16 * g2) * 8 + j] = bi;
        }
    }
    if (l >= 1) {
        const int mt = in / 16, kb = out / 32, ln = (in & 15) + 16 * ((out & 31) >> 3), i = out & 7;
        A.bfrag[((size_t)(bwd_base(l) + mt * bwd_kb(l) + kb) * 64 + ln) * 8 + i] = bv;
    }
}

// One parameter per thread: the gradient (the first-level sums added in slice order, or grad_in), then
// Adam on the FP32 master weight, the inference weight we = ema * we + (1 - ema) * w, and the BF16 fragment
// copies (training fragments from w, inference fragments from we).  Mode 0 sets we = w.
__global__ void __launch_bounds__(256) k_mlp_adam(AdamArgs A) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kP) return;
    if (A.mode == 0) {
        const float w0 = A.w[p];
        A.we[p] = w0;
        write_frags(A, p, w0, w0);
        return;
    }
    float g;
    if (A.grad_in) {
        g = A.grad_in[p];
    } else {
        g = 0.0f;
        const int pi = part_index(p);
        float v[kSlices];
#pragma unroll
        for (int q = 0; q < kSlices; ++q) v[q] = A.part[(size_t)q * kPP + pi];
#pragma unroll
        for (int q = 0; q < kSlices; ++q) g += v[q];
    }
    if (A.mode == 1) {
        A.grad_out[p] = g;
        return;
    }
    const float m = A.b1 * A.m[p] + A.c1 * g;
    const float v = A.b2 * A.v[p] + A.c2 * (g * g);
    const float mh = m / A.bc1, vh = v / A.bc2;
    const float wv = A.w[p] - A.lr * mh / (sqrtf(vh) + A.eps);
    A.m[p] = m;
    A.v[p] = v;
    A.w[p] = wv;
    const float wi = A.ema > 0.0f ? A.ema * A.we[p] + (1.0f - A.ema) * wv : wv;
    A.we[p] = wi;
    write_frags(A, p, wv, wi);
}

// Device statistics of the cache (the fallback decision stays on the device, so a frame needs no sync).
struct DevStats {
    long long frames;
    float last_loss, running_error, baseline;
    int32_t fallback, fallback_frames, n_valid;
};

struct StatsArgs {
    const float *loss_part;
    const int32_t *valid_part;
    int32_t nparts;
    DevStats *st;
    unsigned *qcount;
    float *loss_out;        // grad mode: the loss only
    int32_t warmup;
    float ratio;
};

__global__ void __launch_bounds__(64) k_gc_stats(StatsArgs A) {
    const int lane = threadIdx.x;
    float s = 0.0f, nv = 0.0f;
    for (int q = lane; q < A.nparts; q += 64) {
        s += A.loss_part[q];
        nv += (float)A.valid_part[q];
    }
    s = wave_sum(s);
    nv = wave_sum(nv);
    if (lane != 0) return;
    if (A.loss_out) {
        *A.loss_out = s;
        return;
    }
    DevStats &st = *A.st;
    st.frames += 1;
    st.last_loss = s;
    st.n_valid = (int32_t)nv;
    st.running_error = st.frames == 1 ? s : 0.9f * st.running_error + 0.1f * s;
    if (st.frames <= A.warmup) {
        st.fallback = 1;
        st.baseline = INFINITY;
    } else {
        st.baseline = fminf(st.baseline, st.running_error);
        st.fallback = st.running_error > A.ratio * st.baseline ? 1 : 0;
        st.fallback_frames = st.fallback ? st.fallback_frames + 1 : 0;
        if (st.fallback_frames >= 64) {   // a lasting change (new lighting, edits): accept its error level
            st.baseline = st.running_error;
            st.fallback = 0;
            st.fallback_frames = 0;
        }
    }
    *A.qcount = 0u;
}

// ---- GI kernels -------------------------------------------------------------------------------

struct PathArgs {
    int32_t w, h, tiles_x, ntiles, spp, terminate, bounces;
    uint32_t frame, salt;
    GiPathCfg G;
    float cam[3];
};

__device__ __forceinline__ void gbuffer_point(const float4 *gbpos, const uint32_t *gbn, size_t i, float P[3],
                                              float N[3]) {
    const float4 p4 = gbpos[i];
    const uint32_t nb = gbn[i];
    P[0] = p4.x; P[1] = p4.y; P[2] = p4.z;
    N[0] = unorm8_dec(nb & 255u);
    N[1] = unorm8_dec((nb >> 8) & 255u);
    N[2] = unorm8_dec((nb >> 16) & 255u);
}

__device__ __forceinline__ bool tile_pixel(const PathArgs &A, int32_t &px, int32_t &py) {
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t tx = L % A.tiles_x, ty = L / A.tiles_x;
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    return px < A.w && py < A.h;
}

// Sample k of pixel (px, py): bounce-1 direction cosineHemisphereU(N, ldSample2(pix, k, salt)) and the
// path seed pixelSeed(pix, k, salt).
__device__ __forceinline__ void sample_start(const float N[3], int32_t px, int32_t py, uint32_t k, uint32_t salt,
                                             float dir[3], uint32_t &seed) {
    float u0, u1;
    ld_sample2(px, py, k, salt, u0, u1);
    cosine_hemisphere_u(N, u0, u1, dir);
    seed = pixel_seed(px, py, k, salt);
}

// Plain traced GI: the mean of spp full paths per pixel, two bounces (the reference's giPath) or, with B3,
// three (gi_path3).
template <bool BODIES, bool B3>
__device__ __forceinline__ void plain_pixel(const DevWorld &W, const SkyParams &S, const DevBodies &B, const PathArgs &A,
                                            const float4 *gbpos, const uint32_t *gbn, int32_t px, int32_t py,
                                            float4 *out) {
    const size_t i = (size_t)py * A.w + px;
    float sum[3] = {0.0f, 0.0f, 0.0f};
    const bool sky = gbpos[i].w < 0.0f;
    if (!sky) {
        float P[3], N[3];
        gbuffer_point(gbpos, gbn, i, P, N);
        Cnt c = {0, 0, 0, 0};
        for (int32_t s = 0; s < A.spp; ++s) {
            float dir[3], L[3];
            uint32_t seed;
            sample_start(N, px, py, A.frame * (uint32_t)A.spp + (uint32_t)s, A.salt, dir, seed);
            if (B3) gi_path3<false, BODIES, true>(W, S, A.G, P, N, dir, seed, L, c, &B);
            else gi_path<3, false, false, BODIES, true>(W, S, A.G, P, N, dir, seed, true, L, nullptr, nullptr, c, &B);
#pragma unroll
            for (int a = 0; a < 3; ++a) sum[a] += L[a];
        }
    }
    const float fs = (float)A.spp;
    out[i] = sky ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : make_float4(sum[0] / fs, sum[1] / fs, sum[2] / fs, 1.0f);
}

template <bool BODIES, bool B3>
__global__ void __launch_bounds__(256) k_gi_plain(DevWorld W, SkyParams S, DevBodies B, PathArgs A, const float4 *gbpos,
                                                  const uint32_t *gbn, float4 *out) {
    int32_t px, py;
    if (!tile_pixel(A, px, py)) return;
    plain_pixel<BODIES, B3>(W, S, B, A, gbpos, gbn, px, py, out);
}

__device__ __forceinline__ uint32_t lane_rank(unsigned long long m) {
    return (uint32_t)__popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull));
}

// Appends the lane's record (if `want`) to the query list with one atomic per wavefront.
__device__ __forceinline__ void push_query(bool want, unsigned *qcount, float4 *queries, const float4 r[4]) {
    const unsigned long long m = __ballot(want);
    if (m == 0ull) return;
    unsigned base = 0;
    if ((threadIdx.x & 63) == (unsigned)(__ffsll((long long)m) - 1)) base = atomicAdd(qcount, (unsigned)__popcll(m));
    base = __shfl(base, __ffsll((long long)m) - 1);
    if (want) {
        float4 *q = queries + (size_t)(base + lane_rank(m)) * 4;
        q[0] = r[0]; q[1] = r[1]; q[2] = r[2]; q[3] = r[3];
    }
}

// Cached render, traced part.  Fallback frames: the plain path, written to out.  Otherwise terminate 1:
// per sample the path up to its tail (pre += its light; a query at the tail, or tailv[slot] = 0);
// terminate 0: nothing (k_mlp_infer<2> answers at the G-buffer points).  B3 (three-bounce GI): a path
// whose bounce-1 hit is a body is traced on to its end (gi_path3's continuation) instead of querying.
template <bool BODIES, bool B3>
__global__ void __launch_bounds__(256, 3) k_gc_render(DevWorld W, SkyParams S, DevBodies B, PathArgs A, const float4 *gbpos,
                                                   const uint32_t *gbn, const DevStats *st, float4 *out, float4 *pre,
                                                   float4 *tailv, float4 *queries, unsigned *qcount) {
    int32_t px, py;
    const bool inside = tile_pixel(A, px, py);
    if (st->fallback) {
        if (inside) plain_pixel<BODIES, B3>(W, S, B, A, gbpos, gbn, px, py, out);
        return;
    }
    if (A.terminate == 0) return;   // answered at the G-buffer point by k_mlp_infer<2>
    const size_t i = inside ? (size_t)py * A.w + px : 0;
    const bool sky = !inside || gbpos[i].w < 0.0f;
    float P[3] = {0.0f, 0.0f, 0.0f}, N[3] = {0.0f, 0.0f, 0.0f};
    if (!sky) gbuffer_point(gbpos, gbn, i, P, N);
    float sum[3] = {0.0f, 0.0f, 0.0f};
    Cnt c = {0, 0, 0, 0};
    for (int32_t s = 0; s < A.spp; ++s) {
        const size_t slot = i * (size_t)A.spp + (size_t)s;
        GiTail t;
        t.valid = false;
        if (!sky) {
            float dir[3], L[3];
            uint32_t seed;
            sample_start(N, px, py, A.frame * (uint32_t)A.spp + (uint32_t)s, A.salt, dir, seed);
            gi_path<3, true, false, BODIES, true, B3>(W, S, A.G, P, N, dir, seed, true, L, nullptr, &t, c, &B);
            if (B3 && t.valid && t.body) {   // a moving surface is never answered by the cache
                float d2[3], L2[3];
                cosine_hemisphere(t.n, seed, d2);
                gi_path<3, false, false, BODIES, true>(W, S, A.G, t.p, t.n, d2, seed, false, L2, nullptr, nullptr, c, &B);
#pragma unroll
                for (int a = 0; a < 3; ++a) L[a] += t.thr[a] * L2[a];
                t.valid = false;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) sum[a] += L[a];
        }
        float4 r[4];
        if (t.valid) {
            r[0] = make_float4(t.p[0], t.p[1], t.p[2], __uint_as_float((uint32_t)slot));
            r[1] = make_float4(t.n[0], t.n[1], t.n[2], 0.0f);
            r[2] = make_float4(-t.dir[0], -t.dir[1], -t.dir[2], 0.0f);
            r[3] = make_float4(t.thr[0], t.thr[1], t.thr[2], 0.0f);
        }
        push_query(t.valid, qcount, queries, r);
        if (inside && !t.valid) tailv[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (inside) pre[i] = make_float4(sum[0], sum[1], sum[2], sky ? 0.0f : 1.0f);
}

// out = (pre + sum of the slots' cache terms) / spp; nothing on fallback frames (out is already final).
__global__ void __launch_bounds__(256) k_gc_resolve(const DevStats *st, int32_t npix, int32_t spp, const float4 *pre,
                                                    const float4 *tailv, float4 *out) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= npix || st->fallback) return;
    const float4 p = pre[i];
    float s[3] = {p.x, p.y, p.z};
    for (int32_t k = 0; k < spp; ++k) {
        const float4 t = tailv[(size_t)i * spp + k];
        s[0] += t.x; s[1] += t.y; s[2] += t.z;
    }
    const float fs = (float)spp;
    out[i] = p.w == 0.0f ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : make_float4(s[0] / fs, s[1] / fs, s[2] / fs, 1.0f);
}

struct SampleArgs {
    int32_t w, h, n, terminate, bounces;
    uint32_t frame, salt;
    GiPathCfg G;
    float cam[3];
};

// Training samples: sample i takes pixel floor((i + r) * W*H / n) (r from its own stream
// pcg(i + pcg(frame ^ salt))) and a cosine-hemisphere bounce-1 direction from the same stream.  The target
// is the traced radiance at the query point: terminate 0, the whole path from the G-buffer point (two
// bounces, or three with B3); terminate 1, the rest of the path from the bounce-1 hit (one bounce and its
// sun ray, or two with B3).
template <bool BODIES, bool B3>
__global__ void __launch_bounds__(256) k_gc_samples(DevWorld W, SkyParams S, DevBodies B, SampleArgs A, const float4 *gbpos,
                                                    const uint32_t *gbn, float4 *out) {
    const int32_t i = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    if (i >= A.n) return;
    float4 *r = out + (size_t)i * 4;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const long long npix = (long long)A.w * A.h;
    const float stratum = (float)npix / (float)A.n;
    uint32_t seed = pcg((uint32_t)i + pcg(A.frame ^ A.salt));
    const float u = rnd(seed);
    long long pix = (long long)floorf(((float)i + u) * stratum);
    if (pix > npix - 1) pix = npix - 1;
    if (gbpos[pix].w < 0.0f) {
        r[0] = z; r[1] = z; r[2] = z; r[3] = z;
        return;
    }
    float P[3], N[3], dir[3], L[3];
    gbuffer_point(gbpos, gbn, (size_t)pix, P, N);
    cosine_hemisphere(N, seed, dir);
    Cnt c = {0, 0, 0, 0};
    if (A.terminate == 0) {
        if (B3) gi_path3<false, BODIES, true>(W, S, A.G, P, N, dir, seed, L, c, &B);
        else gi_path<3, false, false, BODIES, true>(W, S, A.G, P, N, dir, seed, true, L, nullptr, nullptr, c, &B);
        float v[3] = {A.cam[0] - P[0], A.cam[1] - P[1], A.cam[2] - P[2]};
        normalize3(v);
        r[0] = make_float4(P[0], P[1], P[2], __uint_as_float(1u));
        r[1] = make_float4(N[0], N[1], N[2], 0.0f);
        r[2] = make_float4(v[0], v[1], v[2], 0.0f);
        r[3] = make_float4(L[0], L[1], L[2], 0.0f);
        return;
    }
    GiTail t;
    gi_path<3, true, false, BODIES, true>(W, S, A.G, P, N, dir, seed, true, L, nullptr, &t, c, &B);
    if (!t.valid) {
        r[0] = z; r[1] = z; r[2] = z; r[3] = z;
        return;
    }
    float d2[3], L2[3];
    cosine_hemisphere(t.n, seed, d2);
    if (B3) gi_path<3, false, false, BODIES, true>(W, S, A.G, t.p, t.n, d2, seed, false, L2, nullptr, nullptr, c, &B);
    else gi_path<1, false, false, BODIES, true>(W, S, A.G, t.p, t.n, d2, seed, false, L2, nullptr, nullptr, c, &B);
    r[0] = make_float4(t.p[0], t.p[1], t.p[2], __uint_as_float(1u));
    r[1] = make_float4(t.n[0], t.n[1], t.n[2], 0.0f);
    r[2] = make_float4(-t.dir[0], -t.dir[1], -t.dir[2], 0.0f);
    r[3] = make_float4(L2[0], L2[1], L2[2], 0.0f);
}

struct InferArgs {
    const float4 *recs;
    long long n;             // records (mode 0), pixels (mode 2)
    const unsigned *count;   // mode 1: the query count on the device; mode 2: counts the answered pixels
    int32_t shift;
    const bf16x8 *ifrag;     // forward fragments in the inference K order
    const float *w;
    uint16_t *out_h;         // mode 0: FP16 RGB(0) per record
    float4 *tailv;           // mode 1: slot -> throughput * max(FP16 output, 0)
    const float4 *gbpos;     // mode 2: the G-buffer
    const uint32_t *gbn;
    float cam[3];
    float4 *out;             // mode 2: RGBA per pixel
    const DevStats *st;      // mode 2: nothing on fallback frames
};

// 64 records per wavefront batch, 4 wavefronts per workgroup looping over batches; the workgroup
// stages the weight fragments and biases in LDS once.  Activations never leave registers: each
// layer's MFMA output tile is converted (bias, ReLU, BF16) into the next layer's B operand.
// MODE 0: n records -> FP16 outputs; MODE 1: the render's query list -> its slots; MODE 2: the cache's
// answer at every non-sky G-buffer pixel (query built from the G-buffer: position, quantised normal,
// direction to the camera) straight into the output (batches of sky only are skipped).
template <int MODE>
__global__ void __launch_bounds__(256, 2) k_mlp_infer(InferArgs A) {
    __shared__ bf16x8 wl[kFwdFrags * 64];
    __shared__ float bl[kBiasLds];
    __shared__ unsigned answered[4];
    if (MODE == 2 && A.st->fallback) return;
    for (int i = threadIdx.x; i < kFwdFrags * 64; i += 256) wl[i] = A.ifrag[i];
    for (int i = threadIdx.x; i < kBiasLds; i += 256) {
        const int l = i >> 6, o = i & 63;
        bl[i] = (l < 5) ? A.w[b_off(l) + o] : (o < 3 ? A.w[b_off(5) + o] : 0.0f);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, g = lane >> 4;
    const long long n = MODE == 1 ? (long long)*A.count : A.n;
    const long long stride = (long long)gridDim.x * 256;
    unsigned n_answered = 0;
    for (long long base = ((long long)blockIdx.x * 4 + wave) * 64; base < n; base += stride) {
        if (MODE == 2) {   // a batch of sky pixels: zeros, no network
            const long long s = base + lane;
            const bool ground = s < n && A.gbpos[s].w >= 0.0f;
            if (__ballot(ground) == 0ull) {
                if (s < n) A.out[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                continue;
            }
        }
        // layer 0's B operand: group 4 kb + g of the features of sample nt*16 + col
        bf16x8 b0[4][3];
        float4 keep0[4], keep3[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const long long s = base + nt * 16 + col;
            float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0, r2 = r0, r3 = r0;
            if (MODE == 2) {
                if (s < n) {
                    const float4 p4 = A.gbpos[s];
                    if (p4.w >= 0.0f) {
                        const uint32_t nb = A.gbn[s];
                        float v[3] = {A.cam[0] - p4.x, A.cam[1] - p4.y, A.cam[2] - p4.z};
                        normalize3(v);
                        r0 = make_float4(p4.x, p4.y, p4.z, __uint_as_float(1u));
                        r1 = make_float4(unorm8_dec(nb & 255u), unorm8_dec((nb >> 8) & 255u),
                                         unorm8_dec((nb >> 16) & 255u), 0.0f);
                        r2 = make_float4(v[0], v[1], v[2], 0.0f);
                    }
                }
            } else if (s < n) {
                const float4 *rec = A.recs + s * 4;
                r0 = rec[0]; r1 = rec[1]; r2 = rec[2];
                if (MODE == 1) r3 = rec[3];
            }
            keep0[nt] = r0;
            keep3[nt] = r3;
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) {
                float f[8];
                encode_group(r0, r1, r2, A.shift, 4 * kb + g, f);
                b0[nt][kb] = to_bf16x8(f);
            }
        }
        f32x4 acc[4][4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) acc[nt][mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) {
                const bf16x8 a = wl[(fwd_base(0) + mt * 3 + kb) * 64 + lane];
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) acc[nt][mt] = mfma(a, b0[nt][kb], acc[nt][mt]);
            }
        bf16x8 b[4][2];
#pragma unroll 1
        for (int l = 1; l < 6; ++l) {
            // epilogue of layer l - 1 into layer l's B operand
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int mt = 2 * kb + (j >> 2), i = j & 3;
                        const float v = acc[nt][mt][i] + bl[(l - 1) * 64 + mt * 16 + 4 * g + i];
                        b[nt][kb][j] = (bf16)(v > 0.0f ? v : 0.0f);
                    }
            const int nmt = l == 5 ? 1 : 4;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) acc[nt][mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                if (mt < nmt) {
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb) {
                        const bf16x8 a = wl[(fwd_base(l) + mt * 2 + kb) * 64 + lane];
#pragma unroll
                        for (int nt = 0; nt < 4; ++nt) acc[nt][mt] = mfma(a, b[nt][kb], acc[nt][mt]);
                    }
                }
            }
        }
        // outputs: lanes 0..15 hold rows 0..2 (R, G, B) of samples nt*16 + col
        if (g == 0) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const long long s = base + nt * 16 + col;
                if (s >= n) continue;
                const __half hr = __float2half_rn(acc[nt][0][0] + bl[5 * 64 + 0]);
                const __half hg = __float2half_rn(acc[nt][0][1] + bl[5 * 64 + 1]);
                const __half hb = __float2half_rn(acc[nt][0][2] + bl[5 * 64 + 2]);
                if (MODE == 0) {
                    ushort4 o;
                    o.x = __half_as_ushort(hr);
                    o.y = __half_as_ushort(hg);
                    o.z = __half_as_ushort(hb);
                    o.w = 0;
                    reinterpret_cast<ushort4 *>(A.out_h)[s] = o;
                } else if (MODE == 2) {
                    if (__float_as_uint(keep0[nt].w) != 0u) {
                        n_answered += 1;
                        A.out[s] = make_float4(fmaxf(__half2float(hr), 0.0f), fmaxf(__half2float(hg), 0.0f),
                                               fmaxf(__half2float(hb), 0.0f), 1.0f);
                    } else {
                        A.out[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    }
                } else {
                    const uint32_t slot = __float_as_uint(keep0[nt].w);
                    const float4 t = keep3[nt];
                    A.tailv[slot] = make_float4(t.x * fmaxf(__half2float(hr), 0.0f), t.y * fmaxf(__half2float(hg), 0.0f),
                                                t.z * fmaxf(__half2float(hb), 0.0f), 0.0f);
                }
            }
        }
    }
    if (MODE == 2) {   // answered pixels: one atomic per workgroup
        float c = (float)n_answered;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (lane == 0) answered[wave] = (unsigned)c;
        __syncthreads();
        if (threadIdx.x == 0) atomicAdd(const_cast<unsigned *>(A.count), ((answered[0] + answered[1]) + answered[2]) + answered[3]);
    }
}

__global__ void __launch_bounds__(64) k_mlp_encode(const float4 *recs, long long n, int32_t shift, float *enc) {
    const long long i = (long long)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const float4 *r = recs + i * 4;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2];
    for (int q = 0; q < 12; ++q) {
        float f[8];
        encode_group(r0, r1, r2, shift, q, f);
        float4 *o = reinterpret_cast<float4 *>(enc + i * kIn + 8 * q);
        o[0] = make_float4(f[0], f[1], f[2], f[3]);
        o[1] = make_float4(f[4], f[5], f[6], f[7]);
    }
}

uint32_t host_pcg(uint32_t v) {
    v = v * 747796405u + 2891336453u;
    v = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
    return (v >> 22u) ^ v;
}

}  // namespace

struct vf_gicache {
    vf_ctx *ctx = nullptr;
    vf_gicache_params p{};
    float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *we = nullptr;
    bf16 *ffrag = nullptr, *bfrag = nullptr, *ifrag = nullptr;
    float *part = nullptr, *part2 = nullptr, *loss_part = nullptr;
    int32_t *valid_part = nullptr;
    int32_t part_rows = 0;
    float4 *samples = nullptr;
    DevStats *st = nullptr;
    unsigned *qcount = nullptr;
    float4 *queries = nullptr, *pre = nullptr, *tailv = nullptr;
    int64_t cap_pix = 0, cap_slots = 0;
    long long steps = 0;
};

namespace {

int gc_fail_null(vf_ctx *ctx, const char *what) { return vf::fail(ctx, VF_E_INVALID, "%s: null argument", what); }

int check_cache(vf_ctx *ctx, vf_gicache *c, const char *what) {
    if (!ctx) return VF_E_INVALID;
    if (!c || c->ctx != ctx) return vf::fail(ctx, VF_E_INVALID, "%s: cache missing or created on another context", what);
    return vf::set_device(ctx);
}

void free_all(vf_gicache *c) {
    void *ptrs[] = {c->w, c->m, c->v, c->grad, c->we, c->ffrag, c->bfrag, c->ifrag, c->part, c->part2, c->loss_part, c->valid_part,
                    c->samples, c->st, c->qcount, c->queries, c->pre, c->tailv};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
}

int launch_adam(vf_gicache *c, int mode, const float *grad_in, float *grad_out, int nparts) {
    if (mode != 0 && !grad_in) {
        hipLaunchKernelGGL(k_mlp_reduce, dim3((kPP / 4 + 255) / 256, kSlices), dim3(256), 0, c->ctx->stream, c->part,
                           nparts, c->part2);
        int rc = launch_check(c->ctx, "k_mlp_reduce");
        if (rc) return rc;
    }
    AdamArgs A;
    A.w = c->w; A.m = c->m; A.v = c->v;
    A.we = c->we;
    A.ema = c->p.weight_ema;
    A.part = c->part2;
    A.nparts = nparts;
    A.grad_in = grad_in;
    A.grad_out = grad_out;
    A.mode = mode;
    A.lr = c->p.lr; A.b1 = c->p.beta1; A.b2 = c->p.beta2;
    A.c1 = 1.0f - c->p.beta1; A.c2 = 1.0f - c->p.beta2;
    A.bc1 = 1.0f; A.bc2 = 1.0f;
    A.eps = c->p.eps;
    if (mode == 2) {
        c->steps += 1;
        A.bc1 = (float)(1.0 - std::pow((double)c->p.beta1, (double)c->steps));
        A.bc2 = (float)(1.0 - std::pow((double)c->p.beta2, (double)c->steps));
    }
    A.ffrag = c->ffrag; A.bfrag = c->bfrag; A.ifrag = c->ifrag;
    hipLaunchKernelGGL(k_mlp_adam, dim3((kP + 255) / 256), dim3(256), 0, c->ctx->stream, A);
    return launch_check(c->ctx, "k_mlp_adam");
}

// Forward + backward over n samples into the partials; returns the number of partial rows used.
int launch_train(vf_gicache *c, const float4 *samples, long long n, int *rows) {
    TrainArgs A;
    A.samples = samples;
    A.n = n;
    A.n_chunks = (int32_t)((n + 63) / 64);
    A.shift = c->p.enc_shift;
    A.inv_n = 1.0f / (float)n;
    A.loss_eps = c->p.loss_eps;
    A.ffrag = reinterpret_cast<const bf16x8 *>(c->ffrag);
    A.bfrag = reinterpret_cast<const bf16x8 *>(c->bfrag);
    A.w = c->w;
    A.part = c->part;
    A.loss_part = c->loss_part;
    A.valid_part = c->valid_part;
    const int g = A.n_chunks < c->part_rows ? A.n_chunks : c->part_rows;
    hipLaunchKernelGGL(k_mlp_train, dim3(g), dim3(64 * kTrainWaves), 0, c->ctx->stream, A);
    *rows = g;
    return launch_check(c->ctx, "k_mlp_train");
}

// Persistent inference grid: two workgroups per compute unit, fewer for small batches.
unsigned infer_groups(const vf_gicache *c, int64_t n) {
    const int64_t need = (n + 255) / 256, cap = 2ll * c->ctx->n_cu;
    return (unsigned)(need < 1 ? 1 : (need < cap ? need : cap));
}

int launch_stats(vf_gicache *c, int rows, float *loss_out) {
    StatsArgs S;
    S.loss_part = c->loss_part;
    S.valid_part = c->valid_part;
    S.nparts = rows;
    S.st = c->st;
    S.qcount = c->qcount;
    S.loss_out = loss_out;
    S.warmup = c->p.warmup_frames;
    S.ratio = c->p.fallback_ratio;
    hipLaunchKernelGGL(k_gc_stats, dim3(1), dim3(64), 0, c->ctx->stream, S);
    return launch_check(c->ctx, "k_gc_stats");
}

void init_weights(const vf_gicache_params &p, std::vector<float> &w) {
    w.assign(kP, 0.0f);
    uint32_t st = host_pcg(p.seed ^ 0x5eedu);
    for (int l = 0; l < 6; ++l) {
        const int fan_in = l == 0 ? 78 : 64;
        float bound = std::sqrt(6.0f / (float)fan_in);
        if (l == 5) bound *= 0.25f;
        for (int i = 0; i < in_dim(l); ++i)
            for (int o = 0; o < out_dim(l); ++o) {
                st = host_pcg(st);
                const float u = (float)(st >> 8u) * (1.0f / 16777216.0f);
                w[w_off(l) + i * out_dim(l) + o] = (l == 0 && i >= 78) ? 0.0f : (2.0f * u - 1.0f) * bound;
            }
    }
}

int upload_params(vf_ctx *ctx, vf_gicache *c, const float *host) {
    VF_HIP(ctx, hipMemcpyAsync(c->w, host, sizeof(float) * kP, hipMemcpyHostToDevice, ctx->stream));
    VF_HIP(ctx, hipMemsetAsync(c->m, 0, sizeof(float) * kP, ctx->stream));
    VF_HIP(ctx, hipMemsetAsync(c->v, 0, sizeof(float) * kP, ctx->stream));
    c->steps = 0;
    int rc = launch_adam(c, 0, nullptr, nullptr, 0);
    if (rc) return rc;
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}

int reset_stats(vf_ctx *ctx, vf_gicache *c) {
    DevStats s{};
    s.baseline = INFINITY;
    s.fallback = 1;
    VF_HIP(ctx, hipMemcpyAsync(c->st, &s, sizeof s, hipMemcpyHostToDevice, ctx->stream));
    VF_HIP(ctx, hipMemsetAsync(c->qcount, 0, sizeof(unsigned), ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}

int check_params(vf_ctx *ctx, const vf_gicache_params *p) {
    if (p->terminate != 0 && p->terminate != 1) return vf::fail(ctx, VF_E_INVALID, "vf_gicache: terminate must be 0 or 1");
    if (p->train_steps < 1 || p->train_steps > 64) return vf::fail(ctx, VF_E_INVALID, "vf_gicache: train_steps must be in [1, 64]");
    if (p->train_samples < 64 || p->train_samples > (1 << 22) || p->train_samples % (64 * p->train_steps))
        return vf::fail(ctx, VF_E_INVALID, "vf_gicache: train_samples %d must be a multiple of 64 x train_steps in [64, 2^22]",
                        p->train_samples);
    if (p->spp < 1 || p->spp > 8) return vf::fail(ctx, VF_E_INVALID, "vf_gicache: spp must be in [1, 8]");
    if (!(p->lr > 0.0f) || !(p->beta1 >= 0.0f && p->beta1 < 1.0f) || !(p->beta2 >= 0.0f && p->beta2 < 1.0f) ||
        !(p->eps > 0.0f) || !(p->loss_eps > 0.0f) || !(p->fallback_ratio >= 1.0f))
        return vf::fail(ctx, VF_E_INVALID, "vf_gicache: optimizer or loss parameter out of range");
    if (p->warmup_frames < 0 || p->enc_shift < 0 || p->enc_shift > 24)
        return vf::fail(ctx, VF_E_INVALID, "vf_gicache: warmup_frames or enc_shift out of range");
    if (!(p->gi_reach > 0.0f) || p->gi_reach > 1e6f) return vf::fail(ctx, VF_E_INVALID, "vf_gicache: gi_reach out of range");
    if (p->bounces != 2 && p->bounces != 3) return vf::fail(ctx, VF_E_INVALID, "vf_gicache: bounces must be 2 or 3");
    if (!(p->weight_ema >= 0.0f && p->weight_ema < 1.0f))
        return vf::fail(ctx, VF_E_INVALID, "vf_gicache: weight_ema must be in [0, 1)");
    return VF_OK;
}

int check_gbuffer(vf_ctx *ctx, const char *what, const vf_pass_ubo *pass, int32_t w, int32_t h, const float *gb_pos,
                  const uint32_t *gb_normal) {
    int rc = vf::check_ready(ctx, what);
    if (rc) return rc;
    if (!ctx->has_scene) return vf::fail(ctx, VF_E_INVALID, "%s: no scene (vf_set_scene)", what);
    if (!pass || !gb_pos || !gb_normal) return gc_fail_null(ctx, what);
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 26)) return vf::fail(ctx, VF_E_INVALID, "%s: bad size %dx%d", what, w, h);
    return vf::check_bodies(ctx, what);
}

PathArgs path_args(const vf_pass_ubo *pass, int32_t w, int32_t h, uint32_t frame, int32_t spp, uint32_t salt, float reach,
                   int32_t nee, int32_t terminate, int32_t bounces) {
    PathArgs A;
    A.w = w; A.h = h;
    A.tiles_x = (w + 15) / 16;
    A.ntiles = A.tiles_x * ((h + 15) / 16);
    A.spp = spp;
    A.terminate = terminate;
    A.bounces = bounces;
    A.frame = frame;
    A.salt = salt;
    A.G.reach = reach;
    A.G.lights_nee = nee ? 1 : 0;
    for (int a = 0; a < 3; ++a) A.cam[a] = pass->cam_pos[a];
    return A;
}

int launch_samples(vf_ctx *ctx, vf_gicache *c, const vf_pass_ubo *pass, int32_t w, int32_t h, const float *gb_pos,
                   const uint32_t *gb_normal, uint32_t frame, float4 *out) {
    SampleArgs A;
    A.w = w; A.h = h;
    A.n = c->p.train_samples;
    A.terminate = c->p.terminate;
    A.bounces = c->p.bounces;
    A.frame = frame;
    A.salt = c->p.salt;
    A.G.reach = c->p.gi_reach;
    A.G.lights_nee = c->p.lights_nee ? 1 : 0;
    for (int a = 0; a < 3; ++a) A.cam[a] = pass->cam_pos[a];
    const DevBodies B = vf::dev_bodies(ctx);
    using Fn = void (*)(DevWorld, SkyParams, DevBodies, SampleArgs, const float4 *, const uint32_t *, float4 *);
    const bool b3 = A.bounces == 3;
    const Fn fn = B.n > 0 ? (b3 ? k_gc_samples<true, true> : k_gc_samples<true, false>)
                          : (b3 ? k_gc_samples<false, true> : k_gc_samples<false, false>);
    hipLaunchKernelGGL(fn, dim3(blocks_for(A.n)), dim3(256), 0, ctx->stream, ctx->W, ctx->S, B, A,
                       reinterpret_cast<const float4 *>(gb_pos), gb_normal, out);
    return launch_check(ctx, "k_gc_samples");
}

}  // namespace

extern "C" {

int vf_gicache_defaults(vf_gicache_params *p) {
    if (!p) return VF_E_INVALID;
    std::memset(p, 0, sizeof *p);
    p->terminate = 0;
    p->train_samples = 65536;
    p->train_steps = 4;
    p->spp = 1;
    p->lr = 1e-2f;
    p->beta1 = 0.9f;
    p->beta2 = 0.9f;
    p->eps = 1e-8f;
    p->loss_eps = 0.01f;
    p->fallback_ratio = 2.0f;
    p->warmup_frames = 16;
    p->enc_shift = 8;
    p->seed = 1337u;
    p->salt = 0x2f6b9a31u;
    p->gi_reach = 48.0f;
    p->lights_nee = 1;
    p->bounces = 2;
    p->weight_ema = 0.9f;
    return VF_OK;
}

int vf_gicache_create(vf_ctx *ctx, const vf_gicache_params *p, vf_gicache **out) {
    if (!ctx) return VF_E_INVALID;
    if (!p || !out) return gc_fail_null(ctx, "vf_gicache_create");
    *out = nullptr;
    int rc = check_params(ctx, p);
    if (rc) return rc;
    if ((rc = vf::no_capture(ctx, "vf_gicache_create"))) return rc;
    if ((rc = vf::set_device(ctx))) return rc;
    vf_gicache *c = new (std::nothrow) vf_gicache();
    if (!c) return vf::fail(ctx, VF_E_OOM, "vf_gicache_create: out of host memory");
    c->ctx = ctx;
    c->p = *p;
    c->part_rows = ctx->n_cu;   // one training workgroup per compute unit (its registers hold the weights)
    hipError_t e = hipSuccess;
#define VF_GC_ALLOC(ptr, bytes)                                              \
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&(ptr)), (bytes))
    VF_GC_ALLOC(c->w, sizeof(float) * kP);
    VF_GC_ALLOC(c->we, sizeof(float) * kP);
    VF_GC_ALLOC(c->m, sizeof(float) * kP);
    VF_GC_ALLOC(c->v, sizeof(float) * kP);
    VF_GC_ALLOC(c->grad, sizeof(float) * kP);
    VF_GC_ALLOC(c->ffrag, sizeof(bf16) * kFwdFrags * 512);
    VF_GC_ALLOC(c->ifrag, sizeof(bf16) * kFwdFrags * 512);
    VF_GC_ALLOC(c->bfrag, sizeof(bf16) * kBwdFrags * 512);
    VF_GC_ALLOC(c->part, sizeof(float) * (size_t)kPP * c->part_rows);
    VF_GC_ALLOC(c->part2, sizeof(float) * (size_t)kPP * kSlices);
    VF_GC_ALLOC(c->loss_part, sizeof(float) * c->part_rows);
    VF_GC_ALLOC(c->valid_part, sizeof(int32_t) * c->part_rows);
    VF_GC_ALLOC(c->samples, sizeof(float4) * 4 * (size_t)p->train_samples);
    VF_GC_ALLOC(c->st, sizeof(DevStats));
    VF_GC_ALLOC(c->qcount, sizeof(unsigned));
#undef VF_GC_ALLOC
    if (e == hipSuccess) e = hipMemsetAsync(c->ffrag, 0, sizeof(bf16) * kFwdFrags * 512, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->ifrag, 0, sizeof(bf16) * kFwdFrags * 512, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->bfrag, 0, sizeof(bf16) * kBwdFrags * 512, ctx->stream);
    if (e != hipSuccess) {
        free_all(c);
        delete c;
        return vf::fail(ctx, e == hipErrorOutOfMemory ? VF_E_OOM : VF_E_DEVICE, "vf_gicache_create: %s", hipGetErrorString(e));
    }
    if ((rc = vf_gicache_reset(ctx, c))) {
        free_all(c);
        delete c;
        return rc;
    }
    *out = c;
    return VF_OK;
}

void vf_gicache_destroy(vf_gicache *c) {
    if (!c) return;
    if (c->ctx) {
        (void)vf::set_device(c->ctx);
        (void)hipStreamSynchronize(c->ctx->stream);
    }
    free_all(c);
    delete c;
}

int vf_gicache_reset(vf_ctx *ctx, vf_gicache *c) {
    int rc = check_cache(ctx, c, "vf_gicache_reset");
    if (rc) return rc;
    if ((rc = vf::no_capture(ctx, "vf_gicache_reset"))) return rc;
    std::vector<float> w;
    init_weights(c->p, w);
    if ((rc = upload_params(ctx, c, w.data()))) return rc;
    return reset_stats(ctx, c);
}

int vf_trace_gi_path_n(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t w, int32_t h, float gi_reach, int32_t lights_nee,
                       const float *gb_pos, const uint32_t *gb_normal, uint32_t frame, int32_t spp, uint32_t salt,
                       int32_t bounces, float *out_rgba) {
    int rc = check_gbuffer(ctx, "vf_trace_gi_path", pass, w, h, gb_pos, gb_normal);
    if (rc) return rc;
    if (!out_rgba) return gc_fail_null(ctx, "vf_trace_gi_path");
    if (spp < 1 || spp > 4096) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi_path: spp %d out of [1, 4096]", spp);
    if (!(gi_reach > 0.0f) || gi_reach > 1e6f) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi_path: gi_reach out of range");
    if (bounces != 2 && bounces != 3) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi_path: bounces must be 2 or 3");
    const PathArgs A = path_args(pass, w, h, frame, spp, salt, gi_reach, lights_nee, 1, bounces);
    if ((rc = vf::begin_timed(ctx))) return rc;
    const DevBodies B = vf::dev_bodies(ctx);
    using Fn = void (*)(DevWorld, SkyParams, DevBodies, PathArgs, const float4 *, const uint32_t *, float4 *);
    const Fn fn = B.n > 0 ? (bounces == 3 ? k_gi_plain<true, true> : k_gi_plain<true, false>)
                          : (bounces == 3 ? k_gi_plain<false, true> : k_gi_plain<false, false>);
    hipLaunchKernelGGL(fn, dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, ctx->S, B, A,
                       reinterpret_cast<const float4 *>(gb_pos), gb_normal, reinterpret_cast<float4 *>(out_rgba));
    if ((rc = launch_check(ctx, "k_gi_plain"))) return rc;
    return vf::end_timed(ctx);
}

int vf_trace_gi_path(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t w, int32_t h, float gi_reach, int32_t lights_nee,
                     const float *gb_pos, const uint32_t *gb_normal, uint32_t frame, int32_t spp, uint32_t salt,
                     float *out_rgba) {
    return vf_trace_gi_path_n(ctx, pass, w, h, gi_reach, lights_nee, gb_pos, gb_normal, frame, spp, salt, 2, out_rgba);
}

int vf_gicache_samples(vf_ctx *ctx, vf_gicache *c, const vf_pass_ubo *pass, int32_t w, int32_t h, const float *gb_pos,
                       const uint32_t *gb_normal, uint32_t frame, vf_gi_record *out) {
    int rc = check_cache(ctx, c, "vf_gicache_samples");
    if (rc) return rc;
    if ((rc = check_gbuffer(ctx, "vf_gicache_samples", pass, w, h, gb_pos, gb_normal))) return rc;
    if (!out) return gc_fail_null(ctx, "vf_gicache_samples");
    return launch_samples(ctx, c, pass, w, h, gb_pos, gb_normal, frame, reinterpret_cast<float4 *>(out));
}

int vf_gicache_frame(vf_ctx *ctx, vf_gicache *c, const vf_pass_ubo *pass, int32_t w, int32_t h, const float *gb_pos,
                     const uint32_t *gb_normal, uint32_t frame, float *out_rgba) {
    int rc = check_cache(ctx, c, "vf_gicache_frame");
    if (rc) return rc;
    if ((rc = check_gbuffer(ctx, "vf_gicache_frame", pass, w, h, gb_pos, gb_normal))) return rc;
    if (!out_rgba) return gc_fail_null(ctx, "vf_gicache_frame");
    const int64_t npix = (int64_t)w * h, nslots = npix * c->p.spp;
    if (c->p.terminate == 1 && (npix > c->cap_pix || nslots > c->cap_slots)) {
        if ((rc = vf::no_capture(ctx, "vf_gicache_frame (first frame of a new size)"))) return rc;
        VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (c->queries) (void)hipFree(c->queries);
        if (c->pre) (void)hipFree(c->pre);
        if (c->tailv) (void)hipFree(c->tailv);
        c->queries = c->pre = c->tailv = nullptr;
        c->cap_pix = c->cap_slots = 0;
        VF_HIP(ctx, hipMalloc(reinterpret_cast<void **>(&c->queries), sizeof(float4) * 4 * (size_t)nslots));
        VF_HIP(ctx, hipMalloc(reinterpret_cast<void **>(&c->pre), sizeof(float4) * (size_t)npix));
        VF_HIP(ctx, hipMalloc(reinterpret_cast<void **>(&c->tailv), sizeof(float4) * (size_t)nslots));
        c->cap_pix = npix;
        c->cap_slots = nslots;
    }
    if ((rc = vf::begin_timed(ctx))) return rc;
    // 1. fresh samples, 2. train_steps Adam steps (the first step's loss feeds the running error)
    if ((rc = launch_samples(ctx, c, pass, w, h, gb_pos, gb_normal, frame, c->samples))) return rc;
    const long long per = c->p.train_samples / c->p.train_steps;
    for (int s = 0; s < c->p.train_steps; ++s) {
        int rows = 0;
        if ((rc = launch_train(c, c->samples + (size_t)s * per * 4, per, &rows))) return rc;
        if (s == 0 && (rc = launch_stats(c, rows, nullptr))) return rc;
        if ((rc = launch_adam(c, 2, nullptr, nullptr, rows))) return rc;
    }
    // 3. render: traced part + queries, the network on the queries, resolve
    const int32_t spp = c->p.spp;
    const PathArgs R =
        path_args(pass, w, h, frame, spp, c->p.salt, c->p.gi_reach, c->p.lights_nee, c->p.terminate, c->p.bounces);
    const DevBodies B = vf::dev_bodies(ctx);
    using Fn = void (*)(DevWorld, SkyParams, DevBodies, PathArgs, const float4 *, const uint32_t *, const DevStats *,
                        float4 *, float4 *, float4 *, float4 *, unsigned *);
    const bool b3 = c->p.bounces == 3;
    const Fn fn = B.n > 0 ? (b3 ? k_gc_render<true, true> : k_gc_render<true, false>)
                          : (b3 ? k_gc_render<false, true> : k_gc_render<false, false>);
    hipLaunchKernelGGL(fn, dim3(R.ntiles), dim3(256), 0, ctx->stream, ctx->W, ctx->S, B, R,
                       reinterpret_cast<const float4 *>(gb_pos), gb_normal, c->st, reinterpret_cast<float4 *>(out_rgba),
                       c->pre, c->tailv, c->queries, c->qcount);
    if ((rc = launch_check(ctx, "k_gc_render"))) return rc;
    InferArgs I;
    std::memset(&I, 0, sizeof I);
    I.count = c->qcount;
    I.shift = c->p.enc_shift;
    I.ifrag = reinterpret_cast<const bf16x8 *>(c->ifrag);
    I.w = c->we;
    if (c->p.terminate == 0) {
        I.n = npix;
        I.gbpos = reinterpret_cast<const float4 *>(gb_pos);
        I.gbn = gb_normal;
        for (int a = 0; a < 3; ++a) I.cam[a] = pass->cam_pos[a];
        I.out = reinterpret_cast<float4 *>(out_rgba);
        I.st = c->st;
        hipLaunchKernelGGL(k_mlp_infer<2>, dim3(infer_groups(c, npix)), dim3(256), 0, ctx->stream, I);
        if ((rc = launch_check(ctx, "k_mlp_infer"))) return rc;
        return vf::end_timed(ctx);
    }
    I.recs = c->queries;
    I.tailv = c->tailv;
    hipLaunchKernelGGL(k_mlp_infer<1>, dim3(infer_groups(c, nslots)), dim3(256), 0, ctx->stream, I);
    if ((rc = launch_check(ctx, "k_mlp_infer"))) return rc;
    hipLaunchKernelGGL(k_gc_resolve, dim3(blocks_for(npix)), dim3(256), 0, ctx->stream, c->st, (int32_t)npix, spp, c->pre,
                       c->tailv, reinterpret_cast<float4 *>(out_rgba));
    if ((rc = launch_check(ctx, "k_gc_resolve"))) return rc;
    return vf::end_timed(ctx);
}

int vf_gicache_read_stats(vf_ctx *ctx, vf_gicache *c, vf_gicache_stats *out) {
    int rc = check_cache(ctx, c, "vf_gicache_read_stats");
    if (rc) return rc;
    if (!out) return gc_fail_null(ctx, "vf_gicache_read_stats");
    if ((rc = vf::no_capture(ctx, "vf_gicache_read_stats"))) return rc;
    DevStats s;
    unsigned q = 0;
    VF_HIP(ctx, hipMemcpyAsync(&s, c->st, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
    VF_HIP(ctx, hipMemcpyAsync(&q, c->qcount, sizeof q, hipMemcpyDeviceToHost, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::memset(out, 0, sizeof *out);
    out->frames = s.frames;
    out->steps = c->steps;
    out->last_loss = s.last_loss;
    out->running_error = s.running_error;
    out->baseline = s.baseline;
    out->fallback = s.fallback;
    out->n_queries = (int32_t)q;
    out->n_valid_samples = s.n_valid;
    return VF_OK;
}

namespace {
int read_params(vf_ctx *ctx, vf_gicache *c, const float *src, float *host_out, const char *what) {
    int rc = check_cache(ctx, c, what);
    if (rc) return rc;
    if (!host_out) return gc_fail_null(ctx, what);
    if ((rc = vf::no_capture(ctx, what))) return rc;
    VF_HIP(ctx, hipMemcpyAsync(host_out, src, sizeof(float) * kP, hipMemcpyDeviceToHost, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}
}  // namespace

int vf_gicache_get_params(vf_ctx *ctx, vf_gicache *c, float *host_out) {
    return read_params(ctx, c, c ? c->w : nullptr, host_out, "vf_gicache_get_params");
}

int vf_gicache_get_render_params(vf_ctx *ctx, vf_gicache *c, float *host_out) {
    return read_params(ctx, c, c ? c->we : nullptr, host_out, "vf_gicache_get_render_params");
}

int vf_gicache_set_params(vf_ctx *ctx, vf_gicache *c, const float *host_in) {
    int rc = check_cache(ctx, c, "vf_gicache_set_params");
    if (rc) return rc;
    if (!host_in) return gc_fail_null(ctx, "vf_gicache_set_params");
    if ((rc = vf::no_capture(ctx, "vf_gicache_set_params"))) return rc;
    return upload_params(ctx, c, host_in);
}

int vf_gicache_encode(vf_ctx *ctx, vf_gicache *c, const vf_gi_record *recs, int64_t n, float *enc) {
    int rc = check_cache(ctx, c, "vf_gicache_encode");
    if (rc) return rc;
    if (n < 0 || n > (1ll << 30)) return vf::fail(ctx, VF_E_INVALID, "vf_gicache_encode: bad count");
    if (n == 0) return VF_OK;
    if (!recs || !enc) return gc_fail_null(ctx, "vf_gicache_encode");
    hipLaunchKernelGGL(k_mlp_encode, dim3(blocks_for(n, 64)), dim3(64), 0, ctx->stream,
                       reinterpret_cast<const float4 *>(recs), (long long)n, c->p.enc_shift, enc);
    return launch_check(ctx, "k_mlp_encode");
}

int vf_gicache_infer(vf_ctx *ctx, vf_gicache *c, const vf_gi_record *recs, int64_t n, uint16_t *out_rgbh) {
    int rc = check_cache(ctx, c, "vf_gicache_infer");
    if (rc) return rc;
    if (n < 0 || n > (1ll << 30)) return vf::fail(ctx, VF_E_INVALID, "vf_gicache_infer: bad count");
    if (n == 0) return VF_OK;
    if (!recs || !out_rgbh) return gc_fail_null(ctx, "vf_gicache_infer");
    InferArgs I;
    std::memset(&I, 0, sizeof I);
    I.recs = reinterpret_cast<const float4 *>(recs);
    I.n = n;
    I.shift = c->p.enc_shift;
    I.ifrag = reinterpret_cast<const bf16x8 *>(c->ifrag);
    I.w = c->we;
    I.out_h = out_rgbh;
    if ((rc = vf::begin_timed(ctx))) return rc;
    hipLaunchKernelGGL(k_mlp_infer<0>, dim3(infer_groups(c, n)), dim3(256), 0, ctx->stream, I);
    if ((rc = launch_check(ctx, "k_mlp_infer"))) return rc;
    return vf::end_timed(ctx);
}

int vf_gicache_grad(vf_ctx *ctx, vf_gicache *c, const vf_gi_record *samples, int64_t n, float *grad, float *loss) {
    int rc = check_cache(ctx, c, "vf_gicache_grad");
    if (rc) return rc;
    if (n <= 0 || n > (1ll << 26)) return vf::fail(ctx, VF_E_INVALID, "vf_gicache_grad: bad count");
    if (!samples || !grad || !loss) return gc_fail_null(ctx, "vf_gicache_grad");
    int rows = 0;
    if ((rc = launch_train(c, reinterpret_cast<const float4 *>(samples), n, &rows))) return rc;
    if ((rc = launch_adam(c, 1, nullptr, grad, rows))) return rc;
    return launch_stats(c, rows, loss);
}

int vf_gicache_adam(vf_ctx *ctx, vf_gicache *c, const float *grad) {
    int rc = check_cache(ctx, c, "vf_gicache_adam");
    if (rc) return rc;
    if (!grad) return gc_fail_null(ctx, "vf_gicache_adam");
    return launch_adam(c, 2, grad, nullptr, 0);
}

int vf_gicache_train(vf_ctx *ctx, vf_gicache *c, const vf_gi_record *samples, int64_t n) {
    int rc = check_cache(ctx, c, "vf_gicache_train");
    if (rc) return rc;
    if (n <= 0 || n > (1ll << 26)) return vf::fail(ctx, VF_E_INVALID, "vf_gicache_train: bad count");
    if (!samples) return gc_fail_null(ctx, "vf_gicache_train");
    int rows = 0;
    if ((rc = vf::begin_timed(ctx))) return rc;
    if ((rc = launch_train(c, reinterpret_cast<const float4 *>(samples), n, &rows))) return rc;
    if ((rc = launch_adam(c, 2, nullptr, nullptr, rows))) return rc;
    return vf::end_timed(ctx);
}

}  // extern "C"
