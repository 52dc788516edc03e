// The warm-start network: a 3-layer MLP (32 inputs, two hidden layers of 128,
// 2 outputs) with BF16 weights and FP32 accumulation on the gfx950 matrix
// cores (v_mfma_f32_16x16x32_bf16), and its training (forward, backward,
// Adam).  It has no counterpart in the reference.
//
// Layout.  Activations are kept transposed, units x samples, so the output
// of one layer is the B operand of the next without a trip through LDS:
//   out^T [units x 16 samples] = W [units x K] * in^T [K x 16 samples].
// In a 16x16x32 MFMA, lane l = 16 g + c holds, for A, row c and K slots
// 8g..8g+7; for B, column c and the same K slots; for C, column c and rows
// 4g..4g+3.  A layer output tile t (16 units) is thus spread as units
// 16t + 4g + i over the four lane groups.  Two tiles (2c, 2c+1) form one
// 32-wide K chunk of the next layer with, in lane group g, the units
//   slot j < 4 : 32c + 4g + j          slot j >= 4 : 32c + 16 + 4g + (j - 4)
// and the A operand (weights) is read with the same unit order, so the
// products pair the right terms whatever order the hardware sums K in.
//
// Parameters (FP32 master copy, KT_MLP_PARAMS floats):
//   W1 [128][32], b1 [128], W2 [128][128], b2 [128], W3 [2][128], b3 [2].
// Device copies used by the kernels: W1, W2, W3 (rows 2..15 zero) as BF16,
// rounded to nearest even, and the biases as FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "kt_internal.h"

namespace kt {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kIn = KT_MLP_INPUTS;     // 32
constexpr int kHid = KT_MLP_HIDDEN;    // 128
constexpr int kOut = KT_MLP_OUTPUTS;   // 2
constexpr int kOutPad = 16;
constexpr uint32_t oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid,
                   oW3 = oB2 + kHid, oB3 = oW3 + kOut * kHid, kParams = oB3 + kOut;
static_assert(kParams == KT_MLP_PARAMS, "parameter count");
// BF16 weight block: W1 | W2 | W3 padded to 16 rows
constexpr uint32_t bW1 = 0, bW2 = bW1 + kHid * kIn, bW3 = bW2 + kHid * kHid, kWeightHalfs = bW3 + kOutPad * kHid;
// FP32 bias block: b1 | b2 | b3
constexpr uint32_t fB1 = 0, fB2 = fB1 + kHid, fB3 = fB2 + kHid, kBiasFloats = fB3 + kOut;

constexpr int kTrainSamples = 32;  // samples per training block (2 waves x 16)
constexpr int kPad = kTrainSamples + 8;

__device__ __forceinline__ uint16_t toBf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float fromBf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

union Frag {
    bf16x8 v;
    uint16_t h[8];
    uint4 q;
    uint2 d[2];
};

__device__ __forceinline__ floatx4 mma(const Frag &a, const Frag &b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
}

// A operand of a 32-wide chunk c of a 128-unit input, rows from a row-major
// [rows][128] BF16 matrix.
__device__ __forceinline__ Frag chunkRow(const uint16_t *__restrict__ w, uint32_t row, int c, int g) {
    Frag f;
    const uint16_t *base = w + row * kHid + 32 * c + 4 * g;
    f.d[0] = *reinterpret_cast<const uint2 *>(base);
    f.d[1] = *reinterpret_cast<const uint2 *>(base + 16);
    return f;
}

// Packs the activations of tiles 2c and 2c+1 (lane's four rows each) into the
// B operand of chunk c.
__device__ __forceinline__ Frag packChunk(const float (*v)[4], int c) {
    Frag f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f.h[i] = toBf16(v[2 * c][i]);
        f.h[4 + i] = toBf16(v[2 * c + 1][i]);
    }
    return f;
}

struct Net {
    const uint16_t *w;  // kWeightHalfs
    const float *b;     // kBiasFloats
};

// Forward pass of one wave over 16 samples whose 32 BF16 inputs the lane
// group g holds in slots 8g..8g+7 (x).  Leaves the post-activation values of
// both hidden layers (lane's column, rows 16t + 4g + i) and the outputs
// (lane group 0 only: o[0], o[1]).
__device__ __forceinline__ void forward(const Net &net, const Frag &x, int c, int g, float (*h1)[4],
                                        float (*h2)[4], float *o) {
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        Frag a;
        a.q = *reinterpret_cast<const uint4 *>(net.w + bW1 + (16 * t + c) * kIn + 8 * g);
        const floatx4 acc = mma(a, x, zero);
#pragma unroll
        for (int i = 0; i < 4; ++i) h1[t][i] = fmaxf(acc[i] + net.b[fB1 + 16 * t + 4 * g + i], 0.f);
    }
    Frag hb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) hb[k] = packChunk(h1, k);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        floatx4 acc = zero;
#pragma unroll
        for (int k = 0; k < 4; ++k) acc = mma(chunkRow(net.w + bW2, 16 * t + c, k, g), hb[k], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) h2[t][i] = fmaxf(acc[i] + net.b[fB2 + 16 * t + 4 * g + i], 0.f);
    }
    floatx4 acc = zero;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = mma(chunkRow(net.w + bW3, c, k, g), packChunk(h2, k), acc);
    o[0] = acc[0] + net.b[fB3 + 0];
    o[1] = acc[1] + net.b[fB3 + 1];
}

// Inference: out[s] = MLP(feat[s]) for s < count (feat [count][32] BF16),
// with the weights in LDS and two 16-sample tiles per wave, grid-stride: each
// weight fragment is read once from LDS and used by two MFMAs.  Rows are
// padded (W1 to 40 halves, W2 and W3 to 136) so the lanes of a fragment read
// spread over the LDS banks.  Same MFMA sequence per tile as forward(), so
// the outputs are the same bit for bit.
constexpr int kS1 = kIn + 8, kS2 = kHid + 8;
constexpr uint32_t lW1 = 0, lW2 = lW1 + kHid * kS1, lW3 = lW2 + kHid * kS2, kLdsHalfs = lW3 + kOutPad * kS2;
static_assert(kLdsHalfs % 8 == 0, "LDS weights in whole 16-byte words");

__device__ __forceinline__ Frag packPair(const float (&lo)[4], const float (&hi)[4]) {
    Frag f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f.h[i] = toBf16(lo[i]);
        f.h[4 + i] = toBf16(hi[i]);
    }
    return f;
}

// A operand of chunk k of a 128-unit input from a padded LDS matrix.
__device__ __forceinline__ Frag ldsChunkRow(const uint16_t *w, uint32_t row, int k, int g) {
    Frag f;
    const uint16_t *base = w + row * kS2 + 32 * k + 4 * g;
    f.d[0] = *reinterpret_cast<const uint2 *>(base);
    f.d[1] = *reinterpret_cast<const uint2 *>(base + 16);
    return f;
}

__global__ void __launch_bounds__(256) mlpInferLds(Net net, const uint16_t *__restrict__ feat, uint64_t count,
                                                   float2 *__restrict__ out) {
    __shared__ alignas(16) uint16_t sW[kLdsHalfs];
    __shared__ float sB[kBiasFloats];
    // stage the weights (16-byte words; the padding columns stay unread)
    for (uint32_t u = threadIdx.x; u < (kHid * kIn + kHid * kHid + kOutPad * kHid) / 8; u += blockDim.x) {
        const uint32_t h = 8 * u;  // half index in the unpadded block
        uint32_t dst;
        if (h < bW2) dst = lW1 + (h / kIn) * kS1 + h % kIn;
        else if (h < bW3) dst = lW2 + ((h - bW2) / kHid) * kS2 + (h - bW2) % kHid;
        else dst = lW3 + ((h - bW3) / kHid) * kS2 + (h - bW3) % kHid;
        *reinterpret_cast<uint4 *>(sW + dst) = *reinterpret_cast<const uint4 *>(net.w + h);
    }
    for (uint32_t u = threadIdx.x; u < kBiasFloats; u += blockDim.x) sB[u] = net.b[u];
    __syncthreads();
    const int lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
    const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x >> 6);
    for (uint64_t pair = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); pair * 32 < count; pair += waves) {
        const uint64_t s0 = pair * 32 + c, s1 = s0 + 16;
        Frag x0, x1;
        x0.q = s0 < count ? *reinterpret_cast<const uint4 *>(feat + s0 * kIn + 8 * g) : make_uint4(0, 0, 0, 0);
        x1.q = s1 < count ? *reinterpret_cast<const uint4 *>(feat + s1 * kIn + 8 * g) : make_uint4(0, 0, 0, 0);
        // layer 1: tiles 2k and 2k + 1 make chunk k of layer 2's input
        Frag hb0[4], hb1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float v0[2][4], v1[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int t = 2 * k + u;
                Frag a;
                a.q = *reinterpret_cast<const uint4 *>(sW + lW1 + (16 * t + c) * kS1 + 8 * g);
                const floatx4 a0 = mma(a, x0, zero), a1 = mma(a, x1, zero);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float bias = sB[fB1 + 16 * t + 4 * g + i];
                    v0[u][i] = fmaxf(a0[i] + bias, 0.f);
                    v1[u][i] = fmaxf(a1[i] + bias, 0.f);
                }
            }
            hb0[k] = packPair(v0[0], v0[1]);
            hb1[k] = packPair(v1[0], v1[1]);
        }
        // layer 2
        Frag hc0[4], hc1[4];
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            float v0[2][4], v1[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int t = 2 * k2 + u;
                floatx4 a0 = zero, a1 = zero;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const Frag a = ldsChunkRow(sW + lW2, 16 * t + c, k, g);
                    a0 = mma(a, hb0[k], a0);
                    a1 = mma(a, hb1[k], a1);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float bias = sB[fB2 + 16 * t + 4 * g + i];
                    v0[u][i] = fmaxf(a0[i] + bias, 0.f);
                    v1[u][i] = fmaxf(a1[i] + bias, 0.f);
                }
            }
            hc0[k2] = packPair(v0[0], v0[1]);
            hc1[k2] = packPair(v1[0], v1[1]);
        }
        // layer 3
        floatx4 o0 = zero, o1 = zero;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const Frag a = ldsChunkRow(sW + lW3, c, k, g);
            o0 = mma(a, hc0[k], o0);
            o1 = mma(a, hc1[k], o1);
        }
        if (g == 0) {
            if (s0 < count) out[s0] = make_float2(o0[0] + sB[fB3 + 0], o0[1] + sB[fB3 + 1]);
            if (s1 < count) out[s1] = make_float2(o1[0] + sB[fB3 + 0], o1[1] + sB[fB3 + 1]);
        }
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct TrainArgs {
    Net net;
    const uint16_t *feat;     // [count][32] BF16
    const float2 *target;     // [count]
    const float *weight;      // [count] per-sample loss weight, or null (1)
    uint64_t count;
    const uint32_t *samples;  // explicit sample list (gradient check) or null: random draw
    uint64_t used;            // samples in this batch
    uint64_t seed, step;
    float invBatch;           // 1 / batch: the loss is the batch mean of |o - t|^2
    float *partial;           // [blocks][kParams]
    float *lossPart;          // [blocks]
};

// One block = 2 waves = 32 samples: forward, backward and this block's
// contribution to every parameter gradient (written to its own slot; the
// slots are summed in block order by mlpReduce).
__global__ void __launch_bounds__(128) mlpTrain(TrainArgs a) {
    __shared__ uint16_t sX[kIn][kPad];
    __shared__ uint16_t sH1[kHid][kPad];
    __shared__ uint16_t sH2[kHid][kPad];
    __shared__ uint16_t sD1[kHid][kPad];
    __shared__ uint16_t sD2[kHid][kPad];
    __shared__ uint16_t sDo[kOutPad][kPad];
    __shared__ float sDof[2][kTrainSamples];
    __shared__ float sLoss[kTrainSamples];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, c = lane & 15, g = lane >> 4;
    const int sl = wave * 16 + c;  // sample slot in the block
    for (int i = tid; i < kOutPad * kPad; i += blockDim.x) (&sDo[0][0])[i] = 0;

    const uint64_t slot = static_cast<uint64_t>(blockIdx.x) * kTrainSamples + sl;
    const bool valid = slot < a.used;
    uint64_t s = 0;
    if (valid) s = a.samples ? a.samples[slot] : mix64(a.seed ^ mix64(a.step * 0x100000001ull + slot)) % a.count;

    Frag x;
    x.q = valid ? *reinterpret_cast<const uint4 *>(a.feat + s * kIn + 8 * g) : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 8; ++j) sX[8 * g + j][sl] = x.h[j];

    float h1[8][4], h2[8][4], o[2];
    forward(a.net, x, c, g, h1, h2, o);
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            sH1[16 * t + 4 * g + i][sl] = toBf16(h1[t][i]);
            sH2[16 * t + 4 * g + i][sl] = toBf16(h2[t][i]);
        }
    __syncthreads();  // sDo zeroed before the output gradient lands in it
    if (g == 0) {
        float e0 = 0.f, e1 = 0.f;
        if (valid) {
            const float2 t = a.target[s];
            e0 = o[0] - t.x;
            e1 = o[1] - t.y;
        }
        const float w = valid && a.weight ? a.weight[s] : 1.f;
        const float d0 = 2.f * w * e0 * a.invBatch, d1 = 2.f * w * e1 * a.invBatch;
        sDof[0][sl] = d0;
        sDof[1][sl] = d1;
        sDo[0][sl] = toBf16(d0);
        sDo[1][sl] = toBf16(d1);
        sLoss[sl] = w * (e0 * e0 + e1 * e1);
    }
    __syncthreads();

    // dh2 = W3^T dO masked by ReLU'(h2): a rank-2 update, on the ALUs
    const float dO0 = sDof[0][sl], dO1 = sDof[1][sl];
    float d2[8][4];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int u = 16 * t + 4 * g + i;
            const float w0 = fromBf16(a.net.w[bW3 + u]), w1 = fromBf16(a.net.w[bW3 + kHid + u]);
            d2[t][i] = h2[t][i] > 0.f ? w0 * dO0 + w1 * dO1 : 0.f;
            sD2[u][sl] = toBf16(d2[t][i]);
        }
    // dh1 = W2^T dh2 masked by ReLU'(h1): A operand = rows of W2^T (strided reads of W2)
    Frag db[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) db[k] = packChunk(d2, k);
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        floatx4 acc = zero;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            Frag w;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                w.h[j] = a.net.w[bW2 + (32 * k + 4 * g + j) * kHid + 16 * t + c];
                w.h[4 + j] = a.net.w[bW2 + (32 * k + 16 + 4 * g + j) * kHid + 16 * t + c];
            }
            acc = mma(w, db[k], acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) sD1[16 * t + 4 * g + i][sl] = toBf16(h1[t][i] > 0.f ? acc[i] : 0.f);
    }
    __syncthreads();

    // weight gradients over the block's 32 samples (one MFMA K step):
    // dW = dOut [units x samples] * In^T [samples x units]
    float *part = a.partial + static_cast<uint64_t>(blockIdx.x) * kParams;
    for (int tile = wave; tile < 64 + 16 + 8; tile += 2) {
        const uint16_t *ar;
        const uint16_t *br;
        uint32_t base, cols;
        int mt, nt, rowsOut = 16;
        if (tile < 64) {  // dW2: dh2 x h1^T
            mt = tile >> 3;
            nt = tile & 7;
            ar = &sD2[16 * mt + c][8 * g];
            br = &sH1[16 * nt + c][8 * g];
            base = oW2;
            cols = kHid;
        } else if (tile < 80) {  // dW1: dh1 x x^T
            mt = (tile - 64) >> 1;
            nt = (tile - 64) & 1;
            ar = &sD1[16 * mt + c][8 * g];
            br = &sX[16 * nt + c][8 * g];
            base = oW1;
            cols = kIn;
        } else {  // dW3: dO x h2^T (rows 0, 1)
            mt = 0;
            nt = tile - 80;
            ar = &sDo[c][8 * g];
            br = &sH2[16 * nt + c][8 * g];
            base = oW3;
            cols = kHid;
            rowsOut = kOut;
        }
        Frag fa, fb;
        fa.q = *reinterpret_cast<const uint4 *>(ar);
        fb.q = *reinterpret_cast<const uint4 *>(br);
        const floatx4 acc = mma(fa, fb, zero);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 16 * mt + 4 * g + i;
            if (row < rowsOut || tile < 80) part[base + row * cols + 16 * nt + c] = acc[i];
        }
    }
    // bias gradients: row sums of the (BF16) output gradients
    {
        const int u = tid;  // 128 threads, one unit each
        float s1 = 0.f, s2 = 0.f;
        for (int j = 0; j < kTrainSamples; ++j) {
            s1 += fromBf16(sD1[u][j]);
            s2 += fromBf16(sD2[u][j]);
        }
        part[oB1 + u] = s1;
        part[oB2 + u] = s2;
        if (u < kOut) {
            float s3 = 0.f;
            for (int j = 0; j < kTrainSamples; ++j) s3 += fromBf16(sDo[u][j]);
            part[oB3 + u] = s3;
        }
        if (u == 0) {
            float l = 0.f;
            for (int j = 0; j < kTrainSamples; ++j) l += sLoss[j];
            a.lossPart[blockIdx.x] = l;
        }
    }
}

__device__ __forceinline__ void storeDevice(uint32_t p, float v, uint16_t *w, float *b) {
    if (p < oB1) {
        w[bW1 + p] = toBf16(v);
    } else if (p < oW2) {
        b[fB1 + p - oB1] = v;
    } else if (p < oB2) {
        w[bW2 + p - oW2] = toBf16(v);
    } else if (p < oW3) {
        b[fB2 + p - oB2] = v;
    } else if (p < oB3) {
        w[bW3 + p - oW3] = toBf16(v);
    } else {
        b[fB3 + p - oB3] = v;
    }
}

// Master FP32 parameters -> BF16 weights and FP32 biases.
__global__ void __launch_bounds__(256) mlpPack(const float *__restrict__ params, uint16_t *w, float *b) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p < kParams) storeDevice(p, params[p], w, b);
}

struct AdamArgs {
    const float *partial;
    const float *lossPart;
    uint32_t blocks;
    float *params, *m, *v;
    uint16_t *w;
    float *b;
    float *grad;      // optional: the summed gradient
    float *lossOut;   // optional: batch mean loss
    float invBatch;
    float lr, beta1, beta2, eps, bc1, bc2;
    int update;       // 0: only sum the gradient
};

// Sums the per-block slots in block order (deterministic) and takes one Adam
// step (Kingma & Ba 2015, bias-corrected) on each parameter.
__global__ void __launch_bounds__(256) mlpReduce(AdamArgs a) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p == 0 && a.lossOut) {
        double l = 0.0;
        for (uint32_t k = 0; k < a.blocks; ++k) l += a.lossPart[k];
        *a.lossOut = static_cast<float>(l * a.invBatch);
    }
    if (p >= kParams) return;
    float g = 0.f;
    for (uint32_t k = 0; k < a.blocks; ++k) g += a.partial[static_cast<uint64_t>(k) * kParams + p];
    if (a.grad) a.grad[p] = g;
    if (!a.update) return;
    const float m = a.beta1 * a.m[p] + (1.f - a.beta1) * g;
    const float v = a.beta2 * a.v[p] + (1.f - a.beta2) * g * g;
    a.m[p] = m;
    a.v[p] = v;
    const float val = a.params[p] - a.lr * (m / a.bc1) / (sqrtf(v / a.bc2) + a.eps);
    a.params[p] = val;
    storeDevice(p, val, a.w, a.b);
}

// features [count][KT_MLP_FEATURES] FP32 -> [count][32] BF16 (zero padded)
__global__ void __launch_bounds__(256) packFeatures(const float *__restrict__ in, uint64_t count,
                                                    uint16_t *__restrict__ out) {
    const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    if (i >= count * kIn) return;
    const uint64_t s = i / kIn;
    const uint32_t f = static_cast<uint32_t>(i % kIn);
    out[i] = f < KT_MLP_FEATURES ? toBf16(in[s * KT_MLP_FEATURES + f]) : 0;
}

inline uint32_t blocksFor(uint64_t n, uint32_t b) { return static_cast<uint32_t>((n + b - 1) / b); }

}  // namespace mlp

int mlpInferLaunch(hipStream_t stream, const uint16_t *w, const float *b, const uint16_t *feat, uint64_t count,
                   float *out) {
    if (count == 0) return KT_OK;
    // one wave per 32 samples, three blocks of four waves per CU
    const uint64_t pairs = (count + 31) / 32;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((pairs + 3) / 4, 768));
    hipLaunchKernelGGL(mlp::mlpInferLds, dim3(blocks), dim3(256), 0, stream, mlp::Net{w, b}, feat, count,
                       reinterpret_cast<float2 *>(out));
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int mlpPackLaunch(hipStream_t stream, const float *params, uint16_t *w, float *b) {
    hipLaunchKernelGGL(mlp::mlpPack, dim3(mlp::blocksFor(mlp::kParams, 256)), dim3(256), 0, stream, params, w, b);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

uint32_t mlpWeightHalfs() { return mlp::kWeightHalfs; }
uint32_t mlpBiasFloats() { return mlp::kBiasFloats; }

}  // namespace kt

// ----------------------------------------------------------------- trainer

struct kt_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    kt::DevBuf<float> params, m, v, bias, partial, lossPart, grad, loss, targetBuf, weightBuf, featIn;
    kt::DevBuf<uint16_t> w, feat;
    kt::DevBuf<uint32_t> samples;
    uint64_t count = 0;
    bool weighted = false;
    uint64_t adamStep = 0;
    bool haveParams = false;
};

namespace kt {
namespace {

int mlpEnsure(kt_mlp *m, uint32_t blocks) {
    KT_TRY(m->partial.reserve(static_cast<size_t>(blocks) * mlp::kParams));
    KT_TRY(m->lossPart.reserve(blocks));
    return KT_OK;
}

int mlpRun(kt_mlp *m, const uint32_t *samples, uint64_t used, uint64_t seed, bool update, float lr, float *gradOut,
           float *lossOut) {
    using namespace mlp;
    const uint32_t blocks = blocksFor(used, kTrainSamples);
    KT_TRY(mlpEnsure(m, blocks));
    TrainArgs ta{};
    ta.net = Net{m->w.p, m->bias.p};
    ta.feat = m->feat.p;
    ta.target = reinterpret_cast<const float2 *>(m->targetBuf.p);
    ta.weight = m->weighted ? m->weightBuf.p : nullptr;
    ta.count = m->count;
    ta.samples = samples;
    ta.used = used;
    ta.seed = seed;
    ta.step = m->adamStep;
    ta.invBatch = 1.f / static_cast<float>(used);
    ta.partial = m->partial.p;
    ta.lossPart = m->lossPart.p;
    hipLaunchKernelGGL(mlpTrain, dim3(blocks), dim3(128), 0, m->stream, ta);
    KT_HIP(hipGetLastError());
    AdamArgs aa{};
    aa.partial = m->partial.p;
    aa.lossPart = m->lossPart.p;
    aa.blocks = blocks;
    aa.params = m->params.p;
    aa.m = m->m.p;
    aa.v = m->v.p;
    aa.w = m->w.p;
    aa.b = m->bias.p;
    aa.grad = gradOut;
    aa.lossOut = lossOut;
    aa.invBatch = ta.invBatch;
    aa.update = update ? 1 : 0;
    if (update) {
        ++m->adamStep;
        aa.lr = lr;
        aa.beta1 = 0.9f;
        aa.beta2 = 0.999f;
        aa.eps = 1e-8f;
        aa.bc1 = 1.f - std::pow(0.9f, static_cast<float>(m->adamStep));
        aa.bc2 = 1.f - std::pow(0.999f, static_cast<float>(m->adamStep));
    }
    hipLaunchKernelGGL(mlpReduce, dim3(blocksFor(kParams, 256)), dim3(256), 0, m->stream, aa);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

}  // namespace
}  // namespace kt

using namespace kt;

#define KT_MLP_GUARD(m)                                                     \
    do {                                                                    \
        if ((m) == nullptr) {                                               \
            setError("null network");                                       \
            return KT_EINVAL;                                               \
        }                                                                   \
        hipError_t kt_se_ = hipSetDevice((m)->device);                      \
        if (kt_se_ != hipSuccess) return hipFail(kt_se_, "hipSetDevice", __FILE__, __LINE__); \
    } while (0)

extern "C" {

int kt_mlp_create(int device, kt_mlp **out) {
    KT_CHECK(out, KT_EINVAL, "kt_mlp_create: null output");
    *out = nullptr;
    int count = 0;
    KT_HIP(hipGetDeviceCount(&count));
    KT_CHECK(device >= 0 && device < count, KT_EINVAL, "kt_mlp_create: no such device");
    KT_HIP(hipSetDevice(device));
    auto *m = new kt_mlp();
    m->device = device;
    int rc = KT_OK;
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = hipFail(e, "hipStreamCreate", __FILE__, __LINE__);
    if (rc == KT_OK) rc = m->params.reserve(mlp::kParams);
    if (rc == KT_OK) rc = m->m.reserve(mlp::kParams);
    if (rc == KT_OK) rc = m->v.reserve(mlp::kParams);
    if (rc == KT_OK) rc = m->grad.reserve(mlp::kParams);
    if (rc == KT_OK) rc = m->loss.reserve(1);
    if (rc == KT_OK) rc = m->bias.reserve(mlp::kBiasFloats);
    if (rc == KT_OK) rc = m->w.reserve(mlp::kWeightHalfs);
    if (rc == KT_OK) {
        e = hipMemsetAsync(m->w.p, 0, m->w.bytes(), m->stream);  // padding rows of W3 stay zero
        if (e == hipSuccess) e = hipMemsetAsync(m->m.p, 0, m->m.bytes(), m->stream);
        if (e == hipSuccess) e = hipMemsetAsync(m->v.p, 0, m->v.bytes(), m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) rc = hipFail(e, "kt_mlp_create", __FILE__, __LINE__);
    }
    if (rc != KT_OK) {
        kt_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return KT_OK;
}

int kt_mlp_destroy(kt_mlp *m) {
    if (!m) return KT_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    m->params.release();
    m->m.release();
    m->v.release();
    m->bias.release();
    m->partial.release();
    m->lossPart.release();
    m->grad.release();
    m->loss.release();
    m->targetBuf.release();
    m->weightBuf.release();
    m->featIn.release();
    m->w.release();
    m->feat.release();
    m->samples.release();
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return KT_OK;
}

int kt_mlp_set_params(kt_mlp *m, const float *params, uint64_t count) {
    KT_MLP_GUARD(m);
    KT_CHECK(params && count == KT_MLP_PARAMS, KT_EINVAL, "kt_mlp_set_params: need KT_MLP_PARAMS floats");
    for (uint64_t i = 0; i < count; ++i)
        KT_CHECK(std::isfinite(params[i]), KT_EINVAL, "kt_mlp_set_params: non-finite parameter");
    KT_HIP(hipMemcpyAsync(m->params.p, params, count * sizeof(float), hipMemcpyHostToDevice, m->stream));
    KT_HIP(hipMemsetAsync(m->m.p, 0, m->m.bytes(), m->stream));
    KT_HIP(hipMemsetAsync(m->v.p, 0, m->v.bytes(), m->stream));
    KT_TRY(mlpPackLaunch(m->stream, m->params.p, m->w.p, m->bias.p));
    KT_HIP(hipStreamSynchronize(m->stream));
    m->adamStep = 0;
    m->haveParams = true;
    return KT_OK;
}

int kt_mlp_get_params(kt_mlp *m, float *params, uint64_t count) {
    KT_MLP_GUARD(m);
    KT_CHECK(params && count == KT_MLP_PARAMS, KT_EINVAL, "kt_mlp_get_params: need KT_MLP_PARAMS floats");
    KT_CHECK(m->haveParams, KT_ESTATE, "kt_mlp_get_params: no parameters set");
    KT_HIP(hipMemcpyAsync(params, m->params.p, count * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    KT_HIP(hipStreamSynchronize(m->stream));
    return KT_OK;
}

int kt_mlp_set_data(kt_mlp *m, const float *features, const float *targets, const float *weights,
                    uint64_t count) {
    KT_MLP_GUARD(m);
    KT_CHECK(count > 0 && features && targets, KT_EINVAL, "kt_mlp_set_data: empty data");
    KT_CHECK(count < (1ull << 32), KT_EINVAL, "kt_mlp_set_data: too many samples");
    if (weights) {
        for (uint64_t i = 0; i < count; ++i)
            KT_CHECK(std::isfinite(weights[i]) && weights[i] >= 0.f, KT_EINVAL,
                     "kt_mlp_set_data: weights must be finite and non-negative");
        KT_TRY(m->weightBuf.reserve(count));
        KT_HIP(hipMemcpyAsync(m->weightBuf.p, weights, count * sizeof(float), hipMemcpyHostToDevice, m->stream));
    }
    m->weighted = weights != nullptr;
    KT_TRY(m->featIn.reserve(count * KT_MLP_FEATURES));
    KT_TRY(m->feat.reserve(count * KT_MLP_INPUTS));
    KT_TRY(m->targetBuf.reserve(count * KT_MLP_OUTPUTS));
    KT_HIP(hipMemcpyAsync(m->featIn.p, features, count * KT_MLP_FEATURES * sizeof(float), hipMemcpyHostToDevice,
                          m->stream));
    KT_HIP(hipMemcpyAsync(m->targetBuf.p, targets, count * KT_MLP_OUTPUTS * sizeof(float), hipMemcpyHostToDevice,
                          m->stream));
    hipLaunchKernelGGL(mlp::packFeatures, dim3(mlp::blocksFor(count * KT_MLP_INPUTS, 256)), dim3(256), 0,
                       m->stream, m->featIn.p, count, m->feat.p);
    KT_HIP(hipGetLastError());
    KT_HIP(hipStreamSynchronize(m->stream));
    m->count = count;
    return KT_OK;
}

int kt_mlp_forward(kt_mlp *m, const float *features, uint64_t count, float *out) {
    KT_MLP_GUARD(m);
    KT_CHECK(m->haveParams, KT_ESTATE, "kt_mlp_forward: no parameters set");
    KT_CHECK(features && out, KT_EINVAL, "kt_mlp_forward: null array");
    if (count == 0) return KT_OK;
    DevBuf<float> in, res;
    DevBuf<uint16_t> packed;
    int rc = in.reserve(count * KT_MLP_FEATURES);
    if (rc == KT_OK) rc = res.reserve(count * KT_MLP_OUTPUTS);
    if (rc == KT_OK) rc = packed.reserve(count * KT_MLP_INPUTS);
    hipError_t e = hipSuccess;
    if (rc == KT_OK) {
        e = hipMemcpyAsync(in.p, features, count * KT_MLP_FEATURES * sizeof(float), hipMemcpyHostToDevice,
                           m->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(mlp::packFeatures, dim3(mlp::blocksFor(count * KT_MLP_INPUTS, 256)), dim3(256), 0,
                               m->stream, in.p, count, packed.p);
            e = hipGetLastError();
        }
        if (e == hipSuccess) rc = mlpInferLaunch(m->stream, m->w.p, m->bias.p, packed.p, count, res.p);
        if (rc == KT_OK && e == hipSuccess)
            e = hipMemcpyAsync(out, res.p, count * KT_MLP_OUTPUTS * sizeof(float), hipMemcpyDeviceToHost,
                               m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    }
    in.release();
    res.release();
    packed.release();
    if (e != hipSuccess) return hipFail(e, "kt_mlp_forward", __FILE__, __LINE__);
    return rc;
}

int kt_mlp_gradient(kt_mlp *m, const uint32_t *samples, uint64_t count, float *grad, double *loss) {
    KT_MLP_GUARD(m);
    KT_CHECK(m->haveParams && m->count > 0, KT_ESTATE, "kt_mlp_gradient: parameters and data must be set");
    KT_CHECK(samples && count > 0 && grad, KT_EINVAL, "kt_mlp_gradient: null or empty argument");
    for (uint64_t i = 0; i < count; ++i)
        KT_CHECK(samples[i] < m->count, KT_EINVAL, "kt_mlp_gradient: sample index out of range");
    KT_TRY(m->samples.reserve(count));
    KT_HIP(hipMemcpyAsync(m->samples.p, samples, count * sizeof(uint32_t), hipMemcpyHostToDevice, m->stream));
    KT_TRY(mlpRun(m, m->samples.p, count, 0, false, 0.f, m->grad.p, m->loss.p));
    float l = 0.f;
    KT_HIP(hipMemcpyAsync(grad, m->grad.p, mlp::kParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    KT_HIP(hipMemcpyAsync(&l, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    KT_HIP(hipStreamSynchronize(m->stream));
    if (loss) *loss = l;
    return KT_OK;
}

int kt_mlp_train(kt_mlp *m, uint64_t steps, uint32_t batch, double lr, uint64_t seed, float *loss) {
    KT_MLP_GUARD(m);
    KT_CHECK(m->haveParams && m->count > 0, KT_ESTATE, "kt_mlp_train: parameters and data must be set");
    KT_CHECK(batch > 0 && batch % mlp::kTrainSamples == 0, KT_EINVAL, "kt_mlp_train: batch must be a multiple of 32");
    KT_CHECK(lr > 0.0 && std::isfinite(lr), KT_EINVAL, "kt_mlp_train: bad learning rate");
    DevBuf<float> losses;
    if (loss) KT_TRY(losses.reserve(std::max<uint64_t>(steps, 1)));
    int rc = KT_OK;
    for (uint64_t s = 0; s < steps && rc == KT_OK; ++s)
        rc = mlpRun(m, nullptr, batch, seed, true, static_cast<float>(lr), nullptr, loss ? losses.p + s : nullptr);
    hipError_t e = hipSuccess;
    if (rc == KT_OK && loss && steps > 0)
        e = hipMemcpyAsync(loss, losses.p, steps * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    losses.release();
    if (e != hipSuccess) return hipFail(e, "kt_mlp_train", __FILE__, __LINE__);
    return rc;
}

}  // extern "C"
