// Learned pruning of YZ station-3 units: a 16-64-64-1 ReLU perceptron run on
// the matrix cores (v_mfma_f32_32x32x8_bf16: bf16 operands, fp32 sums), one
// wave scoring up to 64 units at a time.  No counterpart in the reference: the
// plain chain evaluates every unit.
//
// Layers are computed transposed, hidden x units, so that the accumulator
// layout of one layer is the B-operand layout of the next:
//   A (32 x 8):  lane l holds row l%32, columns 4*(l/32) .. +3
//   B (8 x 32):  lane l holds rows 4*(l/32) .. +3, column l%32
//   C (32 x 32): lane l, register r holds row 8*(r/4) + 4*(l/32) + r%4, column l%32
// With the units as columns, the registers 4s..4s+3 of a layer-1 tile are the
// B operand of layer-2 step s of that tile's rows.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ordev {

constexpr int PRUNE_NF = 16;   // features per unit
constexpr int PRUNE_NH = 64;   // hidden width of both layers
// flat weight vector: w1[64][16], b1[64], w2[64][64], b2[64], w3[64], b3
constexpr int PRUNE_NW = PRUNE_NH * PRUNE_NF + PRUNE_NH + PRUNE_NH * PRUNE_NH + PRUNE_NH + PRUNE_NH + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Weights in per-lane A-operand order ([row tile][k step][lane]), biases and
// the output layer in fp32, the keep threshold and the per-tracklet margin.
struct PrunerDev {
    bf16x4 w1[2][2][64];
    bf16x4 w2[2][8][64];
    alignas(16) float b1[PRUNE_NH];
    alignas(16) float b2[PRUNE_NH];
    alignas(16) float w3[PRUNE_NH];
    float b3, thr;
    float margin;   // keep a unit only within margin of its tracklet's best score (INFINITY: off)
};
static_assert(offsetof(PrunerDev, b1) % 16 == 0 && offsetof(PrunerDev, b2) % 16 == 0 &&
                  offsetof(PrunerDev, w3) % 16 == 0, "bias rows are read as float4");

// fp32 -> bf16, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t bf16_rne(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct UnitFeatures {
    float v[PRUNE_NF];
};

// Score of the unit whose features this lane holds.  Every lane of the wave
// calls it (lanes without a unit pass zeros and ignore the result).  With
// LAYERS, the pre-activations (bias added, before ReLU) of both hidden layers
// of unit u = row0 + (the lane's unit) are also written to pre1/pre2[u * 64 + i]
// for rows below nrow (a diagnostic of the matrix-core lane layout).
template <bool LAYERS>
__device__ __attribute__((noinline)) float prune_eval(const PrunerDev* Wp, UnitFeatures f, float* pre1, float* pre2,
                                                      int64_t row0, int64_t nrow) {
    // the weights live in device global memory: global loads (not flat ones, which
    // also count against the LDS counter the feature shuffles wait on)
    const __attribute__((address_space(1))) PrunerDev* W = (const __attribute__((address_space(1))) PrunerDev*)Wp;
    const int lane = threadIdx.x & 63, half = lane >> 5, col = lane & 31;
    float total[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        // B operand of layer 1: the features of unit nt*32 + col
        float g[PRUNE_NF];
#pragma unroll
        for (int q = 0; q < PRUNE_NF; ++q) g[q] = __shfl(f.v[q], nt * 32 + col, 64);
        bf16x4 xb[2];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
            for (int q = 0; q < 4; ++q) xb[kk][q] = (short)bf16_rne(half ? g[kk * 8 + 4 + q] : g[kk * 8 + q]);
        f32x16 h[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            h[mt] = f32x16{};
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) h[mt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(W->w1[mt][kk][lane], xb[kk], h[mt], 0, 0, 0);
        }
        // bias, ReLU, bf16: the B operand of layer 2, step mt*4 + s (rows
        // mt*32 + 8s + 4*half .. +3, one 16-byte load)
        bf16x4 hb[8];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i0 = mt * 32 + 8 * s + 4 * half;   // 16-byte aligned row of 4
                const float bq[4] = {W->b1[i0], W->b1[i0 + 1], W->b1[i0 + 2], W->b1[i0 + 3]};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float pre = h[mt][4 * s + q] + bq[q];
                    if (LAYERS && row0 + nt * 32 + col < nrow) pre1[(row0 + nt * 32 + col) * PRUNE_NH + i0 + q] = pre;
                    hb[mt * 4 + s][q] = (short)bf16_rne(fmaxf(pre, 0.f));
                }
            }
        float p = 0.f;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            f32x16 o = f32x16{};
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) o = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(W->w2[mt][kk][lane], hb[kk], o, 0, 0, 0);
#pragma unroll
            for (int g4 = 0; g4 < 4; ++g4) {
                const int i0 = mt * 32 + 8 * g4 + 4 * half;
                const float bq[4] = {W->b2[i0], W->b2[i0 + 1], W->b2[i0 + 2], W->b2[i0 + 3]};
                const float wq[4] = {W->w3[i0], W->w3[i0 + 1], W->w3[i0 + 2], W->w3[i0 + 3]};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float pre = o[4 * g4 + q] + bq[q];
                    if (LAYERS && row0 + nt * 32 + col < nrow) pre2[(row0 + nt * 32 + col) * PRUNE_NH + i0 + q] = pre;
                    p += fmaxf(pre, 0.f) * wq[q];
                }
            }
        }
        // lanes col and col + 32 hold the two halves of the rows
        total[nt] = p + __shfl_xor(p, 32, 64);
    }
    return (half ? total[1] : total[0]) + W->b3;
}

__device__ __forceinline__ float prune_score(const PrunerDev* W, UnitFeatures f) {
    return prune_eval<false>(W, f, nullptr, nullptr, 0, 0);
}

// Host-side packing of the flat weight vector (layout above) into PrunerDev.
inline void prune_pack(const float* w, float thr, float margin, PrunerDev* out) {
    const float* w1 = w;
    const float* b1 = w1 + PRUNE_NH * PRUNE_NF;
    const float* w2 = b1 + PRUNE_NH;
    const float* b2 = w2 + PRUNE_NH * PRUNE_NH;
    const float* w3 = b2 + PRUNE_NH;
    const float b3 = w3[PRUNE_NH];
    for (int mt = 0; mt < 2; ++mt)
        for (int l = 0; l < 64; ++l) {
            const int row = mt * 32 + (l % 32);
            for (int kk = 0; kk < 2; ++kk)
                for (int q = 0; q < 4; ++q)
                    out->w1[mt][kk][l][q] = (short)bf16_rne(w1[row * PRUNE_NF + kk * 8 + 4 * (l / 32) + q]);
            for (int kk = 0; kk < 8; ++kk)
                for (int q = 0; q < 4; ++q)
                    out->w2[mt][kk][l][q] = (short)bf16_rne(w2[row * PRUNE_NH + kk * 8 + 4 * (l / 32) + q]);
        }
    for (int i = 0; i < PRUNE_NH; ++i) { out->b1[i] = b1[i]; out->b2[i] = b2[i]; out->w3[i] = w3[i]; }
    out->b3 = b3;
    out->thr = thr;
    out->margin = margin;
}

}  // namespace ordev
