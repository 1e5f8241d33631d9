// mlp_device.h — in-kernel inference of the radiance network (hvx_mlp.h):
// the workgroup stages the BF16 weights in LDS, then each wave evaluates 16
// samples per call with the 16x16x32 BF16 MFMA, all activations staying in
// registers. The positional encoding is computed straight into the first
// layer's operand layout; the FP32 output layer is reduced across lanes from
// the last accumulators. Against the training forward pass of mlp.hip: the
// same encoding bits, BF16 operands, instruction, bias and ReLU after the
// FP32 accumulation and output-layer summation order; layers 2 and 3 feed
// each 32-unit K block to the MFMA in a permuted order (mlp_unit).
#pragma once

#include <hip/hip_runtime.h>

#include "hvx_mfma.h"
#include "hvx_mlp.h"

namespace hvx {

// Device view of a network's parameters (owned by hvx_mlp).
struct MlpView {
    const __bf16* W1;  // [64][128]
    const __bf16* W2;  // [64][64]
    const __bf16* W3;  // [64][64]
    const float* P;    // FP32 parameters (biases, W4, b4 are read from here)
    float lod_max;
};

// Layers 2 and 3 keep their weights with K permuted: W[o][s][g][q] =
// W[o][mlp_unit(s, g, q)], the order in which each lane holds the previous
// layer's units.
struct MlpLds {
    __bf16 W1[kMlpHidden * kMlpIn];
    __bf16 W2[kMlpHidden * kMlpHidden];
    __bf16 W3[kMlpHidden * kMlpHidden];
    float b1[kMlpHidden], b2[kMlpHidden], b3[kMlpHidden];
    float W4[kMlpOut * kMlpHidden];
    float b4[4];
};

// Cooperative copy of the weights into LDS (the caller synchronises).
__device__ __forceinline__ void mlp_stage(MlpLds& L, const MlpView& v) {
    const uint32_t t = threadIdx.x, n = blockDim.x;
    const uint4* s1 = reinterpret_cast<const uint4*>(v.W1);
    uint4* d1 = reinterpret_cast<uint4*>(L.W1);
    for (uint32_t i = t; i < kMlpHidden * kMlpIn / 8; i += n) d1[i] = s1[i];
    for (uint32_t i = t; i < kMlpHidden * kMlpHidden; i += n) {
        const uint32_t o = i / kMlpHidden, r = i % kMlpHidden;
        const uint32_t src = o * kMlpHidden + (uint32_t)mlp_unit((int)(r >> 5), (int)((r >> 3) & 3), (int)(r & 7));
        L.W2[i] = v.W2[src];
        L.W3[i] = v.W3[src];
    }
    for (uint32_t i = t; i < kMlpHidden; i += n) {
        L.b1[i] = v.P[kMlpB1 + i];
        L.b2[i] = v.P[kMlpB2 + i];
        L.b3[i] = v.P[kMlpB3 + i];
    }
    for (uint32_t i = t; i < kMlpOut * kMlpHidden; i += n) L.W4[i] = v.P[kMlpW4 + i];
    if (t < kMlpOut) L.b4[t] = v.P[kMlpB4 + t];
}

// ReLU(acc + bias), rounded to BF16, of the accumulators of one layer: lane
// (g, i) holds unit 16 nb + 4 g + r of sample i in acc[nb][r].
__device__ __forceinline__ void mlp_activate(const f32x4 acc[4], const float* bias, uint32_t g, __bf16 h[4][4]) {
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) h[nb][r] = (__bf16)fmaxf(acc[nb][r] + bias[16 * nb + 4 * g + r], 0.0f);
}

// One 64 -> 64 layer: the B operand of K block s is the lane's own units
// h[2 s][0..3], h[2 s + 1][0..3] (K index 8 g + q <-> mlp_unit(s, g, q)).
__device__ __forceinline__ void mlp_hidden_layer(const __bf16* Wp, const __bf16 h[4][4], uint32_t i, uint32_t g,
                                                 f32x4 acc[4]) {
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        bf16x8 b;
#pragma unroll
        for (int q = 0; q < 8; ++q) b[q] = h[2 * s + (q >> 2)][q & 3];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(Wp + (16 * nb + i) * kMlpHidden + 32 * s + 8 * g);
            acc[nb] = mfma16x16x32(a, b, acc[nb]);
        }
    }
}

// Forward of 16 samples by one wave, entirely in registers (the LDS holds
// only the weights). Lane l holds the inputs x[4] = (dir, lod) of sample
// l & 15 (the four lanes of a sample hold the same values); lanes with
// g = l >> 4 < 3 return channel g of the predicted radiance of that sample,
// the others return 0. Each layer computes its outputs transposed (weights as
// the A operand, samples as the columns), so lane (g, i) ends with 16 units of
// sample i, which are the B operand of the next layer as they are.
__device__ __forceinline__ float mlp_forward_tile(const MlpLds& L, const float x[4], float lod_max) {
    const uint32_t lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    const float xin[4] = {x[0], x[1], x[2], 2.0f * x[3] / lod_max - 1.0f};
    f32x4 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        // features 32 s + 8 g + 2 q + (0 sin, 1 cos): input s, frequency 4 g + q
        float sv[4], cv[4];
        mlp_sincos_direct(xin[s], 4 * (int)g, sv[0], cv[0]);
        mlp_sincos_direct(xin[s], 4 * (int)g + 1, sv[1], cv[1]);
        mlp_sincos_double(sv[0], cv[0], sv[2], cv[2]);
        mlp_sincos_double(sv[1], cv[1], sv[3], cv[3]);
        bf16x8 b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            b[2 * q] = (__bf16)sv[q];
            b[2 * q + 1] = (__bf16)cv[q];
        }
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(L.W1 + (16 * nb + i) * kMlpIn + 32 * s + 8 * g);
            acc[nb] = mfma16x16x32(a, b, acc[nb]);
        }
        // one K block at a time: without the fence the scheduler hoists every
        // block's weight loads and encoding, past the register budget
        __builtin_amdgcn_sched_barrier(0);
    }
    __bf16 h[4][4];
    mlp_activate(acc, L.b1, g, h);
    mlp_hidden_layer(L.W2, h, i, g, acc);
    mlp_activate(acc, L.b2, g, h);
    mlp_hidden_layer(L.W3, h, i, g, acc);
    mlp_activate(acc, L.b3, g, h);
    // output layer in the order of mlp_output: the lane's 16 units, then
    // across the four lane groups (lanes i, i + 16, i + 32, i + 48)
    float y[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float t = (float)h[j >> 2][j & 3] * L.W4[c * kMlpHidden + 16 * (j >> 2) + 4 * g + (j & 3)];
            v = j == 0 ? t : v + t;
        }
        v = v + __shfl_xor(v, 16, 64);
        v = v + __shfl_xor(v, 32, 64);
        y[c] = v + L.b4[c];
    }
    const float yc = g == 0 ? y[0] : (g == 1 ? y[1] : y[2]);
    return g < 3 ? expf(yc) - (float)kMlpEps : 0.0f;
}

}  // namespace hvx
