// mlp_device.h — in-kernel inference of the radiance network (hvx_mlp.h):
// the workgroup stages the BF16 weights in LDS, then each wave evaluates 16
// samples per call with the 16x16x32 BF16 MFMA. The positional encoding is
// computed straight into the first layer's operand layout; hidden
// activations go through a 2 KB per-wave LDS tile between layers. The
// arithmetic matches the training forward pass of mlp.hip (same operands,
// same instruction, bias and ReLU after the FP32 accumulation, BF16
// activations, FP32 output layer).
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
    const uint4* s2 = reinterpret_cast<const uint4*>(v.W2);
    const uint4* s3 = reinterpret_cast<const uint4*>(v.W3);
    uint4* d2 = reinterpret_cast<uint4*>(L.W2);
    uint4* d3 = reinterpret_cast<uint4*>(L.W3);
    for (uint32_t i = t; i < kMlpHidden * kMlpHidden / 8; i += n) {
        d2[i] = s2[i];
        d3[i] = s3[i];
    }
    for (uint32_t i = t; i < kMlpHidden; i += n) {
        L.b1[i] = v.P[kMlpB1 + i];
        L.b2[i] = v.P[kMlpB2 + i];
        L.b3[i] = v.P[kMlpB3 + i];
    }
    for (uint32_t i = t; i < kMlpOut * kMlpHidden; i += n) L.W4[i] = v.P[kMlpW4 + i];
    if (t < kMlpOut) L.b4[t] = v.P[kMlpB4 + t];
}

// ReLU(acc + bias) of a 16 x 64 tile into the wave's LDS tile [16][64].
__device__ __forceinline__ void mlp_store_hidden(const f32x4 acc[4], const float* bias, __bf16* tile) {
    const uint32_t lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float v = fmaxf(acc[nb][r] + bias[16 * nb + i], 0.0f);
            tile[(4 * g + r) * kMlpHidden + 16 * nb + i] = (__bf16)v;
        }
}

__device__ __forceinline__ void mlp_hidden_layer(const __bf16* W, const __bf16* tile, f32x4 acc[4]) {
    const uint32_t lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(tile + i * kMlpHidden + 32 * s + 8 * g);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(W + (16 * nb + i) * kMlpHidden + 32 * s + 8 * g);
            acc[nb] = mfma16x16x32(a, b, acc[nb]);
        }
    }
}

// Forward of 16 samples by one wave. Lane l holds the inputs x[4] = (dir,
// lod) of sample l & 15 (the four lanes of a sample hold the same values);
// lanes with l >> 4 < 3 return channel l >> 4 of the predicted radiance of
// that sample, the others return 0. Every thread of the workgroup must call
// it the same number of times (it synchronises the workgroup).
__device__ __forceinline__ float mlp_forward_tile(const MlpLds& L, __bf16* tile, const float x[4], float lod_max) {
    const uint32_t lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    const float xin[4] = {x[0], x[1], x[2], 2.0f * x[3] / lod_max - 1.0f};
    f32x4 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        // features 32 s + 8 g + q: input s, frequency 4 g + q / 2, sin (q even) or cos
        bf16x8 a;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float arg = mlp_freq<float>(4 * (int)g + q) * xin[s];
            a[2 * q] = (__bf16)sinf(arg);
            a[2 * q + 1] = (__bf16)cosf(arg);
        }
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(L.W1 + (16 * nb + i) * kMlpIn + 32 * s + 8 * g);
            acc[nb] = mfma16x16x32(a, b, acc[nb]);
        }
    }
    __syncthreads();  // the previous call's readers of `tile` are done
    mlp_store_hidden(acc, L.b1, tile);
    __syncthreads();
    mlp_hidden_layer(L.W2, tile, acc);
    __syncthreads();
    mlp_store_hidden(acc, L.b2, tile);
    __syncthreads();
    mlp_hidden_layer(L.W3, tile, acc);
    __syncthreads();
    mlp_store_hidden(acc, L.b3, tile);
    __syncthreads();
    float y = 0.0f;
    if (g < 3) {
        float s = L.b4[g];
        for (int k = 0; k < kMlpHidden; ++k) s += (float)tile[i * kMlpHidden + k] * L.W4[g * kMlpHidden + k];
        y = expf(s) - (float)kMlpEps;
    }
    return y;
}

}  // namespace hvx
