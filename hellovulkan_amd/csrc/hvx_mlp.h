// hvx_mlp.h — definition of the per-environment radiance network, shared by
// the library's kernels and the FP64 oracle.
//
// Input: a unit direction d and a mip level of detail lod in [0, lod_max] of
// the prefiltered specular chain. Positional encoding: x = (d.x, d.y, d.z,
// 2 lod / lod_max - 1); feature 32 j + 2 k + s = s ? cos(f_k x_j) : sin(f_k x_j)
// with f_k = pi 2^(k/2), k = 0..15 (128 features). Then three 64-unit ReLU
// layers and a linear RGB layer that predicts log(radiance + kMlpEps), where
// the radiance is the trilinear sample of the chain at (d, lod) (the value
// the clustered shading reads with textureLod, Ambient.glsl:1-29).
// Parameters (FP32, this order): W1[64][128] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[3][64] b4[3]; W[o][i] multiplies input i of output o.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define HVX_MLP_HD __host__ __device__
#else
#define HVX_MLP_HD
#endif

namespace hvx {

constexpr int kMlpIn = 128, kMlpHidden = 64, kMlpOut = 3, kMlpFreqs = 16;
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpHidden * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;
constexpr int kMlpB3 = kMlpW3 + kMlpHidden * kMlpHidden;
constexpr int kMlpW4 = kMlpB3 + kMlpHidden;
constexpr int kMlpB4 = kMlpW4 + kMlpOut * kMlpHidden;
constexpr int kMlpParams = kMlpB4 + kMlpOut;  // 16771
constexpr double kMlpEps = 1.0e-3;
constexpr double kMlpPi = 3.14159265358979323846;

// f_k = pi 2^(k/2)
template <class T> HVX_MLP_HD inline T mlp_freq(int k) {
    T f = T(kMlpPi);
    for (int i = 0; i < k / 2; ++i) f *= T(2);
    if (k & 1) f *= T(1.41421356237309504880);
    return f;
}

}  // namespace hvx
