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

#include <cmath>
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

// sin and cos of f_k x for k = 0, 1 mod 4, as every kernel (and the
// oracle's BF16 emulation) evaluates them: in FP32, each step an IEEE
// operation or a fused multiply-add, so any IEEE target gives the same bits.
// The argument in revolutions, t = f_k x / 2 pi = x 2^(k/2) / 2, is rounded
// once (exact for even k), reduced exactly to r = t - rint(t) in [-1/2, 1/2]
// and to a quarter turn r = q/4 + f, |f| <= 1/8; sin and cos of a = 2 pi f
// come from their Taylor series through a^9 and a^8 (truncation below 3e-8
// for |a| <= pi/4), then the quarter-turn rotation by q.
HVX_MLP_HD inline void mlp_sincos_direct(float x, int k, float& s, float& c) {
    const float m = ldexpf((k & 1) ? 0.70710678118654752f : 0.5f, k >> 1);  // 2^(k/2) / 2
    const float t = x * m;
    const float r = t - rintf(t);
    const float q = rintf(4.0f * r);
    const float f = fmaf(q, -0.25f, r);
    const float a = f * 6.28318530717958647692f;
    const float a2 = a * a;
    float ps = fmaf(a2, 2.7557319223985891e-6f, -1.9841269841269841e-4f);  // 1/9!, -1/7!
    ps = fmaf(a2, ps, 8.3333333333333333e-3f);                             // 1/5!
    ps = fmaf(a2, ps, -1.6666666666666667e-1f);                            // -1/3!
    const float sa = fmaf(a * a2, ps, a);
    float pc = fmaf(a2, 2.4801587301587302e-5f, -1.3888888888888889e-3f);  // 1/8!, -1/6!
    pc = fmaf(a2, pc, 4.1666666666666667e-2f);                             // 1/4!
    pc = fmaf(a2, pc, -0.5f);
    const float ca = fmaf(a2, pc, 1.0f);
    // quadrant qi = q mod 4 (q in [-2, 2]): (s, c) = (sa, ca), (ca, -sa),
    // (-sa, -ca), (-ca, sa) for qi = 0..3, as selects rather than a branch
    // per quadrant (lanes of one wave hold different quadrants)
    const uint32_t qi = (uint32_t)(int)q;
    const float s0 = (qi & 1u) ? ca : sa, c0 = (qi & 1u) ? sa : ca;
    s = (qi & 2u) ? -s0 : s0;
    c = ((qi + 1u) & 2u) ? -c0 : c0;
}

// Frequency k + 2 from frequency k (f_(k+2) = 2 f_k): sin 2u = 2 sin u cos u,
// cos 2u = 1 - 2 sin^2 u.
HVX_MLP_HD inline void mlp_sincos_double(float s, float c, float& s2, float& c2) {
    s2 = 2.0f * (s * c);
    c2 = fmaf(-2.0f * s, s, 1.0f);
}

// The encoding's sin and cos of f_k x: direct for k = 0, 1 mod 4, one
// doubling of frequency k - 2 for k = 2, 3 mod 4. Against sin/cos of the
// exact f_k x: below 3e-7 absolute for even k, below 4e-5 for odd k (the
// rounding of t and of the FP32 sqrt(2)/2 at up to 90.5 turns dominate).
HVX_MLP_HD inline void mlp_sincos(float x, int k, float& s, float& c) {
    if (k & 2) {
        float s0, c0;
        mlp_sincos_direct(x, k - 2, s0, c0);
        mlp_sincos_double(s0, c0, s, c);
    } else {
        mlp_sincos_direct(x, k, s, c);
    }
}

// Hidden unit fed as K index 8 g + q (g, q < 4, 8) of the 32-wide K block s
// of layers 2 and 3 in the fused inference (mlp_device.h): the unit each
// lane already holds in the previous layer's MFMA accumulators.
HVX_MLP_HD inline int mlp_unit(int s, int g, int q) { return 16 * (2 * s + (q >> 2)) + 4 * g + (q & 3); }

// Output channel of the linear layer: b + sum_u h[u] w[u] in the order the
// fused inference reduces it: four partial sums
// p_g = sum over nb = 0..3, then r = 0..3, of h[16 nb + 4 g + r] w[16 nb + 4 g + r]
// (left to right), then (p_0 + p_1) + (p_2 + p_3), then the bias.
HVX_MLP_HD inline float mlp_output(const float* h, const float* w, float b) {
    float p[4];
    for (int g = 0; g < 4; ++g) {
        float v = h[4 * g] * w[4 * g];
        for (int j = 1; j < 16; ++j) {
            const int u = 16 * (j >> 2) + 4 * g + (j & 3);
            v = v + h[u] * w[u];
        }
        p[g] = v;
    }
    return ((p[0] + p[1]) + (p[2] + p[3])) + b;
}

}  // namespace hvx
