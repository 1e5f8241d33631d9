// hvx_math.h — reproducible scalar math for host and device code.
//
// Every function here is written with IEEE-754 add, subtract, multiply,
// divide and square root only (plus exact bit manipulation), evaluated in the
// order written. Built with FP contraction off on both sides (g++ and hipcc),
// on a CPU with SSE2 arithmetic and on gfx950 (whose fp32/fp64 divide and
// square root are correctly rounded under hipcc's defaults), each function
// returns the same bits on the host and on the GPU. That is what lets light
// lists match bit for bit and shaded/filtered values match to the last ulp.
//
// Methods (all from the mathematics, no vendor library constants):
//   log2f   mantissa/exponent split, m in [sqrt(.5), sqrt(2)],
//           log2(m) = (2/ln2) * atanh(s), s = (m-1)/(m+1), odd Taylor series
//           of atanh to s^11 (truncation < 2e-9 relative).
//   exp2f   x = n + f, |f| <= 1/2, 2^f by its Taylor series in f
//           (coefficients ln2^k / k!, k <= 8), then exact scaling by 2^n.
//   powf    exp2f(y * log2f(x)) for x > 0 (the GLSL definition).
//   sin/cos Cody-Waite reduction by pi/2 with a three-part constant
//           (33 + 33 + rest bits), Taylor series on [-pi/4, pi/4] in double,
//           rounded once to float.
//   atan    reduction |t|<=1 by reciprocal, then t -> (t*sqrt3-1)/(t+sqrt3)
//           (+pi/6) to |t| <= tan(pi/12), odd Taylor series to t^25, double.
//   atan2   quadrant logic of C99 atan2 over atan.
//   acos    atan2(sqrt((1-x)(1+x)), x), double.
// The float versions of the double-evaluated functions are rounded once;
// they are used only outside hot loops (per texel or in per-sample tables).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HVX_HD __host__ __device__ inline
#else
#define HVX_HD inline
#endif

namespace hvx {

HVX_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
HVX_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
HVX_HD uint64_t d2u(double d) { return __builtin_bit_cast(uint64_t, d); }

constexpr float kPiF = 3.14159265358979323846f;   // GLSL PI literal as float
constexpr double kPiD = 3.141592653589793;
constexpr float kInfF = __builtin_huge_valf();

HVX_HD float nanf_() { return u2f(0x7fc00000u); }

// GLSL max/min/clamp: max(x, y) = x < y ? y : x; min(x, y) = y < x ? y : x.
HVX_HD float gmax(float a, float b) { return a < b ? b : a; }
HVX_HD float gmin(float a, float b) { return b < a ? b : a; }
HVX_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
HVX_HD float sqrtf_(float x) { return __builtin_sqrtf(x); }
HVX_HD double sqrt_(double x) { return __builtin_sqrt(x); }
HVX_HD float floorf_(float x) { return __builtin_floorf(x); }
HVX_HD double floor_(double x) { return __builtin_floor(x); }
HVX_HD float absf(float x) { return __builtin_fabsf(x); }
HVX_HD double absd(double x) { return __builtin_fabs(x); }

// ------------------------------------------------------------------ log2f
HVX_HD float log2f_(float x) {
    if (!(x > 0.0f)) {
        return x == 0.0f ? -kInfF : nanf_();  // x < 0 or NaN -> NaN
    }
    uint32_t b = f2u(x);
    if (b >= 0x7f800000u) return x;  // +inf
    int e = 0;
    if (b < 0x00800000u) {  // subnormal: scale into the normal range
        x = x * 8388608.0f;   // 2^23, exact
        b = f2u(x);
        e = -23;
    }
    e += (int)(b >> 23) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 1.41421354f) {
        m = m * 0.5f;  // exact
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);  // |s| <= 0.1716
    const float s2 = s * s;
    float p = 0.2623081892525388f;   // 2/(11 ln2)
    p = p * s2 + 0.3205988979753252f;   // 2/(9 ln2)
    p = p * s2 + 0.4121985831111324f;   // 2/(7 ln2)
    p = p * s2 + 0.5770780163555853f;   // 2/(5 ln2)
    p = p * s2 + 0.9617966939259756f;   // 2/(3 ln2)
    p = p * s2 + 2.8853900817779268f;   // 2/ln2
    return (float)e + s * p;
}

// ------------------------------------------------------------------ exp2f
HVX_HD float exp2f_(float x) {
    if (x != x) return x;
    if (x >= 128.0f) return kInfF;
    if (x < -150.0f) return 0.0f;
    const float n = floorf_(x + 0.5f);
    const float f = x - n;  // exact, |f| <= 1/2 (+1 ulp)
    float p = 1.321548679014431e-06f;   // ln2^8/8!
    p = p * f + 1.5252733804059841e-05f;   // ln2^7/7!
    p = p * f + 0.0001540353039338161f;   // ln2^6/6!
    p = p * f + 0.0013333558146428443f;   // ln2^5/5!
    p = p * f + 0.009618129107628477f;   // ln2^4/4!
    p = p * f + 0.05550410866482158f;   // ln2^3/3!
    p = p * f + 0.24022650695910072f;   // ln2^2/2!
    p = p * f + 0.6931471805599453f;   // ln2
    p = p * f + 1.0f;
    const int ni = (int)n;
    if (ni > 127) return (p * 2.0f) * u2f(254u << 23);
    if (ni >= -126) return p * u2f((uint32_t)(ni + 127) << 23);
    return (p * u2f((uint32_t)(ni + 127 + 64) << 23)) * u2f(63u << 23);  // * 2^-64
}

// GLSL pow(x, y) = exp2(y * log2(x)); x < 0 is undefined in GLSL (NaN here),
// pow(0, y > 0) = 0, pow(x, 0) = 1.
HVX_HD float powf_(float x, float y) {
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : kInfF;
    if (!(x > 0.0f)) return nanf_();
    return exp2f_(y * log2f_(x));
}

// pow(x, 5.0) evaluated as an exact integer power: (x^2)^2 * x.
HVX_HD float pow5f(float x) {
    const float x2 = x * x;
    return (x2 * x2) * x;
}

// -------------------------------------------------------- sin/cos (double)
// pi/2 = P1 + P2 + P3; P1, P2 carry 33 significant bits so k*P1, k*P2 are
// exact for |k| < 2^20.
constexpr double kPio2_1 = 1.5707963267341256;
constexpr double kPio2_2 = 6.077100506303966e-11;
constexpr double kPio2_3 = 2.0222662487959506e-21;
constexpr double kTwoOverPi = 0.6366197723675814;

HVX_HD double sin_poly_(double r) {  // |r| <= pi/4
    const double r2 = r * r;
    double p = -7.647163731819816e-13;   // -1/15!
    p = p * r2 + 1.6059043836821613e-10;  // 1/13!
    p = p * r2 - 2.505210838544172e-08;   // -1/11!
    p = p * r2 + 2.7557319223985893e-06;  // 1/9!
    p = p * r2 - 0.0001984126984126984;   // -1/7!
    p = p * r2 + 0.008333333333333333;    // 1/5!
    p = p * r2 - 0.16666666666666666;     // -1/3!
    return r + (r * r2) * p;
}

HVX_HD double cos_poly_(double r) {  // |r| <= pi/4
    const double r2 = r * r;
    double p = 4.779477332387385e-14;    // 1/16!
    p = p * r2 - 1.1470745597729725e-11;  // -1/14!
    p = p * r2 + 2.08767569878681e-09;    // 1/12!
    p = p * r2 - 2.755731922398589e-07;   // -1/10!
    p = p * r2 + 2.48015873015873e-05;    // 1/8!
    p = p * r2 - 0.001388888888888889;    // -1/6!
    p = p * r2 + 0.041666666666666664;    // 1/4!
    p = p * r2 - 0.5;                     // -1/2!
    return 1.0 + r2 * p;
}

HVX_HD void reduce_pio2_(double x, double* r, int* q) {
    const double k = floor_(x * kTwoOverPi + 0.5);
    *r = ((x - k * kPio2_1) - k * kPio2_2) - k * kPio2_3;
    *q = (int)((int64_t)k & 3);
}

HVX_HD double sin_d(double x) {
    double r;
    int q;
    reduce_pio2_(x, &r, &q);
    switch (q) {
        case 0: return sin_poly_(r);
        case 1: return cos_poly_(r);
        case 2: return -sin_poly_(r);
        default: return -cos_poly_(r);
    }
}

HVX_HD double cos_d(double x) {
    double r;
    int q;
    reduce_pio2_(x, &r, &q);
    switch (q) {
        case 0: return cos_poly_(r);
        case 1: return -sin_poly_(r);
        case 2: return -cos_poly_(r);
        default: return sin_poly_(r);
    }
}

// ------------------------------------------------------ atan family (double)
constexpr double kSqrt3 = 1.7320508075688772;
constexpr double kPio6 = 0.5235987755982989;
constexpr double kPio2 = 1.5707963267948966;

HVX_HD double atan_d(double x) {
    if (x != x) return x;
    const bool neg = x < 0.0;
    double t = neg ? -x : x;
    bool inv = false;
    if (t > 1.0) {
        t = 1.0 / t;
        inv = true;
    }
    double off = 0.0;
    if (t > 0.2679491924311227) {  // tan(pi/12)
        t = (t * kSqrt3 - 1.0) / (t + kSqrt3);
        off = kPio6;
    }
    const double t2 = t * t;
    // odd Taylor series t - t^3/3 + ... - t^23/23 + t^25/25
    double p = 1.0 / 25.0;
    p = p * t2 - 1.0 / 23.0;
    p = p * t2 + 1.0 / 21.0;
    p = p * t2 - 1.0 / 19.0;
    p = p * t2 + 1.0 / 17.0;
    p = p * t2 - 1.0 / 15.0;
    p = p * t2 + 1.0 / 13.0;
    p = p * t2 - 1.0 / 11.0;
    p = p * t2 + 1.0 / 9.0;
    p = p * t2 - 1.0 / 7.0;
    p = p * t2 + 1.0 / 5.0;
    p = p * t2 - 1.0 / 3.0;
    double r = off + (t + (t * t2) * p);
    if (inv) r = kPio2 - r;
    return neg ? -r : r;
}

HVX_HD double atan2_d(double y, double x) {
    if (x != x || y != y) return x + y;
    const bool ysign = (d2u(y) >> 63) != 0;
    const bool xsign = (d2u(x) >> 63) != 0;
    if (y == 0.0) {
        if (!xsign) return y;                 // +-0
        return ysign ? -kPiD : kPiD;          // x < 0 or x = -0
    }
    if (x == 0.0) return ysign ? -kPio2 : kPio2;
    const double a = atan_d(y / x);
    if (!xsign) return a;
    return ysign ? a - kPiD : a + kPiD;
}

HVX_HD double acos_d(double x) {
    if (!(x >= -1.0 && x <= 1.0)) return (double)nanf_();
    return atan2_d(sqrt_((1.0 - x) * (1.0 + x)), x);
}

// Float front ends of the double-evaluated functions: one final rounding.
HVX_HD float sinf_(float x) { return (float)sin_d((double)x); }
HVX_HD float cosf_(float x) { return (float)cos_d((double)x); }
HVX_HD float atan2f_(float y, float x) { return (float)atan2_d((double)y, (double)x); }
HVX_HD float acosf_(float x) { return (float)acos_d((double)x); }

// ------------------------------------------------------------- Hammersley
// PBR/Hammersley.glsl:4-19. Bit reversal, then float(bits) * 2^-32.
HVX_HD uint32_t bitrev32(uint32_t b) {
    b = (b << 16u) | (b >> 16u);
    b = ((b & 0x55555555u) << 1u) | ((b & 0xAAAAAAAAu) >> 1u);
    b = ((b & 0x33333333u) << 2u) | ((b & 0xCCCCCCCCu) >> 2u);
    b = ((b & 0x0F0F0F0Fu) << 4u) | ((b & 0xF0F0F0F0u) >> 4u);
    b = ((b & 0x00FF00FFu) << 8u) | ((b & 0xFF00FF00u) >> 8u);
    return b;
}
HVX_HD float radical_inverse_vdc(uint32_t i) {
    return (float)bitrev32(i) * 2.3283064365386963e-10f;  // 2^-32, exact
}

}  // namespace hvx
