// hvx_contract.h — definition of the dense-contraction cube filter (the MFMA
// IBL path), shared by the library (plan construction, host and device) and
// the FP64 oracle. Double precision throughout.
//
// The reference's specular prefilter (CubeFilterSpecular.frag:34-75) is a
// Monte-Carlo estimate, with 1024 Hammersley samples, of
//     out(n) = int L(l) w(n.l) dl / int w(n.l) dl,   w(c) = D_GGX(h(c)) c  (c > 0)
// with V = R = n: the half vector h of (n, l) has (n.h)^2 = (1 + c)/2, the
// sample density in l is D/4, and each sample carries weight n.l. The
// diffuse filter (CubeFilterDiffuse.frag:35-65) is a grid estimate of the
// same integral with w(c) = c. The contraction path evaluates the integral
// as a deterministic quadrature over cube texels instead:
//
//     out(o) = sum_k W(o, k) L_k / sum_k W(o, k),   W(o, k) = w(n_o . l_k) dw_k
//
// with l_k the texel-centre direction and dw_k = (2/S)^2 / (1 + u^2 + v^2)^1.5
// the texel's solid angle at its centre (face-plane coordinates u, v). Per
// output mip the input is a box-filtered level of the input cube chain whose
// texels resolve the lobe (`contract_levels`):
//   * DELTA (roughness 0): the bilinear seamless sample of level 0 at n_o;
//   * FULL: one dense sum over every texel of one input level;
//   * SPLIT (narrow lobes at large outputs): w = w phi + w (1 - phi) with a
//     smooth angular window phi(theta) (1 inside t0, 0 beyond t1). The near
//     part w phi is summed on the gnomonic plane of the output texel's face
//     at input resolution Si, extended past the face edges (texels there are
//     bilinear seamless samples of the same level). The far part w (1 - phi)
//     is summed over every texel of the coarser level Sf at the Sf x Sf
//     texel centres of each face, then bilinearly (seamless) interpolated to
//     n_o; numerator and normaliser are interpolated alike.
// All faces share one set of weights: in face-local coordinates (u, v, 1)
// every face sees the same geometry (the cube's face frames differ by signed
// axis permutations, which preserve angles and texel grids).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define HVX_CONTRACT_HD __host__ __device__
#else
#define HVX_CONTRACT_HD
#endif

namespace hvx {

enum ContractType { CONTRACT_DIFFUSE = 0, CONTRACT_SPECULAR = 1 };
enum ContractKind { CONTRACT_DELTA = 0, CONTRACT_FULL = 1, CONTRACT_SPLIT = 2 };

constexpr double kContractPi = 3.14159265358979323846;
constexpr uint32_t kContractMaxMips = 16;
// largest dense (FULL) weight matrix of one mip, in entries, before the
// near/far split is used instead
constexpr double kContractFullBudget = 134217728.0;  // 2^27
// coarsest input level of a specular mip: the GGX tails that carry a few
// percent of the weight out to a radian are summed against texels of at
// most pi/128 (measured against the level-0 sum on the sun-and-sky
// environment: mip 3 of 512^2 x 10 within 1.5e-3 at 64^2, 5.8e-3 at 32^2)
constexpr uint32_t kContractMinIn = 64;
constexpr uint32_t kContractTile = 16;               // near-field output tile edge

struct ContractLevel {
    int kind;
    uint32_t So;    // output face of this mip
    uint32_t Si;    // input level face (FULL: the summed level; SPLIT: near level)
    uint32_t Sf;    // SPLIT: far level face and the far grid's face
    double alpha2;  // GGX alpha^2, alpha = roughness^2 (specular)
    double t0, t1;  // SPLIT: angular window radii (radians)
};

HVX_CONTRACT_HD inline uint32_t contract_pow2_at_least(double x) {
    uint32_t p = 1;
    while ((double)p < x && p < (1u << 30)) p <<= 1;
    return p;
}

// w(c) of the filter type at cosine c = n.l (0 for c <= 0).
HVX_CONTRACT_HD inline double contract_weight(int type, double alpha2, double c) {
    if (!(c > 0.0)) return 0.0;
    if (type == CONTRACT_DIFFUSE) return c;
    const double nh2 = 0.5 * (1.0 + c);
    const double d = nh2 * (alpha2 - 1.0) + 1.0;
    return alpha2 / (kContractPi * d * d) * c;
}

// Angular window: 1 for theta <= t0, 0 for theta >= t1, C1 smoothstep between.
HVX_CONTRACT_HD inline double contract_phi(double t0, double t1, double c) {
    const double th = acos(c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c));
    if (th <= t0) return 1.0;
    if (th >= t1) return 0.0;
    const double s = (th - t0) / (t1 - t0);
    return 1.0 - s * s * (3.0 - 2.0 * s);
}

// Face-plane coordinate of texel index x (may lie outside [0, S)).
HVX_CONTRACT_HD inline double contract_coord(int x, uint32_t S) { return (2.0 * x + 1.0) / (double)S - 1.0; }

// Solid angle of the texel centred at face-plane (u, v) of a face of size S.
HVX_CONTRACT_HD inline double contract_dw(double u, double v, uint32_t S) {
    const double r2 = 1.0 + u * u + v * v;
    const double q = 2.0 / (double)S;
    return q * q / (r2 * sqrt(r2));
}

// World direction (unnormalised) of face-plane point (u, v) of face f, in the
// cube convention of texel_dir (storage texel (x, y) of face f <-> (u, v)).
HVX_CONTRACT_HD inline void contract_face_point(int f, double u, double v, double p[3]) {
    switch (f) {
        case 0: p[0] = 1.0; p[1] = -v; p[2] = -u; break;
        case 1: p[0] = -1.0; p[1] = -v; p[2] = u; break;
        case 2: p[0] = u; p[1] = 1.0; p[2] = v; break;
        case 3: p[0] = u; p[1] = -1.0; p[2] = -v; break;
        case 4: p[0] = u; p[1] = -v; p[2] = 1.0; break;
        default: p[0] = -u; p[1] = -v; p[2] = -1.0; break;
    }
}

// Integral of w over the sphere, 2 pi int_0^1 w(c) dc: pi for the cosine
// weight; for GGX 2 a^2 I with I = int_0^1 c / (A + b c)^2 dc, A = (1 + a^2)/2,
// b = (a^2 - 1)/2 (so A + b = a^2): in closed form
// I = (ln(a^2 / A) + A / a^2 - 1) / b^2, or, where |b / A| is small and that
// form cancels, the series I = sum_n (n + 1) / (n + 2) (-b / A)^n / A^2.
HVX_CONTRACT_HD inline double contract_weight_integral(int type, double alpha2) {
    if (type == CONTRACT_DIFFUSE) return kContractPi;
    const double A = 0.5 * (1.0 + alpha2), b = 0.5 * (alpha2 - 1.0), r = -b / A;
    double I;
    if (r > -0.125 && r < 0.125) {
        I = 0.0;
        double p = 1.0;
        for (int n = 0; n < 40; ++n, p *= r) I += (double)(n + 1) / (double)(n + 2) * p;
        I /= A * A;
    } else {
        I = (log(alpha2 / A) + A / alpha2 - 1.0) / (b * b);
    }
    return 2.0 * alpha2 * I;
}

// ------------------------------------------- bright-texel split (optional)
// Level 0 is split per channel into min(L, T) + (L - T)+ at a threshold T
// that at most K texels reach. The contraction filters the clamped part (its
// box levels rebuilt from the clamped level 0); the excess of those texels is
// summed exactly over their level-0 texel centres and added to every output
// texel of the non-delta mips:
//     hot(o) = sum_k w(n_o . l_k) dw_k (L_k - T)+ / contract_weight_integral.
// Delta mips sample the unsplit level 0. T: texels with a positive largest
// channel m are binned by the float pattern of m shifted right by
// kHotBinShift (monotone in m, bins 2^(1/8) wide); with c(b) the texels in
// bins >= b, T is the lower edge of the lowest bin b with c(b) <= K, and a
// texel is hot when m >= T (m > 0). When the brightest occupied bin alone
// holds more than K texels, T is the edge above it and nothing is hot.
// Inputs must be finite.
constexpr uint32_t kHotBinShift = 20;
constexpr uint32_t kHotBins = 2048;  // patterns of positive floats up to +inf (0x7f800000 >> 20 = 2040)
constexpr uint32_t kHotMax = 4096;   // largest K

HVX_CONTRACT_HD inline uint32_t contract_hot_bin(uint32_t float_bits) { return float_bits >> kHotBinShift; }
// T as a float pattern from the first qualifying bin (past the +inf bin: +inf)
HVX_CONTRACT_HD inline uint32_t contract_hot_threshold_bits(uint32_t first_bin) {
    return first_bin > 2040u ? 0x7f800000u : first_bin << kHotBinShift;
}

// Per-mip contraction levels for an input cube of face S (a power of two)
// and an output chain of face So (a power of two) with `mips` levels. The
// roughness of mip m is m / (mips - 1) (PipelineCubeFilter.cpp:305-311).
// Returns the number of levels written (= mips) or 0 on invalid sizes.
inline uint32_t contract_levels(int type, uint32_t S, uint32_t So, uint32_t mips, ContractLevel* out) {
    if (S == 0 || So == 0 || mips == 0 || mips > kContractMaxMips || (S & (S - 1)) || (So & (So - 1))) return 0;
    for (uint32_t m = 0; m < mips; ++m) {
        ContractLevel& L = out[m];
        L = ContractLevel{CONTRACT_FULL, So >> m ? So >> m : 1u, 0, 0, 0.0, 0.0, 0.0};
        if (type == CONTRACT_DIFFUSE) {
            L.Si = S < 32 ? S : 32;
            continue;
        }
        const double r = mips > 1 ? (double)m / (double)(mips - 1) : 0.0;
        const double alpha = r * r;
        L.alpha2 = alpha * alpha;
        if (alpha == 0.0) {
            L.kind = CONTRACT_DELTA;
            L.Si = S;
            continue;
        }
        // input texels of about half the lobe's half width (1.29 alpha)
        uint32_t Si = contract_pow2_at_least((kContractPi / 2.0) / (0.65 * alpha));
        Si = Si < kContractMinIn ? kContractMinIn : Si;
        Si = Si > S ? S : Si;
        L.Si = Si;
        const double full = (double)L.So * L.So * 6.0 * Si * Si;
        if (full <= kContractFullBudget || L.So < kContractTile || Si < 64) continue;
        uint32_t Sf = Si / 4 < 32 ? 32 : Si / 4;
        Sf = Sf > Si / 2 ? Si / 2 : Sf;
        L.kind = CONTRACT_SPLIT;
        L.Sf = Sf;
        // the window ends 20 mean near-level texels out (10 to 20 far texels
        // at Sf = Si / 4): the far level then only sees the tail past it
        const double di = (kContractPi / 2.0) / (double)Si;  // mean angular texel size of the near level
        L.t0 = 10.0 * di;
        L.t1 = 20.0 * di;
    }
    return mips;
}

}  // namespace hvx
