// shade_kernels.hip — clustered Cook-Torrance + split-sum IBL shading of SoA
// fragments on gfx950.
//
// Reference: HelloVulkan/Shaders/ClusteredForward/Scene.frag:60-149 (cluster
// lookup, per-light loop, ambient), ClusteredForward/Radiance.glsl:2-56,
// PBR/PBRHeader.glsl:16-64, PBR/Ambient.glsl:1-29.
//
// Scheduling. Fragments are first binned by cluster (bin_kernel): one
// workgroup per screen tile counting-sorts the tile's pixels by depth slice,
// so the fragment order lists each cluster's fragments together. A shading
// workgroup takes 64 consecutive entries of that order (256 in the learned
// variant) and hands them to lanes in order, so the lanes of a wave mostly share one light list: their
// list entries and light records are wave-uniform and come through the
// scalar cache (with depth drawn independently per pixel, neighbouring pixels
// sit in unrelated clusters, and per-lane gathers of light records bound the
// loop by the L1 tag rate rather than by arithmetic). Grids too deep to bin
// are counting-sorted by light count within each workgroup instead. Each
// fragment is still shaded start to finish by one lane, summing its lights in
// list order, so the order of the fragments changes no result in the exact
// variant (in the fast one it decides which waves are all rough enough for
// the hardware reciprocal square roots: a few ulp). The RGB is
// written at the fragments' binned positions (coalesced) and unbin_kernel
// moves it to pixel order through an LDS image of each tile.
//
// Arithmetic, two variants:
//  exact  every expression as the FP32 oracle evaluates it (IEEE operations,
//         contraction off, correctly rounded divide and square root, the
//         shared log2/exp2/pow) — bit-identical to the oracle.
//  fast   (default) the light vector L, half vector H and N.H stay exact, and
//         so does D's denominator N.H^2 (a^2-1) + 1, because GGX's
//         D = a^2 / (pi (N.H^2 (a^2-1) + 1)^2) amplifies a relative error in
//         N.H by up to 4 (1-a^2) / a^2 (at N.H = 1) and an absolute error in
//         that sum by up to 2 / a^2 (a^2 ~ 6e-6 at roughness 0.05);
//         the remaining per-light terms (N.L, H.V, attenuation, D, G,
//         Fresnel, BRDF combination) use fused multiply-add, hardware
//         reciprocal and hardware exp2/log2, reading a per-call compact light
//         table (position + 1/radius, colour * intensity), and the BRDF's
//         per-fragment constants are applied once to three per-channel sums
//         instead of to every light. Every summed term is non-negative, so
//         the sums keep the terms' ~1e-6 relative accuracy.
#include <hip/hip_runtime.h>

#include "hvx_device.h"
#include "hvx_internal.h"
#include "mlp_device.h"

// build-time tuning switches (defaults are the tuned configuration)
#ifndef HVX_SHADE_XCD_REMAP  // on: 22% less fabric traffic, 1.4% faster (3840x2160, 65,536 lights)
#define HVX_SHADE_XCD_REMAP 1
#endif
#ifndef HVX_SHADE_WAVES_PER_EU  // 5: the plain kernel fits 96 VGPRs, 3.5% faster than 4
#define HVX_SHADE_WAVES_PER_EU 5
#endif
// The learned kernel's workgroups hold the network's weights in LDS (37.9 KB
// of 160 KB per CU), so at most four of its 4-wave workgroups share a CU: the
// register budget is set for those 4 waves per SIMD (an occupancy the LDS
// cannot reach leaves the compiler no register limit at all).
constexpr int kShadeWavesLearned = 4;
#ifndef HVX_SHADE_SKIP  // waves skip lights out of range of all their lanes
#define HVX_SHADE_SKIP 1
#endif

namespace hvx {
namespace {

struct ShadeArgs {
    const float* depth;
    const float* nrm[3];
    const float* alb[3];
    const float* rough;
    const uint8_t* metal;
    uint32_t W, H;
    float ivp[16];
    float cam[3];
    float nearv, farv, scale, bias, scr_x, scr_y;
    uint32_t X, Y, Z;
    hvx_pbr_pc pc;
    const hvx_light* lights;
    const float4* lrec;   // fast: per light (x, y, z, 1/radius), ((r, g, b) * lightIntensity, 0)
    const hvx_light_cell* cells;
    const uint32_t* indices;
    const float4* spec;
    uint32_t spec_size, spec_mips;
    const float4* diff;
    uint32_t diff_size;
    const float2* lut;
    uint32_t lut_w, lut_h;
    float* out;
    // learned prefiltered radiance (shade_kernel<..., true>)
    MlpView mlp;
    const float* cell_err;  // [lod_bins][6][mask_face][mask_face]: the network's relative error per cell
    const float4* cell_lo;  // nullable, same layout: an estimate from below of the network's output per cell (RGB, pad)
    uint32_t mask_face, lod_bins;
    float mlp_tol;          // bound on the relative radiance error the network may cause
    uint32_t* served;       // fragments that used the network (nullable)
    // fragment order grouped by cluster (nullable: identity order) and the
    // tile bounds it is built from: X + 1 column starts, then Y + 1 row starts
    uint32_t* order;
    uint32_t* tiles;
    // binned output (nullable): RGB planes in fragment-order positions, which
    // unbin_kernel moves to pixel order
    float* out_binned;
    uint32_t unbin_cap;  // floats of unbin_kernel's LDS tile image
    uint32_t rclasses;   // binning sub-key classes per slice: 2 splits by roughness, else 1
};

// Fragments at least this rough take hardware-rsqrt light and half vectors
// (fast variant, waves whose every lane qualifies): d ln D / d(N.H) is
// 4 c (1-a^2) / (1 - c^2 (1-a^2)) at N.H = c, at most 4 (1-a^2) / a^2 <= 12.7
// there (a = roughness^2 >= 0.49), so an N.H error of a few ulp (<= 5e-7
// relative) moves D by at most 6.3e-6 relative, within the 1e-5 bound.
constexpr float kRoughFastH = 0.7f;

// Squared distance from which a light's attenuation max(1 - s / r, 0) is 0
// for certain: r^2 (1 + 2^-16), so that s, computed from the squared
// distance dd >= it with a few ulp of error, still satisfies s * (1/r) >= 1.
constexpr float kOutOfRangeMargin = 1.0f + 1.0f / 65536.0f;

__global__ __launch_bounds__(256) void shade_light_prep_kernel(const hvx_light* __restrict__ lights, uint32_t n,
                                                               float intensity, float4* __restrict__ lrec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const hvx_light& l = lights[i];
    lrec[2 * i] = make_float4(l.position[0], l.position[1], l.position[2], 1.0f / l.radius);
    lrec[2 * i + 1] = make_float4(l.color[0] * intensity, l.color[1] * intensity, l.color[2] * intensity,
                                  (l.radius * l.radius) * kOutOfRangeMargin);
}

__device__ __forceinline__ f3 sample_spec_lod(const ShadeArgs& a, f3 dir, float lod) {
    const float q = (float)(a.spec_mips - 1);
    const float lam = gclamp(lod, 0.0f, q);
    const float fl = floorf_(lam);
    const uint32_t l0 = (uint32_t)fl;
    const uint32_t l1 = min(l0 + 1, a.spec_mips - 1);
    const float d = lam - fl;
    const int s0 = (int)max(a.spec_size >> l0, 1u), s1 = (int)max(a.spec_size >> l1, 1u);
    const f3 c0 = sample_cube(a.spec + cube_level_offset_floats(a.spec_size, l0) / 4, s0, dir);
    const f3 c1 = sample_cube(a.spec + cube_level_offset_floats(a.spec_size, l1) / 4, s1, dir);
    return {(1.0f - d) * c0.x + d * c1.x, (1.0f - d) * c0.y + d * c1.y, (1.0f - d) * c0.z + d * c1.z};
}

// Scene.frag:55-78: LinearDepth, log2 slice, tile -> cluster index. The tile
// column and row are non-decreasing in the pixel column and row, so a tile is
// a rectangle of pixels.
__device__ __forceinline__ uint32_t tile_x(const ShadeArgs& a, uint32_t px) {
    const float tsx = a.scr_x / (float)a.X;
    return min((uint32_t)(((float)px + 0.5f) / tsx), a.X - 1);
}
__device__ __forceinline__ uint32_t tile_y(const ShadeArgs& a, uint32_t py) {
    const float tsy = a.scr_y / (float)a.Y;
    return min((uint32_t)(((float)py + 0.5f) / tsy), a.Y - 1);
}
__device__ __forceinline__ uint32_t slice_of(const ShadeArgs& a, uint32_t idx) {
    const float z = a.depth[idx];
    const float lin = (a.nearv * a.farv) / (a.farv + z * (a.nearv - a.farv));
    const float zf = gmax(log2f_(lin) * a.scale + a.bias, 0.0f);
    const uint32_t zi = (uint32_t)zf;
    return zi >= a.Z ? a.Z - 1 : zi;
}
__device__ __forceinline__ uint32_t cluster_of(const ShadeArgs& a, uint32_t idx) {
    return tile_x(a, idx % a.W) + tile_y(a, idx / a.W) * a.X + slice_of(a, idx) * a.X * a.Y;
}

__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }

// Reflection vector and specular level of detail of a fragment, evaluated
// exactly as shade_one does (Ambient.glsl:17-22).
__device__ __forceinline__ f3 frag_reflect(const ShadeArgs& a, uint32_t idx, float* lod) {
    const uint32_t px = idx % a.W, py = idx / a.W;
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    const float z = a.depth[idx];
    float wh[4];
    matvec4(a.ivp, (fx / (float)a.W) * 2.0f - 1.0f, (fy / (float)a.H) * 2.0f - 1.0f, z, 1.0f, wh);
    const f3 wp = {wh[0] / wh[3], wh[1] / wh[3], wh[2] / wh[3]};
    const f3 N = {a.nrm[0][idx], a.nrm[1][idx], a.nrm[2][idx]};
    const f3 V = normalize3(sub3(make3(a.cam[0], a.cam[1], a.cam[2]), wp));
    const f3 I = {-V.x, -V.y, -V.z};
    const float dNI = dot3(N, I);
    *lod = a.rough[idx] * a.pc.max_reflection_lod;
    return sub3(I, mul3(N, 2.0f * dNI));
}

// The cell of (R, lod) in the network's error grid: the cube texel of R at
// the cell-grid resolution (face selection as for sampling) and lod's bin.
__device__ __forceinline__ size_t mlp_cell_index(const ShadeArgs& a, f3 R, float lod) {
    const CubeCoord cc = cube_select(R);
    const uint32_t mf = a.mask_face;
    const uint32_t x = min((uint32_t)(cc.s * (float)mf), mf - 1), y = min((uint32_t)(cc.t * (float)mf), mf - 1);
    const uint32_t b = min((uint32_t)(lod / a.mlp.lod_max * (float)a.lod_bins), a.lod_bins - 1);
    return ((size_t)(b * 6 + (uint32_t)cc.face) * mf + y) * mf + x;
}
// The network's measured relative error for (R, lod).
__device__ __forceinline__ float mlp_cell_error(const ShadeArgs& a, f3 R, float lod) {
    return a.cell_err[mlp_cell_index(a, R, lod)];
}

// What a learned-shading lane keeps from the light loop to the final write:
// kD * diffuse and the prefiltered radiance's weight per channel, the direct
// light, the reflection vector and the roughness.
struct ShadeTail {
    float kdd[3], sw[3], Lo[3];
    f3 R;
    float rough;
};

// TAIL = false: shades fragment idx and writes its RGB with the map's
// prefiltered radiance. TAIL = true: stops before the specular ambient term
// and leaves the rest in *tail (shade_tail_write finishes it).
template <bool EXACT, bool TAIL = false>
__device__ __forceinline__ void shade_one(const ShadeArgs& a, uint32_t idx, uint32_t oidx, uint32_t off,
                                          uint32_t count, ShadeTail* tail = nullptr) {
    const float PI = kPiF;
    const uint32_t nfrag = a.W * a.H;
    // the RGB goes to pixel idx of out, or to position oidx of the binned planes
    float* const dst = a.out_binned ? a.out_binned : a.out;
    const uint32_t di = a.out_binned ? oidx : idx;
    const uint32_t px = idx % a.W, py = idx / a.W;
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    const float z = a.depth[idx];
    float wh[4];
    matvec4(a.ivp, (fx / (float)a.W) * 2.0f - 1.0f, (fy / (float)a.H) * 2.0f - 1.0f, z, 1.0f, wh);
    const f3 wp = {wh[0] / wh[3], wh[1] / wh[3], wh[2] / wh[3]};

    const float alb[3] = {a.alb[0][idx], a.alb[1][idx], a.alb[2][idx]};
    const float metallic = (float)a.metal[idx] / 255.0f;
    const float rough = a.rough[idx];
    const float rr = rough + 1.0f;
    const float alphaRough = (rr * rr) / 8.0f;  // AlphaDirectLighting
    const f3 N = {a.nrm[0][idx], a.nrm[1][idx], a.nrm[2][idx]};
    const f3 V = normalize3(sub3(make3(a.cam[0], a.cam[1], a.cam[2]), wp));
    const float NoV = gmax(dot3(N, V), 0.0f);
    const float base = a.pc.base_reflectivity;
    const float oneMinusMetal = 1.0f - metallic;
    float F0[3], Lo[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        F0[c] = base * (1.0f - metallic) + alb[c] * metallic;
        Lo[c] = alb[c] * a.pc.albedo_multiplier;
    }
    const float alpha = rough * rough;
    const float alpha2 = alpha * alpha;

    if constexpr (EXACT) {
        const float GV = NoV / (NoV * (1.0f - alphaRough) + alphaRough);
        for (uint32_t k = 0; k < count; ++k) {
            const hvx_light& l = a.lights[a.indices[off + k]];
            const float4 lpos = *reinterpret_cast<const float4*>(l.position);
            const float4 lcol = *reinterpret_cast<const float4*>(l.color);
            const float lrad = l.radius;
            const f3 d = sub3(make3(lpos.x, lpos.y, lpos.z), wp);
            const f3 L = normalize3(d);
            const f3 Hh = normalize3(add3(V, L));
            const float NoH = gmax(dot3(N, Hh), 0.0f);
            const float NoL = gmax(dot3(N, L), 0.0f);
            const float HoV = gmax(dot3(Hh, V), 0.0f);
            const float dist = length3(d);
            const float att = gmax(1.0f - dist / lrad, 0.0f) / powf_(dist, a.pc.light_falloff);
            const float rad[3] = {(lcol.x * att) * a.pc.light_intensity, (lcol.y * att) * a.pc.light_intensity,
                                  (lcol.z * att) * a.pc.light_intensity};
            const float NoH2 = NoH * NoH;
            float den = NoH2 * (alpha2 - 1.0f) + 1.0f;
            den = (PI * den) * den;
            const float D = alpha2 / den;
            const float GL = NoL / (NoL * (1.0f - alphaRough) + alphaRough);
            const float G = GL * GV;
            const float p5 = pow5f(gclamp(1.0f - HoV, 0.0f, 1.0f));
            const float denom = ((4.0f * NoV) * NoL) + 0.0001f;
            const float DG = D * G;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float F = F0[c] + (1.0f - F0[c]) * p5;
                const float spec = (DG * F) / denom;
                const float kD = (1.0f - F) * oneMinusMetal;
                const float diffuse = (kD * alb[c]) / PI;
                Lo[c] = Lo[c] + ((diffuse + spec) * rad[c]) * NoL;
            }
        }
    } else {
        const float omk = 1.0f - alphaRough;
        const float GV = NoV * rcp_fast(__builtin_fmaf(NoV, omk, alphaRough));
        const float a2m1 = alpha2 - 1.0f;
        const float nov4 = 4.0f * NoV;
        const float negf = -a.pc.light_falloff;
        const float aGVpi = (alpha2 * GV) * (1.0f / kPiF);
        // A light's BRDF term per channel, (1 - F) q + F DGr with
        // F = F0 + (1 - F0) p5 and q = (1 - metallic) albedo / pi, is
        // q (1 - F0) (1 - p5) + F0 DGr + (1 - F0) p5 DGr: three per-light
        // weights times per-fragment constants. The loop sums colour times
        // each weight (S1, S2, S3, all terms non-negative) and the constants
        // are applied once after it.
        float S1[3] = {0.0f, 0.0f, 0.0f}, S2[3] = {0.0f, 0.0f, 0.0f}, S3[3] = {0.0f, 0.0f, 0.0f};
        // wave-uniform: every lane rough enough for the fast light vectors
        const bool fastH = __all(rough >= kRoughFastH);
        // one light's term, added to Lo
        auto light = [&](const float4 A, const float4 B) {
            const f3 d = {A.x - wp.x, A.y - wp.y, A.z - wp.z};
            float s, NoH;
            f3 L, Hh;
            // A light out of range of every lane (squared distance at least
            // B.w = r^2 (1 + 2^-16)) has attenuation exactly 0 in each: its
            // terms add nothing, so the wave skips it.
            if (fastH) {
                // L and H from the hardware reciprocal square root (a few ulp)
                const float dd = __builtin_fmaf(d.z, d.z, __builtin_fmaf(d.y, d.y, d.x * d.x));
                if (HVX_SHADE_SKIP && !__any(dd < B.w)) return;
                const float r = __builtin_amdgcn_rsqf(dd);
                s = dd * r;
                L = mul3(d, r);
                const f3 h = add3(V, L);
                Hh = mul3(h, __builtin_amdgcn_rsqf(__builtin_fmaf(h.z, h.z, __builtin_fmaf(h.y, h.y, h.x * h.x))));
                NoH = fmaxf(__builtin_fmaf(N.z, Hh.z, __builtin_fmaf(N.y, Hh.y, N.x * Hh.x)), 0.0f);
            } else {
                // exact: L, H and N.H (max with 0 only differs in the sign of
                // a zero, which no later term can observe)
                const float dd = dot3(d, d);
                if (HVX_SHADE_SKIP && !__any(dd < B.w)) return;
                s = sqrt_rn_normal(dd);
                L = mul3(d, rcp_rn_normal(s));
                const f3 h = add3(V, L);
                Hh = mul3(h, rcp_rn_normal(sqrt_rn_normal(dot3(h, h))));
                NoH = fmaxf(dot3(N, Hh), 0.0f);
            }
            const float NoL = fmaxf(__builtin_fmaf(N.z, L.z, __builtin_fmaf(N.y, L.y, N.x * L.x)), 0.0f);
            const float HoV = __builtin_fmaf(Hh.z, V.z, __builtin_fmaf(Hh.y, V.y, Hh.x * V.x));
            // fast: N.L and H.V fused (they enter only smooth factors), then
            // attenuation, D, G, Fresnel and the BRDF sum
            const float att = fmaxf(__builtin_fmaf(-s, A.w, 1.0f), 0.0f) *
                              __builtin_amdgcn_exp2f(negf * __builtin_amdgcn_logf(s));
            // den = NoH^2 (a^2 - 1) + 1 exactly as the oracle rounds it: near
            // N.H = 1 it cancels down to ~a^2, so one rounding more or less
            // moves D by up to 1e-7 / a^2
            const float den = (NoH * NoH) * a2m1 + 1.0f;
            // D * G_L * G_V / (4 NoV NoL + 1e-4) with one reciprocal of the
            // product of the three denominators (each bounded away from 0);
            // a^2 G_V / pi is a per-fragment constant
            const float gden = __builtin_fmaf(NoL, omk, alphaRough);
            const float sden = __builtin_fmaf(nov4, NoL, 0.0001f);
            const float x = 1.0f - HoV;
            const float x2 = x * x;
            const float p5 = (x2 * x2) * x;
            const float DGr = (aGVpi * NoL) * rcp_fast(((den * den) * gden) * sden);
            const float w = att * NoL;
            const float w1 = __builtin_fmaf(-w, p5, w);  // w (1 - p5)
            const float w2 = w * DGr;
            const float w3 = w2 * p5;
            const float col[3] = {B.x, B.y, B.z};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                S1[c] = __builtin_fmaf(col[c], w1, S1[c]);
                S2[c] = __builtin_fmaf(col[c], w2, S2[c]);
                S3[c] = __builtin_fmaf(col[c], w3, S3[c]);
            }
        };
        const uint32_t off0 = __builtin_amdgcn_readfirstlane(off), cnt0 = __builtin_amdgcn_readfirstlane(count);
        if (__all((off == off0) & (count == cnt0))) {
            // Every active lane walks the same list (the usual case with
            // cluster-binned fragments): its entries and light records are
            // wave-uniform and come through the scalar cache into SGPRs.
            // Both tables are read-only for the kernel's lifetime, which the
            // constant address space states.
            typedef const __attribute__((address_space(4))) float cfloat;
            typedef const __attribute__((address_space(4))) uint32_t cuint;
            cfloat* R = (cfloat*)a.lrec;
            cuint* I = (cuint*)(a.indices + off0);
            // one 32-byte light record per scalar load; software-pipelined
            // like the per-lane loop below (records of k+2, k+3 and entries
            // of k+4, k+5 in flight while lights k, k+1 are shaded)
            auto rec = [&](uint32_t li, float4& A, float4& B) {
                cfloat* r = R + 8 * (size_t)li;
                A = make_float4(r[0], r[1], r[2], r[3]);
                B = make_float4(r[4], r[5], r[6], r[7]);
            };
            if (cnt0 > 0) {
                const uint32_t c1 = cnt0 - 1;
                float4 A0n, B0n, A1n, B1n;
                rec(I[0], A0n, B0n);
                rec(I[min(1u, c1)], A1n, B1n);
                uint32_t i2 = I[min(2u, c1)], i3 = I[min(3u, c1)];
                for (uint32_t k = 0; k < cnt0; k += 2) {
                    const float4 A0 = A0n, B0 = B0n, A1 = A1n, B1 = B1n;
                    rec(i2, A0n, B0n);
                    rec(i3, A1n, B1n);
                    i2 = I[min(k + 4, c1)];
                    i3 = I[min(k + 5, c1)];
                    __asm__ volatile("" ::: "memory");
                    light(A0, B0);
                    if (k + 1 < cnt0) light(A1, B1);
                }
            }
        } else {
            // Lanes on different lists: per-lane loads, unrolled by two and
            // software-pipelined (while lights k, k+1 are shaded, the records
            // of k+2, k+3 and the list entries of k+4, k+5 are in flight).
            // Positions past the end of a list re-read its last entry
            // (count >= 1 inside the loop), which keeps the loads
            // unconditional; the empty asm keeps them in this iteration
            // (without it they are sunk to their first use, one iteration
            // later).
            float4 A0n = make_float4(0.f, 0.f, 0.f, 0.f), B0n = A0n, A1n = A0n, B1n = A0n;
            uint32_t i2 = 0, i3 = 0;
            const uint32_t c1 = count > 0 ? count - 1 : 0;
            if (count > 0) {
                const uint32_t l0 = a.indices[off], l1 = a.indices[off + min(1u, c1)];
                A0n = a.lrec[2 * l0];
                B0n = a.lrec[2 * l0 + 1];
                A1n = a.lrec[2 * l1];
                B1n = a.lrec[2 * l1 + 1];
                i2 = a.indices[off + min(2u, c1)];
                i3 = a.indices[off + min(3u, c1)];
            }
            for (uint32_t k = 0; k < count; k += 2) {
                const float4 A0 = A0n, B0 = B0n, A1 = A1n, B1 = B1n;
                A0n = a.lrec[2 * i2];
                B0n = a.lrec[2 * i2 + 1];
                A1n = a.lrec[2 * i3];
                B1n = a.lrec[2 * i3 + 1];
                i2 = a.indices[off + min(k + 4, c1)];
                i3 = a.indices[off + min(k + 5, c1)];
                __asm__ volatile("" ::: "memory");
                light(A0, B0);
                if (k + 1 < count) light(A1, B1);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float om = 1.0f - F0[c];
            const float qom = ((oneMinusMetal * alb[c]) * (1.0f / kPiF)) * om;
            Lo[c] = __builtin_fmaf(qom, S1[c], __builtin_fmaf(F0[c], S2[c], __builtin_fmaf(om, S3[c], Lo[c])));
        }
    }

    // Ambient (PBR/Ambient.glsl:1-29), as the oracle evaluates it
    const float p5v = pow5f(gclamp(1.0f - NoV, 0.0f, 1.0f));
    const f3 irr = sample_cube(a.diff, (int)a.diff_size, N);
    const f3 I = {-V.x, -V.y, -V.z};
    const float dNI = dot3(N, I);
    const f3 R = sub3(I, mul3(N, 2.0f * dNI));
    if constexpr (!TAIL) {  // plain: the map's prefiltered radiance
        const f3 pre = sample_spec_lod(a, R, rough * a.pc.max_reflection_lod);
        const float2 brdf = sample_2d_rg_repeat(a.lut, (int)a.lut_w, (int)a.lut_h, NoV, rough);
        const float irr3[3] = {irr.x, irr.y, irr.z}, pre3[3] = {pre.x, pre.y, pre.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float F = F0[c] + (gmax(1.0f - rough, F0[c]) - F0[c]) * p5v;
            const float kD = (1.0f - F) * oneMinusMetal;
            const float diffuse = irr3[c] * alb[c];
            const float specular = pre3[c] * (F * brdf.x + brdf.y);
            const float ambient = (kD * diffuse + specular) * 1.0f;
            dst[(size_t)c * nfrag + di] = (ambient + 0.0f) + Lo[c];
        }
    } else {
        (void)dst;
        (void)di;
        const float2 brdf = sample_2d_rg_repeat(a.lut, (int)a.lut_w, (int)a.lut_h, NoV, rough);
        const float irr3[3] = {irr.x, irr.y, irr.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float F = F0[c] + (gmax(1.0f - rough, F0[c]) - F0[c]) * p5v;
            const float kD = (1.0f - F) * oneMinusMetal;
            const float diffuse = irr3[c] * alb[c];
            tail->kdd[c] = kD * diffuse;
            tail->sw[c] = F * brdf.x + brdf.y;  // weight of the prefiltered radiance
            tail->Lo[c] = Lo[c];
        }
        tail->R = R;
        tail->rough = rough;
    }
}

// The learned radiance replaces the map sample where the error it can cause
// in the shaded radiance, e/(1-e) * learned * weight per channel (e = the
// cell's measured relative error of the network), stays within tol of the
// radiance without the specular ambient term (a lower bound of the shaded
// value, every term being non-negative).
__device__ __forceinline__ bool learned_serves(const ShadeArgs& a, const ShadeTail& t, float cell_err, f3 learned) {
    if (!(cell_err >= 0.0f && cell_err < 1.0f)) return false;
    const float k = cell_err / (1.0f - cell_err);
    const float l3[3] = {learned.x, learned.y, learned.z};
    bool use = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) use = use && k * l3[c] * t.sw[c] <= a.mlp_tol * (t.kdd[c] + t.Lo[c]);
    return use;
}

// The final write of a learned-shading fragment (the plain expression, with
// `pre` the learned or the map's prefiltered radiance).
__device__ __forceinline__ void shade_tail_write(const ShadeArgs& a, uint32_t idx, uint32_t oidx, const ShadeTail& t,
                                                 f3 pre) {
    const uint32_t nfrag = a.W * a.H;
    float* const dst = a.out_binned ? a.out_binned : a.out;
    const uint32_t di = a.out_binned ? oidx : idx;
    const float pre3[3] = {pre.x, pre.y, pre.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float specular = pre3[c] * t.sw[c];
        const float ambient = (t.kdd[c] + specular) * 1.0f;
        dst[(size_t)c * nfrag + di] = (ambient + 0.0f) + t.Lo[c];
    }
}

// Fragments per workgroup (3840x2160, 65,536 lights): plain shading 64 (2.07 ms; 128:
// 2.09, 256: 2.11, 1024: 2.34); learned shading 256 (2.30 ms; 128: 3.65, 512:
// 2.65), since each of its workgroups first stages the network's weights.
#ifndef HVX_SHADE_CHUNK
#define HVX_SHADE_CHUNK 64
#endif
#ifndef HVX_SHADE_CHUNK_LEARNED
#define HVX_SHADE_CHUNK_LEARNED 256
#endif
template <bool LEARNED> struct ShadeChunk {
    static constexpr int kChunk = LEARNED ? HVX_SHADE_CHUNK_LEARNED : HVX_SHADE_CHUNK;
    static constexpr int kThreads = kChunk < 256 ? kChunk : 256;  // lanes per workgroup
    static constexpr int kRounds = kChunk / kThreads;            // fragments per lane
    static_assert(kChunk % kThreads == 0 && kThreads % 64 == 0, "whole waves, whole rounds");
};
constexpr int kBuckets = 64;   // light-count classes (count / 2, saturating)
constexpr uint32_t kXcds = 8;  // MI355X accelerator complex dies

// ------------------------------------------------------- cluster binning
constexpr uint32_t kBinThreads = 1024;
constexpr uint32_t kBinMaxZ = 1024;  // depth slices a tile's LDS histogram holds

// Tile bounds: tiles[c] = first pixel column of tile column c (c <= X, with
// tiles[X] = W), then the same for rows. One thread per pixel column or row;
// tile_x is non-decreasing, so the thread of the first column of a tile
// writes that tile's start (and the starts of any empty tiles before it).
__global__ __launch_bounds__(256) void tile_bounds_kernel(ShadeArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.W + a.H) return;
    const bool col = i < a.W;
    const uint32_t p = col ? i : i - a.W, n = col ? a.W : a.H, T = col ? a.X : a.Y;
    uint32_t* b = col ? a.tiles : a.tiles + a.X + 1;
    const uint32_t c = col ? tile_x(a, p) : tile_y(a, p);
    const uint32_t first = p == 0 ? 0 : (col ? tile_x(a, p - 1) : tile_y(a, p - 1)) + 1;
    for (uint32_t t = first; t <= c; ++t) b[t] = p;
    if (p == n - 1)
        for (uint32_t t = c + 1; t <= T; ++t) b[t] = n;
}

// One workgroup per screen tile and RGB plane: moves the tile's segment of
// the binned plane to pixel order through an LDS image of the tile, so that
// both the shading kernel's stores (fragment order) and these (tile rows)
// are coalesced. A workgroup per plane rather than per tile: with two
// tile images per CU, 576 tiles (3840x2160) would leave a second round of
// 64 workgroups; 1,728 shorter ones balance across the CUs.
__global__ __launch_bounds__(kBinThreads) void unbin_kernel(ShadeArgs a) {
    extern __shared__ float s_img[];
    const uint32_t tile = blockIdx.x / 3, c = blockIdx.x % 3;
    const uint32_t cx = tile % a.X, cy = tile / a.X, t = threadIdx.x;
    const uint32_t x0 = a.tiles[cx], x1 = a.tiles[cx + 1];
    const uint32_t y0 = a.tiles[a.X + 1 + cy], y1 = a.tiles[a.X + 1 + cy + 1];
    const uint32_t w = x1 - x0, area = w * (y1 - y0), nfrag = a.W * a.H;
    if (area == 0) return;  // uniform over the workgroup
    const uint32_t start = y0 * a.W + (y1 - y0) * x0;
    if (area > a.unbin_cap) {  // larger than the bound the image was sized by: scatter directly
        for (uint32_t i = t; i < area; i += kBinThreads)
            a.out[(size_t)c * nfrag + a.order[start + i]] = a.out_binned[(size_t)c * nfrag + start + i];
        return;
    }
    for (uint32_t i = t; i < area; i += kBinThreads) {
        const uint32_t idx = a.order[start + i];
        s_img[(idx / a.W - y0) * w + (idx % a.W - x0)] = a.out_binned[(size_t)c * nfrag + start + i];
    }
    __syncthreads();
    for (uint32_t i = t; i < area; i += kBinThreads)
        a.out[(size_t)c * nfrag + (y0 + i / w) * a.W + x0 + i % w] = s_img[i];
}

// Binning key of a fragment: its depth slice, then (rclasses = 2) whether it
// is rough enough for the fast light vectors, so a cluster's fragments are
// contiguous and its rough ones fill whole waves.
__device__ __forceinline__ uint32_t bin_of(const ShadeArgs& a, uint32_t idx) {
    const uint32_t z = slice_of(a, idx);
    return a.rclasses == 2 ? 2 * z + (a.rough[idx] >= kRoughFastH ? 1u : 0u) : z;
}

// One workgroup per screen tile: counting sort of the tile's pixels by depth
// slice into the tile's segment of the order. Tiles are laid out row band by
// row band, so tile (cx, cy) starts at y0 * W + (y1 - y0) * x0. Within a
// slice the order is that of the LDS atomics (see the note on the fast
// variant at the top of this file).
__global__ __launch_bounds__(kBinThreads) void bin_kernel(ShadeArgs a) {
    __shared__ uint32_t s_hist[kBinMaxZ];
    __shared__ uint32_t s_wave[kBinThreads / 64];
    const uint32_t cx = blockIdx.x % a.X, cy = blockIdx.x / a.X, t = threadIdx.x;
    const uint32_t x0 = a.tiles[cx], x1 = a.tiles[cx + 1];
    const uint32_t y0 = a.tiles[a.X + 1 + cy], y1 = a.tiles[a.X + 1 + cy + 1];
    const uint32_t w = x1 - x0, area = w * (y1 - y0);
    if (area == 0) return;  // uniform over the workgroup
    const uint32_t start = y0 * a.W + (y1 - y0) * x0;
    for (uint32_t z = t; z < a.Z * a.rclasses; z += kBinThreads) s_hist[z] = 0;
    __syncthreads();
    const uint32_t nb = a.Z * a.rclasses;  // bins: (slice, roughness class)
    // the bins of a thread's first kBinCache pixels stay in registers for
    // the second pass (a 120 x 120 tile takes 15 rounds)
    constexpr uint32_t kBinCache = 16;
    uint32_t cache[kBinCache];
#pragma unroll
    for (uint32_t k = 0; k < kBinCache; ++k) {
        const uint32_t i = t + k * kBinThreads;
        cache[k] = 0;
        if (i < area) {
            const uint32_t idx = (y0 + i / w) * a.W + x0 + i % w;
            cache[k] = bin_of(a, idx);
            atomicAdd(&s_hist[cache[k]], 1u);
        }
    }
    for (uint32_t i = t + kBinCache * kBinThreads; i < area; i += kBinThreads) {
        const uint32_t idx = (y0 + i / w) * a.W + x0 + i % w;
        atomicAdd(&s_hist[bin_of(a, idx)], 1u);
    }
    __syncthreads();
    // exclusive scan of the Z <= kBinThreads slice counts: per wave, then
    // over the wave totals
    const uint32_t v = t < nb ? s_hist[t] : 0u, lane = t & 63, wv = t >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(inc, d, 64);
        if ((int)lane >= d) inc += u;
    }
    if (lane == 63) s_wave[wv] = inc;
    __syncthreads();
    if (t < 64) {
        const uint32_t wt = t < kBinThreads / 64 ? s_wave[t] : 0u;
        uint32_t wi = wt;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const uint32_t u = __shfl_up(wi, d, 64);
            if ((int)t >= d) wi += u;
        }
        if (t < kBinThreads / 64) s_wave[t] = wi - wt;
    }
    __syncthreads();
    if (t < nb) s_hist[t] = start + s_wave[wv] + inc - v;
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kBinCache; ++k) {
        const uint32_t i = t + k * kBinThreads;
        if (i < area) a.order[atomicAdd(&s_hist[cache[k]], 1u)] = (y0 + i / w) * a.W + x0 + i % w;
    }
    for (uint32_t i = t + kBinCache * kBinThreads; i < area; i += kBinThreads) {
        const uint32_t idx = (y0 + i / w) * a.W + x0 + i % w;
        a.order[atomicAdd(&s_hist[bin_of(a, idx)], 1u)] = idx;
    }
}


template <bool EXACT, bool LEARNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LEARNED ? kShadeWavesLearned : HVX_SHADE_WAVES_PER_EU, 8))) void shade_kernel(
    ShadeArgs a) {
    constexpr int kChunk = ShadeChunk<LEARNED>::kChunk, kThreads = ShadeChunk<LEARNED>::kThreads,
                  kRounds = ShadeChunk<LEARNED>::kRounds;
    __shared__ uint32_t s_idx[kChunk];
    __shared__ uint32_t s_off[kChunk];
    __shared__ uint16_t s_cnt[kChunk];
    __shared__ uint16_t s_order[kChunk];
    __shared__ uint32_t s_hist[kBuckets];
    const uint32_t nfrag = a.W * a.H;
    // XCD-aware chunk order: workgroups are dealt to the 8 XCDs round-robin,
    // so the remap gives XCD k a contiguous run of chunks. With binned
    // fragments a run is a run of screen tiles, so the workgroups that gather
    // one tile's attributes (each 128-byte line holds pixels of many slices,
    // shaded by different workgroups) share one L2 instead of eight.
    const uint32_t nblk = gridDim.x, per = (nblk + kXcds - 1) / kXcds;
    const uint32_t xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const uint32_t full = nblk % kXcds == 0 ? kXcds : nblk % kXcds;  // XCDs with `per` blocks
#if HVX_SHADE_XCD_REMAP
    const uint32_t chunk = xcd < full ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
#else
    const uint32_t chunk = blockIdx.x;
    (void)xcd; (void)slot; (void)full;
#endif
    const uint32_t base = chunk * kChunk;
    const uint32_t t = threadIdx.x;
    if (t < kBuckets) s_hist[t] = 0;
    __syncthreads();
    uint32_t bucket[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const uint32_t local = r * kThreads + t, pos = base + local;
        const uint32_t idx = pos >= nfrag ? nfrag : a.order ? a.order[pos] : pos;
        s_idx[local] = idx;
        uint32_t off = 0, cnt = 0;
        if (idx < nfrag) {
            const hvx_light_cell cell = a.cells[cluster_of(a, idx)];
            off = cell.offset;
            cnt = min(cell.count, 65535u);
        }
        s_off[local] = off;
        s_cnt[local] = (uint16_t)cnt;
        bucket[r] = min(cnt >> 1, (uint32_t)kBuckets - 1);
        if (a.order) s_order[local] = (uint16_t)local;
        else atomicAdd(&s_hist[bucket[r]], 1u);
    }
    __syncthreads();
    // Unbinned fragments: counting sort by light count, so the lanes of a
    // wave walk lists of nearly equal length. Binned fragments keep their
    // order, which groups each cluster's fragments.
    if (!a.order) {
        if (t < 64) {  // exclusive scan of the bucket sizes (one wave)
            const uint32_t v = s_hist[t];
            uint32_t inc = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t u = __shfl_up(inc, d, 64);
                if ((int)t >= d) inc += u;
            }
            s_hist[t] = inc - v;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const uint32_t slot = atomicAdd(&s_hist[bucket[r]], 1u);
            s_order[slot] = (uint16_t)(r * kThreads + t);
        }
        __syncthreads();
    }
    if constexpr (LEARNED) {
        // Each lane first shades its fragment up to the specular ambient term.
        // The network is then evaluated only for candidate lanes, those whose
        // cell could serve even at the cell's lower estimate of the network's
        // output (a larger output only makes serving harder), packed into
        // 16-sample tiles; tile T holds candidates 16 T .. 16 T + 15 of the
        // wave, and the results return to their lanes by shuffles. The final
        // decision is made on the network's actual output, so the lower
        // estimate only decides where the network is evaluated.
        __shared__ MlpLds s_mlp;
        __shared__ uint32_t s_served;
        __shared__ uint8_t s_cand[kThreads / 64][64];
        mlp_stage(s_mlp, a.mlp);
        if (t == 0) s_served = 0;
        __syncthreads();
        const uint32_t lane = t & 63, i = lane & 15, wv = t >> 6;
        uint32_t mine = 0;
#pragma unroll 1
        for (int r = 0; r < kRounds; ++r) {
            const uint32_t local = s_order[r * kThreads + t];
            const uint32_t idx = s_idx[local];
            const bool live = idx < nfrag;
            float x[4] = {0.0f, 0.0f, 1.0f, 0.0f};
            float err = -1.0f;
            f3 lo = {0.0f, 0.0f, 0.0f};
            ShadeTail tl;
            if (live) {
                float lod;
                const f3 R = frag_reflect(a, idx, &lod);
                x[0] = R.x;
                x[1] = R.y;
                x[2] = R.z;
                x[3] = lod;
                const size_t cell = mlp_cell_index(a, R, lod);
                err = a.cell_err[cell];
                if (a.cell_lo) {
                    const float4 l = a.cell_lo[cell];
                    lo = f3{l.x, l.y, l.z};
                }
                shade_one<EXACT, true>(a, idx, base + local, s_off[local], s_cnt[local], &tl);
            }
            const bool cand = live && learned_serves(a, tl, err, lo);
            const unsigned long long cm = __ballot(cand);
            const uint32_t nc = (uint32_t)__popcll(cm);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(cm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)cm, 0u));
            if (cand) s_cand[wv][rank] = (uint8_t)lane;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            f3 learned = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
            for (uint32_t T = 0; 16 * T < nc; ++T) {
                const uint32_t q = 16 * T + i;
                const int src = (int)s_cand[wv][q < nc ? q : 16 * T];
                const float xt[4] = {__shfl(x[0], src, 64), __shfl(x[1], src, 64), __shfl(x[2], src, 64),
                                     __shfl(x[3], src, 64)};
                const float y = mlp_forward_tile(s_mlp, xt, a.mlp.lod_max);  // lane (c, j): channel c of sample j
                const int j = (int)(rank & 15u);
                const float c0 = __shfl(y, j, 64), c1 = __shfl(y, 16 + j, 64), c2 = __shfl(y, 32 + j, 64);
                if (cand && (rank >> 4) == T) learned = f3{c0, c1, c2};
            }
            if (live) {
                const bool use = cand && learned_serves(a, tl, err, learned);
                const f3 pre = use ? learned : sample_spec_lod(a, tl.R, tl.rough * a.pc.max_reflection_lod);
                shade_tail_write(a, idx, base + local, tl, pre);
                mine += use ? 1u : 0u;
            }
        }
        if (a.served) {
            if (mine) atomicAdd(&s_served, mine);
            __syncthreads();
            if (t == 0 && s_served) atomicAdd(a.served, s_served);
        }
    } else {
#pragma unroll 1
        for (int r = 0; r < kRounds; ++r) {
            const uint32_t local = s_order[r * kThreads + t];
            const uint32_t idx = s_idx[local];
            if (idx < nfrag) shade_one<EXACT>(a, idx, base + local, s_off[local], s_cnt[local]);
        }
    }
}

}  // namespace

bool shade_binned(const hvx_cf_ubo& ubo) { return ubo.slice_count_z <= kBinMaxZ; }

// LDS bytes of unbin_kernel's tile image: a bound on the largest tile (a tile
// spans at most ceil(W / X) + 1 columns and ceil(H / Y) + 1 rows)
static size_t unbin_lds_bytes(const hvx_frag_soa& fr, const hvx_cf_ubo& ubo) {
    const size_t tw = (fr.width + ubo.slice_count_x - 1) / ubo.slice_count_x + 1;
    const size_t th = (fr.height + ubo.slice_count_y - 1) / ubo.slice_count_y + 1;
    return tw * th * sizeof(float);
}
constexpr size_t kUnbinMaxLds = 64 * 1024;

size_t shade_bin_scratch_bytes(const hvx_frag_soa& fr, const hvx_cf_ubo& ubo) {
    const size_t F = (size_t)fr.width * fr.height;
    const size_t out = unbin_lds_bytes(fr, ubo) <= kUnbinMaxLds ? 3 * F : 0;  // binned RGB planes
    return (F + ubo.slice_count_x + ubo.slice_count_y + 2 + out) * sizeof(uint32_t);
}

hipError_t launch_shade(const hvx_frag_soa& fr, const hvx_camera_ubo& cam, const hvx_cf_ubo& ubo,
                        const hvx_pbr_pc& pc, const hvx_light* d_lights, uint32_t n_lights,
                        const hvx_light_cell* d_cells, const uint32_t* d_indices, const hvx_ibl_maps& ibl,
                        float* d_rgb, bool exact, float4* light_scratch, uint32_t* bin_scratch,
                        const LearnedShade* learned, hipStream_t s, hipEvent_t ev_begin, hipEvent_t ev_end) {
    ShadeArgs a{};
    a.depth = fr.depth;
    for (int c = 0; c < 3; ++c) {
        a.nrm[c] = fr.normal[c];
        a.alb[c] = fr.albedo[c];
    }
    a.rough = fr.roughness;
    a.metal = fr.metallic;
    a.W = fr.width;
    a.H = fr.height;
    for (int k = 0; k < 16; ++k) a.ivp[k] = fr.inv_view_proj[k];
    for (int k = 0; k < 3; ++k) a.cam[k] = cam.position[k];
    a.nearv = ubo.camera_near;
    a.farv = ubo.camera_far;
    a.scale = ubo.slice_scaling;
    a.bias = ubo.slice_bias;
    a.scr_x = ubo.screen_size[0];
    a.scr_y = ubo.screen_size[1];
    a.X = ubo.slice_count_x;
    a.Y = ubo.slice_count_y;
    a.Z = ubo.slice_count_z;
    a.pc = pc;
    a.lights = d_lights;
    a.lrec = light_scratch;
    a.cells = d_cells;
    a.indices = d_indices;
    a.spec = reinterpret_cast<const float4*>(ibl.specular);
    a.spec_size = ibl.specular_size;
    a.spec_mips = ibl.specular_mips;
    a.diff = reinterpret_cast<const float4*>(ibl.diffuse);
    a.diff_size = ibl.diffuse_size;
    a.lut = reinterpret_cast<const float2*>(ibl.brdf_lut);
    a.lut_w = ibl.lut_width;
    a.lut_h = ibl.lut_height;
    a.out = d_rgb;
    if (learned) {
        a.mlp = learned->mlp;
        a.cell_err = learned->cell_err;
        a.cell_lo = learned->cell_lo;
        a.mask_face = learned->mask_face;
        a.lod_bins = learned->lod_bins;
        a.mlp_tol = learned->tol;
        a.served = learned->served;
    }
    const uint32_t F = fr.width * fr.height;
    const int chunk = learned ? ShadeChunk<true>::kChunk : ShadeChunk<false>::kChunk;
    const dim3 grid((F + chunk - 1) / chunk), block(learned ? ShadeChunk<true>::kThreads : ShadeChunk<false>::kThreads);
    const size_t unbin_lds = unbin_lds_bytes(fr, ubo);
    const bool binned = bin_scratch && shade_binned(ubo), staged = binned && unbin_lds <= kUnbinMaxLds;
    if (binned) {
        a.order = bin_scratch;
        a.tiles = bin_scratch + F;
        a.rclasses = !exact && 2 * a.Z <= kBinMaxZ ? 2u : 1u;
        if (staged) {
            a.out_binned = reinterpret_cast<float*>(bin_scratch + F + a.X + a.Y + 2);
            a.unbin_cap = (uint32_t)(unbin_lds / sizeof(float));
        }
        hipLaunchKernelGGL(tile_bounds_kernel, dim3((fr.width + fr.height + 255) / 256), dim3(256), 0, s, a);
        hipLaunchKernelGGL(bin_kernel, dim3(a.X * a.Y), dim3(kBinThreads), 0, s, a);
    }
    if (!exact && n_lights > 0)
        hipLaunchKernelGGL(shade_light_prep_kernel, dim3((n_lights + 255) / 256), dim3(256), 0, s, d_lights, n_lights,
                           pc.light_intensity, light_scratch);
    if (ev_begin) (void)hipEventRecord(ev_begin, s);
    if (exact) {
        if (learned) hipLaunchKernelGGL((shade_kernel<true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((shade_kernel<true, false>), grid, block, 0, s, a);
    } else {
        if (learned) hipLaunchKernelGGL((shade_kernel<false, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((shade_kernel<false, false>), grid, block, 0, s, a);
    }
    if (ev_end) (void)hipEventRecord(ev_end, s);
    if (staged) hipLaunchKernelGGL(unbin_kernel, dim3(a.X * a.Y * 3), dim3(kBinThreads), unbin_lds, s, a);
    return hipGetLastError();
}

}  // namespace hvx
