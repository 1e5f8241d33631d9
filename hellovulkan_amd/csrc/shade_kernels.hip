// shade_kernels.hip — clustered Cook-Torrance + split-sum IBL shading of SoA
// fragments on gfx950.
//
// Reference: HelloVulkan/Shaders/ClusteredForward/Scene.frag:60-149 (cluster
// lookup, per-light loop, ambient), ClusteredForward/Radiance.glsl:2-56,
// PBR/PBRHeader.glsl:16-64, PBR/Ambient.glsl:1-29.
//
// Scheduling. A workgroup owns 1024 consecutive fragments. It first resolves
// every fragment's cluster and light count, then counting-sorts the
// fragments by light count in LDS and hands them to lanes in that order, so
// the 64 lanes of a wave walk light lists of nearly equal length (with depth
// drawn independently per pixel, neighbouring pixels sit in unrelated
// clusters; unsorted, a wave runs to its longest list). Each fragment is still
// shaded start to finish by one lane, summing its lights in list order.
//
// Arithmetic, two variants:
//  exact  every expression as the FP32 oracle evaluates it (IEEE operations,
//         contraction off, correctly rounded divide and square root, the
//         shared log2/exp2/pow) — bit-identical to the oracle.
//  fast   (default) the light vector L, half vector H and N.H, N.L, H.V stay
//         exact, and so does D's denominator N.H^2 (a^2-1) + 1, because GGX's
//         D = a^2 / (pi (N.H^2 (a^2-1) + 1)^2) amplifies an error in N.H or in
//         that sum by up to ~1/a^2 (a^2 ~ 6e-6 at roughness 0.05);
//         the remaining per-light terms (attenuation, D, G, Fresnel, BRDF
//         combination) use fused multiply-add, hardware reciprocal and
//         hardware exp2/log2, reading a per-call compact light table
//         (position + 1/radius, colour * intensity). Every light term is
//         non-negative, so the sum keeps the terms' ~1e-6 relative accuracy.
#include <hip/hip_runtime.h>

#include "hvx_device.h"
#include "hvx_internal.h"
#include "mlp_device.h"

// build-time tuning switches (defaults are the tuned configuration)
#ifndef HVX_SHADE_ONE_RCP
#define HVX_SHADE_ONE_RCP 1
#endif
#ifndef HVX_SHADE_XCD_REMAP  // measured 4% slower on MI355X (4.81 vs 4.60 ms per frame): off
#define HVX_SHADE_XCD_REMAP 0
#endif
#ifndef HVX_SHADE_WAVES_PER_EU
#define HVX_SHADE_WAVES_PER_EU 4
#endif
#ifndef HVX_SHADE_PREFETCH  // lights in flight ahead of the one being shaded (1 or 2)
#define HVX_SHADE_PREFETCH 1
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
    const float4* lpos;   // fast: (x, y, z, 1/radius)
    const float4* lcol;   // fast: (r, g, b) * lightIntensity
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
    uint32_t mask_face, lod_bins;
    float mlp_tol;          // bound on the relative radiance error the network may cause
    uint32_t* served;       // fragments that used the network (nullable)
};

__global__ __launch_bounds__(256) void shade_light_prep_kernel(const hvx_light* __restrict__ lights, uint32_t n,
                                                               float intensity, float4* __restrict__ lpos,
                                                               float4* __restrict__ lcol) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const hvx_light& l = lights[i];
    lpos[i] = make_float4(l.position[0], l.position[1], l.position[2], 1.0f / l.radius);
    lcol[i] = make_float4(l.color[0] * intensity, l.color[1] * intensity, l.color[2] * intensity, 0.0f);
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

// Scene.frag:55-78: LinearDepth, log2 slice, tile -> cluster index
__device__ __forceinline__ uint32_t cluster_of(const ShadeArgs& a, uint32_t idx) {
    const uint32_t px = idx % a.W, py = idx / a.W;
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    const float z = a.depth[idx];
    const float lin = (a.nearv * a.farv) / (a.farv + z * (a.nearv - a.farv));
    const float zf = gmax(log2f_(lin) * a.scale + a.bias, 0.0f);
    uint32_t zi = (uint32_t)zf;
    if (zi >= a.Z) zi = a.Z - 1;
    const float tsx = a.scr_x / (float)a.X, tsy = a.scr_y / (float)a.Y;
    const uint32_t cx = min((uint32_t)(fx / tsx), a.X - 1), cy = min((uint32_t)(fy / tsy), a.Y - 1);
    return cx + cy * a.X + zi * a.X * a.Y;
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

// The network's measured relative error for (R, lod): the cell of the cube
// texel of R at the cell-grid resolution (face selection as for sampling)
// and of lod's bin.
__device__ __forceinline__ float mlp_cell_error(const ShadeArgs& a, f3 R, float lod) {
    const CubeCoord cc = cube_select(R);
    const uint32_t mf = a.mask_face;
    const uint32_t x = min((uint32_t)(cc.s * (float)mf), mf - 1), y = min((uint32_t)(cc.t * (float)mf), mf - 1);
    const uint32_t b = min((uint32_t)(lod / a.mlp.lod_max * (float)a.lod_bins), a.lod_bins - 1);
    return a.cell_err[((size_t)(b * 6 + (uint32_t)cc.face) * mf + y) * mf + x];
}

// Returns whether the learned radiance was used (cell_err >= 0 offers it:
// the network's value `learned` and the measured relative error of its cell).
template <bool EXACT>
__device__ __forceinline__ bool shade_one(const ShadeArgs& a, uint32_t idx, uint32_t off, uint32_t count,
                                          float cell_err = -1.0f, f3 learned = f3{0.0f, 0.0f, 0.0f}) {
    const float PI = kPiF;
    const uint32_t nfrag = a.W * a.H;
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
        float q[3], om[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            q[c] = (oneMinusMetal * alb[c]) * (1.0f / kPiF);
            om[c] = 1.0f - F0[c];
        }
        // software-pipelined: the next light's record is in flight while the
        // current one is shaded
        float4 An = make_float4(0.f, 0.f, 0.f, 0.f), Bn = An;
        if (count > 0) {
            const uint32_t li = a.indices[off];
            An = a.lpos[li];
            Bn = a.lcol[li];
        }
#if HVX_SHADE_PREFETCH >= 2
        float4 An2 = An, Bn2 = Bn;
        if (count > 1) {
            const uint32_t li = a.indices[off + 1];
            An2 = a.lpos[li];
            Bn2 = a.lcol[li];
        }
#endif
        for (uint32_t k = 0; k < count; ++k) {
            const float4 A = An, B = Bn;
#if HVX_SHADE_PREFETCH >= 2
            An = An2;
            Bn = Bn2;
            if (k + 2 < count) {
                const uint32_t li = a.indices[off + k + 2];
                An2 = a.lpos[li];
                Bn2 = a.lcol[li];
            }
#else
            if (k + 1 < count) {
                const uint32_t li = a.indices[off + k + 1];
                An = a.lpos[li];
                Bn = a.lcol[li];
            }
#endif
            // exact: L, H and the cosines (max with 0 only differs in the sign
            // of a zero, which no later term can observe)
            const f3 d = {A.x - wp.x, A.y - wp.y, A.z - wp.z};
            const float s = sqrt_rn_normal(dot3(d, d));
            const f3 L = mul3(d, rcp_rn_normal(s));
            const f3 h = add3(V, L);
            const f3 Hh = mul3(h, rcp_rn_normal(sqrt_rn_normal(dot3(h, h))));
            const float NoH = fmaxf(dot3(N, Hh), 0.0f);
            const float NoL = fmaxf(dot3(N, L), 0.0f);
            const float HoV = fmaxf(dot3(Hh, V), 0.0f);
            // fast: attenuation, D, G, Fresnel and the BRDF sum
            const float att = fmaxf(__builtin_fmaf(-s, A.w, 1.0f), 0.0f) *
                              __builtin_amdgcn_exp2f(negf * __builtin_amdgcn_logf(s));
            // den = NoH^2 (a^2 - 1) + 1 exactly as the oracle rounds it: near
            // N.H = 1 it cancels down to ~a^2, so one rounding more or less
            // moves D by up to 1e-7 / a^2
            const float den = (NoH * NoH) * a2m1 + 1.0f;
            // D * G_L * G_V / (4 NoV NoL + 1e-4) with one reciprocal of the
            // product of the three denominators (each bounded away from 0)
            const float gden = __builtin_fmaf(NoL, omk, alphaRough);
            const float sden = __builtin_fmaf(nov4, NoL, 0.0001f);
            const float x = fmaxf(1.0f - HoV, 0.0f);  // HoV >= 0 already
            const float x2 = x * x;
            const float p5 = (x2 * x2) * x;
#if HVX_SHADE_ONE_RCP
            const float DGr = (alpha2 * NoL * GV) * rcp_fast((((PI * den) * den) * gden) * sden);
#else
            const float DGr = (alpha2 * rcp_fast((PI * den) * den)) * (NoL * rcp_fast(gden)) * (GV * rcp_fast(sden));
#endif
            const float w = att * NoL;
            const float col[3] = {B.x, B.y, B.z};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float F = __builtin_fmaf(om[c], p5, F0[c]);
                const float brdf = __builtin_fmaf(1.0f - F, q[c], DGr * F);
                Lo[c] = __builtin_fmaf(brdf * col[c], w, Lo[c]);
            }
        }
    }

    // Ambient (PBR/Ambient.glsl:1-29), as the oracle evaluates it
    const float p5v = pow5f(gclamp(1.0f - NoV, 0.0f, 1.0f));
    const f3 irr = sample_cube(a.diff, (int)a.diff_size, N);
    const f3 I = {-V.x, -V.y, -V.z};
    const float dNI = dot3(N, I);
    const f3 R = sub3(I, mul3(N, 2.0f * dNI));
    if (!(cell_err >= 0.0f)) {  // plain: the map's prefiltered radiance
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
            a.out[(size_t)c * nfrag + idx] = (ambient + 0.0f) + Lo[c];
        }
        return false;
    }
    const float2 brdf = sample_2d_rg_repeat(a.lut, (int)a.lut_w, (int)a.lut_h, NoV, rough);
    const float irr3[3] = {irr.x, irr.y, irr.z};
    float kD[3], diffuse[3], sw[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float F = F0[c] + (gmax(1.0f - rough, F0[c]) - F0[c]) * p5v;
        kD[c] = (1.0f - F) * oneMinusMetal;
        diffuse[c] = irr3[c] * alb[c];
        sw[c] = F * brdf.x + brdf.y;  // weight of the prefiltered radiance
    }
    // The learned radiance replaces the map sample where the error it can
    // cause in the shaded radiance, e/(1-e) * learned * weight per channel
    // (e = the cell's measured relative error of the network), stays within
    // tol of the radiance without the specular ambient term (a lower bound
    // of the shaded value, every term being non-negative).
    bool use = false;
    if (cell_err < 1.0f) {
        const float k = cell_err / (1.0f - cell_err);
        const float l3[3] = {learned.x, learned.y, learned.z};
        use = true;
#pragma unroll
        for (int c = 0; c < 3; ++c) use = use && k * l3[c] * sw[c] <= a.mlp_tol * (kD[c] * diffuse[c] + Lo[c]);
    }
    const f3 pre = use ? learned : sample_spec_lod(a, R, rough * a.pc.max_reflection_lod);
    const float pre3[3] = {pre.x, pre.y, pre.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float specular = pre3[c] * sw[c];
        const float ambient = (kD[c] * diffuse[c] + specular) * 1.0f;
        a.out[(size_t)c * nfrag + idx] = (ambient + 0.0f) + Lo[c];
    }
    return use;
}

#ifndef HVX_SHADE_CHUNK
#define HVX_SHADE_CHUNK 1024
#endif
constexpr int kChunk = HVX_SHADE_CHUNK;   // fragments per workgroup
constexpr int kBuckets = 64;   // light-count classes (count / 2, saturating)
constexpr uint32_t kXcds = 8;  // MI355X accelerator complex dies


template <bool EXACT, bool LEARNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HVX_SHADE_WAVES_PER_EU, 8))) void shade_kernel(
    ShadeArgs a) {
    __shared__ uint32_t s_off[kChunk];
    __shared__ uint16_t s_cnt[kChunk];
    __shared__ uint16_t s_order[kChunk];
    __shared__ uint32_t s_hist[kBuckets];
    const uint32_t nfrag = a.W * a.H;
    // Optional XCD-aware chunk order: workgroups are dealt to the 8 XCDs
    // round-robin, so the remap would give XCD k a contiguous run of chunks (a
    // horizontal screen band) to confine its L2 working set. Measured slower
    // than the plain order for this workload, so it is compiled out by default.
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
    uint32_t bucket[kChunk / 256];
#pragma unroll
    for (int r = 0; r < kChunk / 256; ++r) {
        const uint32_t local = r * 256 + t, idx = base + local;
        uint32_t off = 0, cnt = 0;
        if (idx < nfrag) {
            const hvx_light_cell cell = a.cells[cluster_of(a, idx)];
            off = cell.offset;
            cnt = min(cell.count, 65535u);
        }
        s_off[local] = off;
        s_cnt[local] = (uint16_t)cnt;
        bucket[r] = min(cnt >> 1, (uint32_t)kBuckets - 1);
        atomicAdd(&s_hist[bucket[r]], 1u);
    }
    __syncthreads();
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
    for (int r = 0; r < kChunk / 256; ++r) {
        const uint32_t slot = atomicAdd(&s_hist[bucket[r]], 1u);
        s_order[slot] = (uint16_t)(r * 256 + t);
    }
    __syncthreads();
    if constexpr (LEARNED) {
        // Each wave evaluates the network for the 64 fragments it is about to
        // shade, as four 16-sample tiles; tile T holds the fragments of lanes
        // 16 T .. 16 T + 15 and its results return to those lanes by shuffles.
        __shared__ MlpLds s_mlp;
        __shared__ uint32_t s_served;
        mlp_stage(s_mlp, a.mlp);
        if (t == 0) s_served = 0;
        __syncthreads();
        const uint32_t lane = t & 63, i = lane & 15;
        uint32_t mine = 0;
#pragma unroll 1
        for (int r = 0; r < kChunk / 256; ++r) {
            const uint32_t local = s_order[r * 256 + t];
            const uint32_t idx = base + local;
            float x[4] = {0.0f, 0.0f, 1.0f, 0.0f};
            float err = -1.0f;
            if (idx < nfrag) {
                float lod;
                const f3 R = frag_reflect(a, idx, &lod);
                x[0] = R.x;
                x[1] = R.y;
                x[2] = R.z;
                x[3] = lod;
                err = mlp_cell_error(a, R, lod);
            }
            f3 learned = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
            for (int T = 0; T < 4; ++T) {
                const int src = 16 * T + (int)i;
                const float et = __shfl(err, src, 64);
                // tiles whose cells all exclude the network (error >= 1) skip it
                if (__ballot(et >= 0.0f && et < 1.0f) == 0) continue;
                const float xt[4] = {__shfl(x[0], src, 64), __shfl(x[1], src, 64), __shfl(x[2], src, 64),
                                     __shfl(x[3], src, 64)};
                const float y = mlp_forward_tile(s_mlp, xt, a.mlp.lod_max);  // lane (c, i): channel c of sample i
                const float c0 = __shfl(y, (int)i, 64), c1 = __shfl(y, 16 + (int)i, 64),
                            c2 = __shfl(y, 32 + (int)i, 64);
                if ((lane >> 4) == (uint32_t)T) learned = f3{c0, c1, c2};
            }
            if (idx < nfrag) mine += shade_one<EXACT>(a, idx, s_off[local], s_cnt[local], err, learned) ? 1u : 0u;
        }
        if (a.served) {
            if (mine) atomicAdd(&s_served, mine);
            __syncthreads();
            if (t == 0 && s_served) atomicAdd(a.served, s_served);
        }
    } else {
#pragma unroll 1
        for (int r = 0; r < kChunk / 256; ++r) {
            const uint32_t local = s_order[r * 256 + t];
            const uint32_t idx = base + local;
            if (idx < nfrag) shade_one<EXACT>(a, idx, s_off[local], s_cnt[local]);
        }
    }
}

}  // namespace

hipError_t launch_shade(const hvx_frag_soa& fr, const hvx_camera_ubo& cam, const hvx_cf_ubo& ubo,
                        const hvx_pbr_pc& pc, const hvx_light* d_lights, uint32_t n_lights,
                        const hvx_light_cell* d_cells, const uint32_t* d_indices, const hvx_ibl_maps& ibl,
                        float* d_rgb, bool exact, float4* light_scratch, const LearnedShade* learned,
                        hipStream_t s) {
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
    a.lpos = light_scratch;
    a.lcol = light_scratch + n_lights;
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
        a.mask_face = learned->mask_face;
        a.lod_bins = learned->lod_bins;
        a.mlp_tol = learned->tol;
        a.served = learned->served;
    }
    const uint32_t F = fr.width * fr.height;
    const dim3 grid((F + kChunk - 1) / kChunk);
    if (exact) {
        if (learned) hipLaunchKernelGGL((shade_kernel<true, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((shade_kernel<true, false>), grid, dim3(256), 0, s, a);
    } else {
        if (n_lights > 0) {
            hipLaunchKernelGGL(shade_light_prep_kernel, dim3((n_lights + 255) / 256), dim3(256), 0, s, d_lights,
                               n_lights, pc.light_intensity, light_scratch, light_scratch + n_lights);
        }
        if (learned) hipLaunchKernelGGL((shade_kernel<false, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((shade_kernel<false, false>), grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace hvx
