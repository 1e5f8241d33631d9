// ibl_kernels.hip — image-based-lighting precompute on gfx950.
//
// Reference shaders (HelloVulkan/Shaders/IBL) and their kernels here:
//   BRDFLUT.comp:26-71             -> brdf_lut_kernel (workgroup per LUT row;
//                                     the row's GGX half-vectors staged in LDS)
//   Equirect2Cube.frag:87-107      -> equirect_kernel (lane per cube texel)
//   VulkanImage::GenerateMipmap    -> mip_kernel (2x2 box, lane per texel)
//   CubeFilterDiffuse.frag:35-136  -> diffuse_stage_kernel (each bilinear
//                                     footprint's texels in one 64-byte
//                                     entry) and diffuse_kernel (32 lanes per
//                                     output texel evaluate samples, loads one
//                                     round ahead; one lane per texel and
//                                     channel sums them in order; the
//                                     phi/theta sin/cos grid in LDS)
//   CubeFilterSpecular.frag:34-146 -> ggx_table_kernel (tangent-space GGX
//                                     samples per mip, computed once),
//                                     specular_delta_kernel (mip 0, roughness
//                                     0: the n identical samples evaluated
//                                     once, lane per texel) and
//                                     specular_split_kernel (mips >= 1: each
//                                     texel's samples spread over 16 lanes, the
//                                     terms summed in sample order through
//                                     LDS; tap loads one round ahead)
// The per-sample sequence and every summation order are the reference's, so
// outputs reproduce the FP32 oracle to the bit; the precomputed tables hold
// exactly the values the per-sample code would compute (same functions, same
// inputs), they are only hoisted out of the per-texel loops.
#include <hip/hip_runtime.h>

#include "hvx_device.h"
#include "hvx_internal.h"

namespace hvx {
namespace {

// ImportanceSampleGGX (IBL/Header.glsl:14-28): tangent-space half vector of
// Hammersley sample i of n at roughness `rough`.
__device__ __forceinline__ float4 ggx_tangent_sample(uint32_t i, uint32_t n, float rough) {
    const float xix = (float)i / (float)n;
    const float xiy = radical_inverse_vdc(i);
    const float a = rough * rough;
    const float phi = (2.0f * kPiF) * xix;
    const float cosTheta = sqrtf_((1.0f - xiy) / (1.0f + (a * a - 1.0f) * xiy));
    const float sinTheta = sqrtf_(1.0f - cosTheta * cosTheta);
    return make_float4(cosf_(phi) * sinTheta, sinf_(phi) * sinTheta, cosTheta, 0.0f);
}

// tangent frame of IBL/Header.glsl:31-35
struct Frame {
    f3 t, b;
};
__device__ __forceinline__ Frame ggx_frame(f3 N) {
    const f3 up = absf(N.z) < 0.999f ? make3(0.0f, 0.0f, 1.0f) : make3(1.0f, 0.0f, 0.0f);
    const f3 t = normalize3(cross3(up, N));
    return {t, cross3(N, t)};
}

__device__ __forceinline__ f3 ggx_world(const Frame& fr, f3 N, float4 h) {
    const f3 sv = add3(add3(mul3(fr.t, h.x), mul3(fr.b, h.y)), mul3(N, h.z));
    return normalize3(sv);
}

// ---------------------------------------------------------------- BRDF LUT
constexpr int kLutChunk = 1024;

__global__ __launch_bounds__(256) void brdf_lut_kernel(uint32_t W, uint32_t H, uint32_t n, float2* __restrict__ lut) {
    __shared__ float4 s_h[kLutChunk];
    const uint32_t y = blockIdx.x;
    const float rough = 1.0f - ((float)y + 0.5f) / (float)H;
    const f3 N = make3(0.0f, 0.0f, 1.0f);
    const Frame fr = ggx_frame(N);
    const float alpha = (rough * rough) / 2.0f;  // AlphaIBL
    // each lane owns up to 4 texels of the row (W <= 1024)
    constexpr int kPer = 4;
    float lx[kPer], ly[kPer], nov[kPer];
    f3 V[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const uint32_t x = threadIdx.x + q * blockDim.x;
        nov[q] = ((float)min(x, W - 1) + 0.5f) / (float)W;
        V[q] = make3(sqrtf_(1.0f - nov[q] * nov[q]), 0.0f, nov[q]);
        lx[q] = 0.0f;
        ly[q] = 0.0f;
    }
    for (uint32_t c0 = 0; c0 < n; c0 += kLutChunk) {
        const uint32_t cn = min((uint32_t)kLutChunk, n - c0);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < cn; i += blockDim.x) {
            const float4 h = ggx_tangent_sample(c0 + i, n, rough);
            const f3 w = ggx_world(fr, N, h);
            s_h[i] = make_float4(w.x, w.y, w.z, 0.0f);
        }
        __syncthreads();
        for (uint32_t i = 0; i < cn; ++i) {
            const float4 hh = s_h[i];
            const f3 Hh = make3(hh.x, hh.y, hh.z);
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const f3 L = normalize3(sub3(mul3(Hh, 2.0f * dot3(V[q], Hh)), V[q]));
                const float dotNL = gmax(dot3(N, L), 0.0f);
                if (dotNL > 0.0f) {
                    const float dotNV = gmax(dot3(N, V[q]), 0.0f);
                    const float dotVH = gmax(dot3(V[q], Hh), 0.0f);
                    const float dotNH = gmax(dot3(Hh, N), 0.0f);
                    const float GL = dotNL / (dotNL * (1.0f - alpha) + alpha);
                    const float GV = dotNV / (dotNV * (1.0f - alpha) + alpha);
                    const float G = GL * GV;
                    const float G_Vis = (G * dotVH) / (dotNH * dotNV);
                    const float Fc = pow5f(1.0f - dotVH);
                    lx[q] = lx[q] + (1.0f - Fc) * G_Vis;
                    ly[q] = ly[q] + Fc * G_Vis;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const uint32_t x = threadIdx.x + q * blockDim.x;
        if (x < W) lut[(size_t)y * W + x] = make_float2(lx[q] / (float)n, ly[q] / (float)n);
    }
}

// ------------------------------------------------------------ equirect->cube
__global__ __launch_bounds__(256) void equirect_kernel(const float4* __restrict__ eq, uint32_t w, uint32_t h,
                                                       float4* __restrict__ cube, uint32_t S) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 6u * S * S) return;
    const uint32_t x = id % S, y = (id / S) % S, face = id / (S * S);
    const f3 d = texel_dir((int)face, (int)x, (int)y, (int)S);
    const float s = 0.5f + (0.5f * atan2f_(d.z, d.x)) / kPiF;
    const float t = 1.0f - acosf_(d.y) / kPiF;
    const f3 c = sample_2d_rgba_repeat(eq, (int)w, (int)h, s, t);
    cube[id] = make_float4(c.x, c.y, c.z, 1.0f);
}

// ---------------------------------------------------------------- mip chain
// grid dimension y: chains `pstride` float4s apart (one chain: y = 1)
__global__ __launch_bounds__(256) void mip_kernel(const float4* __restrict__ src, uint32_t sp, float4* __restrict__ dst,
                                                  uint32_t sn, size_t pstride) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 6u * sn * sn) return;
    src += blockIdx.y * pstride;
    dst += blockIdx.y * pstride;
    const uint32_t x = id % sn, y = (id / sn) % sn, f = id / (sn * sn);
    const float4* p = src + ((size_t)f * sp + 2 * y) * sp + 2 * x;
    const float4 a = p[0], b = p[1], c = p[sp], d = p[sp + 1];
    dst[id] = make_float4(((a.x + b.x) + (c.x + d.x)) * 0.25f, ((a.y + b.y) + (c.y + d.y)) * 0.25f,
                          ((a.z + b.z) + (c.z + d.z)) * 0.25f, ((a.w + b.w) + (c.w + d.w)) * 0.25f);
}

// Orders a wave's LDS stores before its later LDS loads of what other lanes
// of the same wave stored (no workgroup barrier needed).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- diffuse
// The reference sums 252 x 63 = 15,876 samples per output texel in one
// sequential loop (phi outer, theta inner). A lane per texel leaves the GPU
// idle (6,144 texels at 32^2), so here a workgroup owns TEX texels and its
// THREADS lanes evaluate THREADS / TEX consecutive samples of each texel per
// round; the weighted values go to LDS and one "summer" lane per texel and
// channel adds them in the reference's order. Results are bit-identical to
// the sequential loop. The kernel waits on gathers (at 1024^2 every sample
// reads about two fresh 128-byte lines; a few percent of its time issues
// arithmetic), so samples read their footprint from a staged copy
// (diffuse_stage_kernel, one line per sample), a lane's loads are issued
// one round ahead of their use, and the LDS rounds are double-buffered: the
// gathers of round r+1 are in flight while round r is weighted, stored and
// summed. 32^2 from 1024^2 (staging included): 1.79 ms with 8 texels per
// 256-lane workgroup (staging 0.12 ms of it; 1.80 with 4 texels and the
// staging's stores one entry per lane); 2.01 ms when only footprints inside a
// face were staged (the others gathered per lane); gathered from the cube
// (the path for stream capture before the context holds a staging buffer)
// 2.66 ms with 4 texels (3.11 / 3.31 / 2.77 ms with 8 / 16 / 2, 3.02 with 4
// per 128-lane workgroup); 3.54 ms with 16 texels and no loads ahead.
// Staged loads two rounds ahead (79 VGPRs): no faster (1.81 ms).
constexpr int kMaxPhi = 512, kMaxTheta = 128;
#ifndef HVX_DIFFUSE_TEX
#define HVX_DIFFUSE_TEX 8
#endif
constexpr int kDTex = HVX_DIFFUSE_TEX;  // texels per workgroup
static_assert(kDTex == 2 || kDTex == 4 || kDTex == 8 || kDTex == 16, "texels per workgroup: 2, 4, 8 or 16");
#ifndef HVX_DIFFUSE_THREADS
#define HVX_DIFFUSE_THREADS 256
#endif
constexpr int kDThreads = HVX_DIFFUSE_THREADS;  // lanes per workgroup
static_assert(kDThreads == 128 || kDThreads == 256, "lanes per workgroup: 128 or 256");
#ifndef HVX_DIFFUSE_WAVES_PER_EU
#define HVX_DIFFUSE_WAVES_PER_EU 4
#endif

// Staged footprints: for every bilinear footprint of level 0, first texel
// (i0, j0) with -1 <= i0, j0 <= S - 1 on face f, one 64-byte entry at
// (f (S+1) + j0 + 1) (S+1) + i0 + 1 holding the RGB of its four taps
// exactly as sample_cube reads them (texels across an edge folded onto the
// neighbouring face, a tap off both edges the mean of three texels):
// (t00.rgb, t11.r) (t10.rgb, t11.g) (t01.rgb, t11.b), then 16 bytes of zeros.
// A sample then reads one 128-byte line instead of two rows of texels (2.3
// lines on average), and the bilinear sums are unchanged to the bit.
__global__ __launch_bounds__(256) void diffuse_stage_kernel(const float4* __restrict__ cube, uint32_t S,
                                                           float4* __restrict__ quads) {
    // lane g writes 16-byte part g % 4 of entry g / 4: a wave's stores are
    // one contiguous kilobyte
    const uint32_t Se = S + 1, per = Se * Se;
    const uint32_t g = blockIdx.x * 256u + threadIdx.x, e = g / 4u, k = g % 4u;
    if (e >= 6u * per) return;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // part 3: zeros (whole lines written)
    if (k < 3u) {
        // part k: tap k (t00, t10, t01) and channel k of t11
        const uint32_t f = e / per, r = e % per;
        const int j0 = (int)(r / Se) - 1, i0 = (int)(r % Se) - 1, Si = (int)S;
        const int ti = i0 + (k == 1u ? 1 : 0), tj = j0 + (k == 2u ? 1 : 0);
        f3 t, t11;
        if (i0 >= 0 && j0 >= 0 && i0 + 1 < Si && j0 + 1 < Si) {
            const float4 a = cube[((size_t)f * S + tj) * S + ti];
            const float4 d = cube[((size_t)f * S + j0 + 1) * S + i0 + 1];
            t = {a.x, a.y, a.z};
            t11 = {d.x, d.y, d.z};
        } else {
            t = cube_texel_seamless(cube, Si, (int)f, ti, tj);
            t11 = cube_texel_seamless(cube, Si, (int)f, i0 + 1, j0 + 1);
        }
        v = make_float4(t.x, t.y, t.z, k == 0u ? t11.x : (k == 1u ? t11.y : t11.z));
    }
    quads[g] = v;
}

// One lane's sample of one round: direction, theta terms, taps and loads
// (gathered: q[k] = tap k's texel; staged: q[0..2] = the footprint's entry).
struct DiffSample {
    f3 L;
    float ct, st;
    CubeTaps taps;
    float4 q[4];
};

// STAGED: the samples read staged footprints (three 16-byte loads from one
// line); otherwise their four taps from the cube. Four lanes loading one
// entry together (16 lines per load instruction instead of 64, the parts
// passed on through LDS) measured slower: 2.36 ms against 1.82.
template <int TEX, int THREADS, bool STAGED>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(HVX_DIFFUSE_WAVES_PER_EU, 8))) void
diffuse_kernel(const float4* __restrict__ cube, uint32_t S, const float4* __restrict__ quads, float4* __restrict__ out,
               uint32_t Sd) {
    constexpr uint32_t SLOT = THREADS / TEX;  // samples per texel per round
    __shared__ float s_phi[kMaxPhi], s_theta[kMaxTheta];
    __shared__ float s_sp[kMaxPhi], s_cp[kMaxPhi], s_st[kMaxTheta], s_ct[kMaxTheta];
    __shared__ float s_buf[2][SLOT][TEX][3];
    __shared__ uint32_t s_np, s_nt;
    const float TWO_PI = kPiF * 2.0f, HALF_PI = kPiF * 0.5f;
    const float delta = 0.025f;
    const uint32_t l = threadIdx.x;
    if (l == 0) {
        uint32_t np = 0;
        for (float phi = 0.0f; phi < TWO_PI && np < (uint32_t)kMaxPhi; phi = phi + delta) s_phi[np++] = phi;
        uint32_t nt = 0;
        for (float th = 0.0f; th < HALF_PI && nt < (uint32_t)kMaxTheta; th = th + delta) s_theta[nt++] = th;
        s_np = np;
        s_nt = nt;
    }
    __syncthreads();
    const uint32_t np = s_np, nt = s_nt, total = np * nt;
    for (uint32_t i = l; i < np; i += blockDim.x) {
        s_sp[i] = sinf_(s_phi[i]);
        s_cp[i] = cosf_(s_phi[i]);
    }
    for (uint32_t i = l; i < nt; i += blockDim.x) {
        s_st[i] = sinf_(s_theta[i]);
        s_ct[i] = cosf_(s_theta[i]);
    }
    __syncthreads();
    const uint32_t ntex = 6u * Sd * Sd;
    // sample lane: texel l % TEX, slot l / TEX
    const uint32_t ti = l % TEX, slot = l / TEX;
    const uint32_t texel = min(blockIdx.x * TEX + ti, ntex - 1);
    const uint32_t x = texel % Sd, y = (texel / Sd) % Sd, face = texel / (Sd * Sd);
    const f3 N = texel_dir((int)face, (int)x, (int)y, (int)Sd);
    f3 up = make3(0.0f, 1.0f, 0.0f);
    const f3 right = normalize3(cross3(up, N));
    up = normalize3(cross3(N, right));
    // summer lane: l < 3 * TEX -> texel l / 3, channel l % 3
    const bool summer = l < 3u * TEX;
    const uint32_t st_i = l / 3, st_c = l % 3;
    float acc = 0.0f;
    const uint32_t rounds = (total + SLOT - 1) / SLOT;
    // sample r * SLOT + slot; slots past the last sample evaluate the last
    // one again (in range, never summed), so the loads stay unconditional
    auto begin = [&](DiffSample& p, uint32_t r) {
        const uint32_t sidx = min(r * SLOT + slot, total - 1);
        const uint32_t ip = sidx / nt, it = sidx % nt;
        const float sp = s_sp[ip], cp = s_cp[ip];
        p.st = s_st[it];
        p.ct = s_ct[it];
        const float tx = p.st * cp, ty = p.st * sp, tz = p.ct;
        p.L = add3(add3(mul3(right, tx), mul3(up, ty)), mul3(N, tz));
        p.taps = cube_taps((int)S, p.L);
        if constexpr (STAGED) {
            // i0, j0 lie in [-1, S - 1] (the face coordinates are in [0, 1]);
            // the clamp only keeps the index inside the table
            const int i0 = min(max(p.taps.i0, -1), (int)S - 1), j0 = min(max(p.taps.j0, -1), (int)S - 1);
            const uint32_t e = ((uint32_t)p.taps.face * (S + 1) + (uint32_t)(j0 + 1)) * (S + 1) + (uint32_t)(i0 + 1);
#pragma unroll
            for (int k = 0; k < 3; ++k) p.q[k] = quads[4 * (size_t)e + k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) p.q[k] = cube[p.taps.idx[k]];
        }
    };
    // weight round r's sample, store it, and (summers) add the round's terms
    // in sample order; every lane of the workgroup calls it for every r
    auto sum_round = [&](uint32_t r, const DiffSample& p) {
        float (*buf)[TEX][3] = s_buf[r & 1];
        f3 c;
        if constexpr (STAGED) {
            // sample_cube's bilinear sum of the entry's taps
            const float4 q0 = p.q[0], q1 = p.q[1], q2 = p.q[2];
            const CubeTaps& t = p.taps;
            c = {((t.w00 * q0.x + t.w10 * q1.x) + t.w01 * q2.x) + t.w11 * q0.w,
                 ((t.w00 * q0.y + t.w10 * q1.y) + t.w01 * q2.y) + t.w11 * q1.w,
                 ((t.w00 * q0.z + t.w10 * q1.z) + t.w01 * q2.z) + t.w11 * q2.w};
        } else {
            c = cube_combine(cube, (int)S, p.L, p.taps, p.q);
        }
        buf[slot][ti][0] = (c.x * p.ct) * p.st;
        buf[slot][ti][1] = (c.y * p.ct) * p.st;
        buf[slot][ti][2] = (c.z * p.ct) * p.st;
        __syncthreads();
        if (summer && r < rounds) {
            const uint32_t kmax = min(SLOT, total - r * SLOT);
            for (uint32_t k = 0; k < kmax; ++k) acc = acc + buf[k][st_i][st_c];
        }
    };
    DiffSample A, B;
    begin(A, 0);
    for (uint32_t r = 0; r < rounds; r += 2) {
        begin(B, r + 1);
        sum_round(r, A);
        begin(A, r + 2);
        sum_round(r + 1, B);
    }
    if (summer) {
        const uint32_t t = blockIdx.x * TEX + st_i;
        if (t < ntex) {
            const float inv = 1.0f / (float)total;
            reinterpret_cast<float*>(out + t)[st_c] = (kPiF * acc) * inv;
            if (st_c == 0) reinterpret_cast<float*>(out + t)[3] = 1.0f;
        }
    }
}

// ---------------------------------------------------------------- specular
__global__ __launch_bounds__(256) void ggx_table_kernel(uint32_t mips, uint32_t n, float4* __restrict__ table) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= mips * n) return;
    const uint32_t m = id / n, i = id % n;
    // PipelineCubeFilter.cpp:305-311: roughness = float(i) / float(mips - 1)
    const float rough = mips <= 1 ? 0.0f : (float)m / (float)(mips - 1);
    table[id] = ggx_tangent_sample(i, n, rough);
}

constexpr int kMaxMips = 16;

// One output texel of the GGX prefilter (CubeFilterSpecular.frag:34-75) at
// direction-grid position (face, x, y) of a D x D viewport; `tab` holds the
// mip's n tangent-space samples; rough0 marks the roughness-0 mip.
__device__ __forceinline__ float4 specular_texel_value(const float4* __restrict__ cube, uint32_t S,
                                                       const float4* __restrict__ tab, uint32_t n, bool rough0,
                                                       int face, int x, int y, int D) {
    const f3 N = texel_dir(face, x, y, D);
    const f3 V = N;
    const Frame fr = ggx_frame(N);
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, wsum = 0.0f;
    if (rough0) {
        // roughness 0 (PipelineCubeFilter.cpp:305-311 gives mip 0 roughness 0):
        // cos(theta) = sqrt((1-y)/(1-y)) = 1 and sin(theta) = 0 for every
        // Hammersley sample, so each sample's half vector is normalize(N + (+-0))
        // and all n samples produce the same L, NoL and texel value. Evaluate it
        // once and form the reference's sequential sums of the n identical terms.
        const f3 Hh = ggx_world(fr, N, tab[0]);
        const f3 L = normalize3(sub3(mul3(Hh, 2.0f * dot3(V, Hh)), V));
        const float NoL = gmax(dot3(N, L), 0.0f);
        if (NoL > 0.0f) {
            const f3 c = sample_cube(cube, (int)S, L);
            const float t0 = c.x * NoL, t1 = c.y * NoL, t2 = c.z * NoL;
            for (uint32_t i = 0; i < n; ++i) {
                acc0 = acc0 + t0;
                acc1 = acc1 + t1;
                acc2 = acc2 + t2;
                wsum = wsum + NoL;
            }
        }
    } else {
        for (uint32_t i = 0; i < n; ++i) {
            const f3 Hh = ggx_world(fr, N, tab[i]);
            const f3 L = normalize3(sub3(mul3(Hh, 2.0f * dot3(V, Hh)), V));
            const float NoL = gmax(dot3(N, L), 0.0f);
            if (NoL > 0.0f) {
                const f3 c = sample_cube(cube, (int)S, L);
                acc0 = acc0 + c.x * NoL;
                acc1 = acc1 + c.y * NoL;
                acc2 = acc2 + c.z * NoL;
                wsum = wsum + NoL;
            }
        }
    }
    return make_float4(acc0 / wsum, acc1 / wsum, acc2 / wsum, 1.0f);
}

// Mip 0 has roughness 0 (PipelineCubeFilter.cpp:305-311): one lane per texel.
__global__ __launch_bounds__(256) void specular_delta_kernel(const float4* __restrict__ cube, uint32_t S,
                                                             float4* __restrict__ out, uint32_t So,
                                                             const float4* __restrict__ table, uint32_t n) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 6u * So * So) return;
    const uint32_t x = id % So, y = (id / So) % So, face = id / (So * So);
    out[id] = specular_texel_value(cube, S, table, n, true, (int)face, (int)x, (int)y, (int)So);
}

// Mips of roughness > 0, samples spread over lanes. A wave owns kSpecTPW
// texels of a tile and evaluates kSpecSPR = 64 / kSpecTPW consecutive
// samples of each per round; every lane writes its sample's four terms
// (c * NoL, NoL; -0 for a rejected sample, the exact identity of x + y) to
// LDS, and one "summer" lane per texel and term adds that round's terms in
// sample order, so every sum is the reference's sequential sum (bit-exact).
// Few texels per wave keep the texels in flight per XCD few, so the GGX lobes
// being read at any time share the XCD's L2: tiles are taken in 16x16-tile
// supertiles, and the workgroups of one XCD (workgroup b runs on XCD b % 8)
// get one contiguous run of tiles.
#ifndef HVX_SPEC_TPW
#define HVX_SPEC_TPW 4
#endif
constexpr int kSpecTPW = HVX_SPEC_TPW;                   // texels per wave: 4, 8 or 16
constexpr int kSpecSPR = 64 / kSpecTPW;                  // samples per texel per round
constexpr int kSpecTW = kSpecTPW >= 8 ? 8 : 4;           // tile width (texels)
constexpr int kSpecRPW = kSpecTPW / kSpecTW;             // texel rows per wave
constexpr int kSpecTH = 4 * kSpecRPW;                    // tile height: four waves
constexpr int kSpecRow = 4 * kSpecSPR + 4;               // floats per texel row of the LDS buffer (padded)
static_assert(kSpecTPW == 4 || kSpecTPW == 8 || kSpecTPW == 16, "HVX_SPEC_TPW is 4, 8 or 16");

struct SpecSplitArgs {
    const float4* cube;
    uint32_t S;
    float4* out;
    uint32_t So, mips, n, emulate;
    uint32_t n_tiles, per_xcd;          // padded tiles of mips 1.. and tiles per XCD run
    uint32_t tile_begin[kMaxMips + 1];  // first tile of mip m (mips 1..), tile_begin[1] = 0
    const float4* table;
};

// One GGX sample of the specular loop, split so its texel loads overlap other
// work: spec_begin forms L and NoL and issues the four tap loads (always, so
// the loads are not under a branch; a rejected sample's taps are valid
// texels whose values are not used), spec_term forms the sample's terms.
struct SpecSample {
    f3 L;
    float nol;
    CubeTaps taps;
    float4 q[4];
};

__device__ __forceinline__ void spec_begin(SpecSample& p, const float4* __restrict__ cube, int S, const Frame& fr,
                                           f3 N, float4 h) {
    // both vectors are unit length up to rounding: normalize3_rn is exact here
    const f3 Hh = normalize3_rn(add3(add3(mul3(fr.t, h.x), mul3(fr.b, h.y)), mul3(N, h.z)));  // ggx_world
    p.L = normalize3_rn(sub3(mul3(Hh, 2.0f * dot3(N, Hh)), N));                            // V = N
    p.nol = gmax(dot3(N, p.L), 0.0f);
    p.taps = cube_taps(S, p.L);
#pragma unroll
    for (int k = 0; k < 4; ++k) p.q[k] = cube[p.taps.idx[k]];
}

// (c * NoL, NoL) of a sample that counts (NoL > 0); -0 in every term of one
// that does not (a rejected or padding sample): x + (-0) == x for every x.
__device__ __forceinline__ float4 spec_term(const SpecSample& p, bool live, const float4* __restrict__ cube, int S) {
    if (!(live && p.nol > 0.0f)) return make_float4(-0.0f, -0.0f, -0.0f, -0.0f);
    const f3 c = cube_combine(cube, S, p.L, p.taps, p.q);
    return make_float4(c.x * p.nol, c.y * p.nol, c.z * p.nol, p.nol);
}

// Tile geometry of one mip of size Sm: tiles of kSpecTW x kSpecTH texels,
// taken in 16 x 16-tile supertiles (row-major supertiles, row-major tiles in
// each); padded tile k of a face maps to (tx, ty), outside the face when the
// supertile is partial.
struct SpecTileGeom {
    uint32_t tiles_x, tiles_y, super_x, per_face;  // per_face: padded tiles per face
};
__host__ __device__ inline SpecTileGeom spec_tile_geom(uint32_t Sm) {
    SpecTileGeom g;
    g.tiles_x = (Sm + kSpecTW - 1) / kSpecTW;
    g.tiles_y = (Sm + kSpecTH - 1) / kSpecTH;
    g.super_x = (g.tiles_x + 15) / 16;
    g.per_face = g.super_x * ((g.tiles_y + 15) / 16) * 256u;
    return g;
}
// Padded tile index -> (face, tx, ty); false for a padding slot.
__host__ __device__ inline bool spec_tile_at(const SpecTileGeom& g, uint32_t k, uint32_t& face, uint32_t& tx,
                                             uint32_t& ty) {
    face = k / g.per_face;
    const uint32_t kk = k % g.per_face, ks = kk / 256u, kw = kk % 256u;
    tx = (ks % g.super_x) * 16u + kw % 16u;
    ty = (ks / g.super_x) * 16u + kw / 16u;
    return tx < g.tiles_x && ty < g.tiles_y;
}

// One tile of the sample-split specular loop (four waves, each independent):
// texels of face `face` of a mip of size Sm whose directions lie on a D x D
// grid, out_level that mip's texels, tab its n tangent-space samples, s_buf
// the workgroup's [4][2][kSpecTPW][kSpecRow] summation buffer.
__device__ __forceinline__ void specular_split_tile(const float4* __restrict__ cube, int S,
                                                    float4* __restrict__ out_level, uint32_t Sm, uint32_t D,
                                                    const float4* __restrict__ tab, uint32_t n, uint32_t face,
                                                    uint32_t tx, uint32_t ty, float* s_buf) {
    const uint32_t wave = threadIdx.x / 64u, lane = threadIdx.x % 64u;
    // evaluating lane: texel t of the wave, sample slot s
    const uint32_t t = lane / kSpecSPR, s = lane % kSpecSPR;
    const uint32_t x = min(tx * kSpecTW + t % kSpecTW, Sm - 1),
                   y = min(ty * kSpecTH + wave * kSpecRPW + t / kSpecTW, Sm - 1);
    const f3 N = texel_dir((int)face, (int)x, (int)y, (int)D);
    const Frame fr = ggx_frame(N);
    // summing lane: texel st of the wave, term sc (R, G, B, NoL)
    const bool summer = lane < 4u * kSpecTPW;
    const uint32_t st = lane / 4u, sc = lane % 4u;
    float acc = 0.0f;
    constexpr uint32_t kHalf = kSpecTPW * kSpecRow;  // floats between the two buffers
    float* const wbuf = s_buf + wave * 2u * kHalf + t * kSpecRow + 4u * s;
    const float* const rbuf = s_buf + wave * 2u * kHalf + st * kSpecRow + sc;
    // rounds in pairs (A: even, B: odd); sample i of round r is r * SPR + s,
    // table entries are fetched two rounds ahead and tap loads one round
    // ahead of their use
    const uint32_t rounds = (n + kSpecSPR - 1) / kSpecSPR;
    auto tab_at = [&](uint32_t r) { return tab[min(r * kSpecSPR + s, n - 1)]; };
    auto sum_round = [&](uint32_t parity, float4 term) {
        *reinterpret_cast<float4*>(wbuf + parity * kHalf) = term;
        wave_lds_sync();
        if (summer) {
#pragma unroll
            for (int q = 0; q < kSpecSPR; ++q) acc = acc + rbuf[parity * kHalf + 4 * q];
        }
    };
    SpecSample A, B;
    float4 hA = tab_at(0), hB = tab_at(1);
    spec_begin(A, cube, S, fr, N, hA);
    hA = tab_at(2);
    for (uint32_t r = 0; r < rounds; r += 2) {
        spec_begin(B, cube, S, fr, N, hB);
        hB = tab_at(r + 3);
        sum_round(0, spec_term(A, r * kSpecSPR + s < n, cube, S));
        spec_begin(A, cube, S, fr, N, hA);
        hA = tab_at(r + 4);
        sum_round(1, spec_term(B, (r + 1) * kSpecSPR + s < n, cube, S));
    }
    const float wsum = __shfl(acc, (int)((lane & ~3u) | 3u));
    const uint32_t xs = tx * kSpecTW + st % kSpecTW, ys = ty * kSpecTH + wave * kSpecRPW + st / kSpecTW;
    if (summer && xs < Sm && ys < Sm) {
        float* o = reinterpret_cast<float*>(out_level + (face * Sm + ys) * Sm + xs);
        o[sc] = sc == 3u ? 1.0f : acc / wsum;
    }
    // the next tile of this workgroup reuses the buffers: keep its first
    // writes after these reads
    wave_lds_sync();
}

#ifndef HVX_SPEC_WAVES_PER_EU
#define HVX_SPEC_WAVES_PER_EU 4
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HVX_SPEC_WAVES_PER_EU, 8))) void
specular_split_kernel(SpecSplitArgs a) {
    __shared__ __align__(16) float s_buf[4][2][kSpecTPW][kSpecRow];
    const uint32_t b = blockIdx.x;
    const uint32_t tile = (b % 8u) * a.per_xcd + b / 8u;
    if (tile >= a.n_tiles) return;
    uint32_t m = 1;
    while (tile >= a.tile_begin[m + 1]) ++m;
    const uint32_t Sm = max(a.So >> m, 1u);
    uint32_t face, tx, ty;
    if (!spec_tile_at(spec_tile_geom(Sm), tile - a.tile_begin[m], face, tx, ty)) return;  // padding
    specular_split_tile(a.cube, (int)a.S, a.out + cube_level_offset_floats(a.So, m) / 4, Sm, a.emulate ? a.So : Sm,
                        a.table + (size_t)m * a.n, a.n, face, tx, ty, &s_buf[0][0][0][0]);
}

// ------------------------------------------------- batched probe prefilter
// Persistent: the workgroups on XCD x (workgroup b runs on XCD b % 8) pull
// items from queue x, which serves the contiguous run [xcd_begin[x],
// xcd_begin[x+1]) of the item list; the host orders items by probe, mip and
// tile and cuts the runs at equal estimated cost, so each XCD works through
// neighbouring tiles of a few probes and shares their texels in its L2.
// Items of mip 0 are texel ranges (lane per texel, as specular_delta_kernel);
// other items are runs of padded tiles (specular_split_tile).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HVX_SPEC_WAVES_PER_EU, 8))) void
probe_batch_kernel(const ProbeJob* __restrict__ jobs, const ProbeItem* __restrict__ items,
                   const uint32_t* __restrict__ xcd_begin, const float4* __restrict__ tables,
                   uint32_t* __restrict__ queues) {
    __shared__ __align__(16) float s_buf[4][2][kSpecTPW][kSpecRow];
    __shared__ uint32_t s_item;
    const uint32_t x = blockIdx.x % 8u;
    const uint32_t begin = xcd_begin[x], end = xcd_begin[x + 1];
    for (;;) {
        if (threadIdx.x == 0) s_item = begin + atomicAdd(queues + x, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= end) break;
        const ProbeItem it = items[item];
        const ProbeJob& j = jobs[it.probe];
        const uint32_t m = it.mip;
        const uint32_t Sm = max(j.So >> m, 1u);
        const float4* tab = tables + j.table + (size_t)m * j.n;
        float4* out = j.out + cube_level_offset_floats(j.So, m) / 4;
        if (m == 0) {
            for (uint32_t t = threadIdx.x; t < it.count; t += blockDim.x) {
                const uint32_t local = it.first + t;
                const uint32_t tx = local % Sm, ty = (local / Sm) % Sm, face = local / (Sm * Sm);
                out[local] = specular_texel_value(j.cube, j.S, tab, j.n, true, (int)face, (int)tx, (int)ty, (int)Sm);
            }
        } else {
            const SpecTileGeom g = spec_tile_geom(Sm);
            for (uint32_t k = it.first; k < it.first + it.count; ++k) {
                uint32_t face, tx, ty;
                if (spec_tile_at(g, k, face, tx, ty))
                    specular_split_tile(j.cube, (int)j.S, out, Sm, Sm, tab, j.n, face, tx, ty, &s_buf[0][0][0][0]);
            }
        }
    }
}

}  // namespace

// ================================================================ launchers
hipError_t launch_brdf_lut(uint32_t w, uint32_t h, uint32_t n, float* d_lut, hipStream_t s) {
    hipLaunchKernelGGL(brdf_lut_kernel, dim3(h), dim3(256), 0, s, w, h, n, reinterpret_cast<float2*>(d_lut));
    return hipGetLastError();
}

hipError_t launch_equirect_to_cube(const float* d_eq, uint32_t w, uint32_t h, float* d_cube, uint32_t S,
                                   hipStream_t s) {
    const uint32_t T = 6u * S * S;
    hipLaunchKernelGGL(equirect_kernel, dim3((T + 255) / 256), dim3(256), 0, s, reinterpret_cast<const float4*>(d_eq),
                       w, h, reinterpret_cast<float4*>(d_cube), S);
    return hipGetLastError();
}

hipError_t launch_cube_mips(float* d_chain, uint32_t S, uint32_t levels, hipStream_t s, uint32_t n,
                            size_t chain_stride_floats) {
    for (uint32_t m = 1; m < levels; ++m) {
        const uint32_t sp = std::max(S >> (m - 1), 1u), sn = std::max(S >> m, 1u);
        const float4* src = reinterpret_cast<const float4*>(d_chain + cube_level_offset_floats(S, m - 1));
        float4* dst = reinterpret_cast<float4*>(d_chain + cube_level_offset_floats(S, m));
        const uint32_t T = 6u * sn * sn;
        hipLaunchKernelGGL(mip_kernel, dim3((T + 255) / 256, n), dim3(256), 0, s, src, sp, dst, sn,
                           chain_stride_floats / 4);
    }
    return hipGetLastError();
}

// dst + p * dst_stride <- src[p], `count` float4s, for p < n (grid dimension y)
__global__ __launch_bounds__(256) void gather_copy_kernel(const float4* const* __restrict__ src, size_t count,
                                                          float4* __restrict__ dst, size_t dst_stride) {
    const float4* from = src[blockIdx.y];
    float4* to = dst + blockIdx.y * dst_stride;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) to[i] = from[i];
}

hipError_t launch_gather_copy(const float* const* d_src, size_t floats, float* d_dst, size_t dst_stride_floats,
                              uint32_t n, hipStream_t s) {
    const size_t count = floats / 4;
    const uint32_t blocks = (uint32_t)std::min<size_t>((count + 255) / 256, 1024);
    hipLaunchKernelGGL(gather_copy_kernel, dim3(std::max(blocks, 1u), n), dim3(256), 0, s,
                       reinterpret_cast<const float4* const*>(d_src), count, reinterpret_cast<float4*>(d_dst),
                       dst_stride_floats / 4);
    return hipGetLastError();
}

size_t diffuse_stage_size(uint32_t S) {
    const size_t b = (size_t)6 * (S + 1) * (S + 1) * 64;
    return b <= kDiffuseStageMaxBytes ? b : 0;
}

hipError_t launch_diffuse(const float* d_cube, uint32_t S, float* d_out, uint32_t Sd, float* d_stage, hipStream_t s) {
    const float4* cube = reinterpret_cast<const float4*>(d_cube);
    float4* quads = diffuse_stage_size(S) ? reinterpret_cast<float4*>(d_stage) : nullptr;
    const uint32_t T = 6u * Sd * Sd;
    float4* out = reinterpret_cast<float4*>(d_out);
    if (quads) {
        const uint32_t parts = 4u * 6u * (S + 1) * (S + 1);  // 16-byte parts of the entries
        hipLaunchKernelGGL(diffuse_stage_kernel, dim3((parts + 255) / 256), dim3(256), 0, s, cube, S, quads);
        hipLaunchKernelGGL((diffuse_kernel<kDTex, kDThreads, true>), dim3((T + kDTex - 1) / kDTex), dim3(kDThreads), 0,
                           s, cube, S, quads, out, Sd);
    } else {
        hipLaunchKernelGGL((diffuse_kernel<kDTex, kDThreads, false>), dim3((T + kDTex - 1) / kDTex), dim3(kDThreads),
                           0, s, cube, S, nullptr, out, Sd);
    }
    return hipGetLastError();
}

hipError_t launch_specular(const float* d_cube, uint32_t S, float* d_out, uint32_t So, uint32_t mips, uint32_t n,
                           bool emulate_static_viewport, float4* ggx_table, hipStream_t s) {
    if (mips > (uint32_t)kMaxMips) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ggx_table_kernel, dim3((mips * n + 255) / 256), dim3(256), 0, s, mips, n, ggx_table);
    const float4* cube = reinterpret_cast<const float4*>(d_cube);
    float4* out = reinterpret_cast<float4*>(d_out);
    // mip 0 (roughness 0) covers its whole face whether or not the viewport
    // is emulated: its size is the base size
    hipLaunchKernelGGL(specular_delta_kernel, dim3((6u * So * So + 255) / 256), dim3(256), 0, s, cube, S, out, So,
                       ggx_table, n);
    if (mips > 1) {
        SpecSplitArgs a{};
        a.cube = cube;
        a.S = S;
        a.out = out;
        a.So = So;
        a.mips = mips;
        a.n = n;
        a.emulate = emulate_static_viewport ? 1u : 0u;
        a.table = ggx_table;
        uint32_t t = 0;
        for (uint32_t m = 1; m <= (uint32_t)kMaxMips; ++m) {
            a.tile_begin[m] = t;
            if (m < mips) t += 6u * spec_tile_geom(std::max(So >> m, 1u)).per_face;
        }
        a.n_tiles = t;
        a.per_xcd = (t + 7) / 8;
        hipLaunchKernelGGL(specular_split_kernel, dim3(a.per_xcd * 8), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_ggx_tables(uint32_t mips, uint32_t n, float4* table, hipStream_t s) {
    hipLaunchKernelGGL(ggx_table_kernel, dim3((mips * n + 255) / 256), dim3(256), 0, s, mips, n, table);
    return hipGetLastError();
}

hipError_t launch_probe_batch(const ProbeJob* d_jobs, const ProbeItem* d_items, const uint32_t* d_xcd_begin,
                              const float4* d_tables, uint32_t persistent_wgs, uint32_t* d_queues, hipStream_t s) {
    const uint32_t grid = std::max((persistent_wgs + 7) / 8, 1u) * 8;  // every XCD gets workgroups
    hipLaunchKernelGGL(probe_batch_kernel, dim3(grid), dim3(256), 0, s, d_jobs, d_items, d_xcd_begin, d_tables,
                       d_queues);
    return hipGetLastError();
}

// Batch work items (probe, mip, range) for the probes of `jobs`, in probe,
// mip, tile order: mip 0 as ranges of `texels` texels, other mips as runs of
// padded tiles holding about texels / 16 tiles each; xcd_begin[0..8] cuts the
// list into eight runs of equal estimated cost.
void make_probe_items(const ProbeJob* jobs, uint32_t n_probes, uint32_t texels, std::vector<ProbeItem>& items,
                      uint32_t xcd_begin[9]) {
    items.clear();
    std::vector<double> cost;
    const uint32_t per_item = std::max(texels / (uint32_t)(kSpecTW * kSpecTH), 1u);
    for (uint32_t p = 0; p < n_probes; ++p) {
        const ProbeJob& j = jobs[p];
        for (uint32_t m = 0; m < j.mips; ++m) {
            const uint32_t Sm = std::max(j.So >> m, 1u);
            if (m == 0) {
                const uint32_t T = 6u * Sm * Sm;
                for (uint32_t f = 0; f < T; f += texels) {
                    const uint32_t c = std::min(texels, T - f);
                    items.push_back({p, 0, f, c});
                    cost.push_back((double)c * (200.0 + 4.0 * j.n));  // one sample and n sums per texel
                }
                continue;
            }
            const SpecTileGeom g = spec_tile_geom(Sm);
            const uint32_t slots = 6u * g.per_face;
            uint32_t k = 0;
            while (k < slots) {
                uint32_t first = k, valid = 0;
                for (; k < slots && valid < per_item; ++k) {
                    uint32_t f, tx, ty;
                    if (spec_tile_at(g, k, f, tx, ty)) valid += 1;
                }
                if (valid) {
                    items.push_back({p, m, first, k - first});
                    cost.push_back((double)valid * kSpecTW * kSpecTH * j.n * 216.0);
                }
            }
        }
    }
    double total = 0.0;
    for (double c : cost) total += c;
    uint32_t x = 1;
    double run = 0.0;
    xcd_begin[0] = 0;
    for (uint32_t i = 0; i < (uint32_t)items.size() && x < 8; ++i) {
        run += cost[i];
        while (x < 8 && run >= total * x / 8.0) xcd_begin[x++] = i + 1;
    }
    while (x < 8) xcd_begin[x++] = (uint32_t)items.size();
    xcd_begin[8] = (uint32_t)items.size();
}

}  // namespace hvx
