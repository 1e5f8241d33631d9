// hvx_device.h — device-side vector algebra and software texture sampling.
//
// The evaluation order of every expression is fixed and matches the scalar
// definitions the path is specified by: dot(a,b) = (ax*bx + ay*by) + az*bz;
// mat*vec row r = ((m[0][r]*x + m[1][r]*y) + m[2][r]*z) + m[3][r]*w;
// normalize(v) = v * (1 / sqrt(dot(v, v))). All translation units that include
// this header are compiled with -ffp-contract=off, so nothing is fused unless a
// kernel asks for it explicitly.
//
// Texture sampling follows the Vulkan specification's texel filtering:
// LINEAR filter (u = s*W - 1/2, i0 = floor(u), alpha = frac(u)), REPEAT wrap
// for 2D images, cube face selection by major axis (ties X > Y > Z) and
// seamless cube edges (the texel across an edge is found by folding the face
// plane over the edge; a corner texel is the mean of its three neighbours).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hvx_math.h"

namespace hvx {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 make3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul3(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float length3(f3 a) { return sqrtf_(dot3(a, a)); }
__device__ __forceinline__ f3 normalize3(f3 a) {
    const float inv = 1.0f / sqrtf_(dot3(a, a));
    return mul3(a, inv);
}

// column-major 4x4 times (x, y, z, w)
__device__ __forceinline__ void matvec4(const float* m, float x, float y, float z, float w, float o[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        o[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
    }
}

// ------------------------------------------- correctly rounded fast paths
// Correctly rounded square root of x in [2^-102, FLT_MAX]: Markstein's
// iteration from the hardware reciprocal square root y ~ 1/sqrt(x):
// s0 = rn(x y), then one correction s0 + (x - s0^2) (y/2) with the residual
// fused, so it is exact. Five operations where correcting the hardware square
// root by both neighbours' residuals takes nine. Verified equal to sqrtf on
// every float of [2^-102, 2^128) on gfx950 (tests/test_gpu_fastmath.py);
// below 2^-102 the residual is subnormal and the result can be an ulp off.
__device__ __forceinline__ float sqrt_rn_normal(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float s0 = x * y;
    const float r = __builtin_fmaf(-s0, s0, x);
    return __builtin_fmaf(r, 0.5f * y, s0);
}

// Correctly rounded reciprocal of a normal b with |b| in [2^-125, 2^125]:
// hardware estimate (within 1 ulp) plus one Newton-Raphson correction with
// fused multiply-add, y' = y + y (1 - b y): the fused residual is rounded
// once, so y' carries only an O(ulp^2) error before its single rounding.
// Verified equal to 1.0f / b on every float b with |b| in [2^-125, 2^125)
// on gfx950 (tests/test_gpu_fastmath.py).
__device__ __forceinline__ float rcp_rn_normal(float b) {
    const float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    return __builtin_fmaf(e, y, y);
}

// normalize3 of a vector whose squared length is a float in [2^-102, FLT_MAX]
// (every near-unit vector): the same correctly rounded square root and
// reciprocal, by the fast paths above, so the result is bit-identical.
__device__ __forceinline__ f3 normalize3_rn(f3 a) { return mul3(a, rcp_rn_normal(sqrt_rn_normal(dot3(a, a)))); }

// ------------------------------------------------------------ cube sampling
struct CubeCoord {
    int face;
    float s, t;
};

__device__ __forceinline__ CubeCoord cube_select(f3 r) {
    const float ax = absf(r.x), ay = absf(r.y), az = absf(r.z);
    CubeCoord c;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (r.x >= 0.0f) { c.face = 0; sc = -r.z; tc = -r.y; }
        else { c.face = 1; sc = r.z; tc = -r.y; }
    } else if (ay >= az) {
        ma = ay;
        if (r.y >= 0.0f) { c.face = 2; sc = r.x; tc = r.z; }
        else { c.face = 3; sc = r.x; tc = -r.z; }
    } else {
        ma = az;
        if (r.z >= 0.0f) { c.face = 4; sc = r.x; tc = -r.y; }
        else { c.face = 5; sc = -r.x; tc = -r.y; }
    }
    c.s = 0.5f * (sc / ma + 1.0f);
    c.t = 0.5f * (tc / ma + 1.0f);
    return c;
}

// Folds coordinate c (if |c| > S) onto the plane of the major coordinate mj:
// c -> sign(c) * S, mj -> sign(mj) * (2S - |c|).
__device__ __forceinline__ bool edge_fold(int& c, int& mj, int S) {
    const int ac = c < 0 ? -c : c;
    if (ac <= S) return false;
    mj = mj > 0 ? (2 * S - ac) : -(2 * S - ac);
    c = c > 0 ? S : -S;
    return true;
}

// Texel (face, i, j) with at most one of i, j one step outside [0, S):
// returns the linear texel index of the texel across the edge.
__device__ __forceinline__ int cube_edge_index_inl(int f, int i, int j, int S) {
    const int a = 2 * i + 1 - S, b = 2 * j + 1 - S;
    // 3D point (doubled-integer units) of the texel centre on face f's plane
    int p0, p1, p2;
    switch (f) {
        case 0: p0 = S; p1 = -b; p2 = -a; break;
        case 1: p0 = -S; p1 = -b; p2 = a; break;
        case 2: p0 = a; p1 = S; p2 = b; break;
        case 3: p0 = a; p1 = -S; p2 = -b; break;
        case 4: p0 = a; p1 = -b; p2 = S; break;
        default: p0 = -a; p1 = -b; p2 = -S; break;
    }
    // fold the out-of-range coordinate onto the neighbouring face
    switch (f >> 1) {
        case 0: if (!edge_fold(p1, p0, S)) edge_fold(p2, p0, S); break;
        case 1: if (!edge_fold(p0, p1, S)) edge_fold(p2, p1, S); break;
        default: if (!edge_fold(p0, p2, S)) edge_fold(p1, p2, S); break;
    }
    const int ax = p0 < 0 ? -p0 : p0, ay = p1 < 0 ? -p1 : p1, az = p2 < 0 ? -p2 : p2;
    int nf, na, nb;
    if (ax >= ay && ax >= az) {
        if (p0 > 0) { nf = 0; na = -p2; nb = -p1; }
        else { nf = 1; na = p2; nb = -p1; }
    } else if (ay >= az) {
        if (p1 > 0) { nf = 2; na = p0; nb = p2; }
        else { nf = 3; na = p0; nb = -p2; }
    } else {
        if (p2 > 0) { nf = 4; na = p0; nb = -p1; }
        else { nf = 5; na = -p0; nb = -p1; }
    }
    const int ni = (na + S - 1) >> 1, nj = (nb + S - 1) >> 1;
    return (nf * S + nj) * S + ni;
}
__device__ __noinline__ int cube_edge_index(int f, int i, int j, int S) { return cube_edge_index_inl(f, i, j, S); }

// RGB of texel (f, i, j) of a cube level with seamless edges.
__device__ __noinline__ f3 cube_texel_seamless(const float4* __restrict__ level, int S, int f, int i, int j) {
    // taps of a finite direction are within one texel of the face
    i = i < -1 ? -1 : (i > S ? S : i);
    j = j < -1 ? -1 : (j > S ? S : j);
    const bool iin = (unsigned)i < (unsigned)S, jin = (unsigned)j < (unsigned)S;
    if (iin && jin) {
        const float4 v = level[(f * S + j) * S + i];
        return {v.x, v.y, v.z};
    }
    if (iin || jin) {
        const float4 v = level[cube_edge_index(f, i, j, S)];
        return {v.x, v.y, v.z};
    }
    const int ci = i < 0 ? 0 : S - 1, cj = j < 0 ? 0 : S - 1;
    const float4 v0 = level[(f * S + cj) * S + ci];
    const float4 va = level[cube_edge_index(f, i, cj, S)];
    const float4 vb = level[cube_edge_index(f, ci, j, S)];
    return {((v0.x + va.x) + vb.x) / 3.0f, ((v0.y + va.y) + vb.y) / 3.0f, ((v0.z + va.z) + vb.z) / 3.0f};
}

// Bilinear seamless sample of one cube level (RGB).
__device__ __forceinline__ f3 sample_cube(const float4* __restrict__ level, int S, f3 dir) {
    const CubeCoord cc = cube_select(dir);
    const float u = cc.s * (float)S - 0.5f;
    const float v = cc.t * (float)S - 0.5f;
    const float fu = floorf_(u), fv = floorf_(v);
    const float a = u - fu, b = v - fv;
    const int i0 = (int)fu, j0 = (int)fv;
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
    const float w01 = (1.0f - a) * b, w11 = a * b;
    f3 t00, t10, t01, t11;
    if (i0 >= 0 && j0 >= 0 && i0 + 1 < S && j0 + 1 < S) {
        const float4* row0 = level + (cc.face * S + j0) * S + i0;
        const float4 q00 = row0[0], q10 = row0[1], q01 = row0[S], q11 = row0[S + 1];
        t00 = {q00.x, q00.y, q00.z};
        t10 = {q10.x, q10.y, q10.z};
        t01 = {q01.x, q01.y, q01.z};
        t11 = {q11.x, q11.y, q11.z};
    } else {
        t00 = cube_texel_seamless(level, S, cc.face, i0, j0);
        t10 = cube_texel_seamless(level, S, cc.face, i0 + 1, j0);
        t01 = cube_texel_seamless(level, S, cc.face, i0, j0 + 1);
        t11 = cube_texel_seamless(level, S, cc.face, i0 + 1, j0 + 1);
    }
    return {((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x,
            ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y,
            ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z};
}

// The same bilinear seamless sample in two steps, so the texel loads of one
// sample can be in flight while other work runs: cube_taps finds the four
// taps' texel indices and weights (and issues nothing), the caller loads the
// taps, cube_combine forms the sum exactly as sample_cube does. A tap at a
// cube corner (both coordinates off the face) is the mean of three texels;
// such samples are marked and cube_combine recomputes them with sample_cube.
struct CubeTaps {
    int idx[4];  // taps (i0,j0), (i0+1,j0), (i0,j0+1), (i0+1,j0+1); a corner tap holds 0
    float w00, w10, w01, w11;
    int face, i0, j0;  // the footprint's face and first texel
    bool interior;     // all four taps on face `face` (the usual case)
    bool corner;
};

// Linear texel index of tap (f, i, j), i and j at most one step off the
// face; -1 for a corner tap.
__device__ __forceinline__ int cube_tap_index(int S, int f, int i, int j) {
    i = i < -1 ? -1 : (i > S ? S : i);
    j = j < -1 ? -1 : (j > S ? S : j);
    const bool iin = (unsigned)i < (unsigned)S, jin = (unsigned)j < (unsigned)S;
    if (iin && jin) return (f * S + j) * S + i;
    if (iin || jin) return cube_edge_index_inl(f, i, j, S);
    return -1;
}

__device__ __forceinline__ CubeTaps cube_taps(int S, f3 dir) {
    const CubeCoord cc = cube_select(dir);
    const float u = cc.s * (float)S - 0.5f;
    const float v = cc.t * (float)S - 0.5f;
    const float fu = floorf_(u), fv = floorf_(v);
    const float a = u - fu, b = v - fv;
    const int i0 = (int)fu, j0 = (int)fv;
    CubeTaps t;
    t.w00 = (1.0f - a) * (1.0f - b);
    t.w10 = a * (1.0f - b);
    t.w01 = (1.0f - a) * b;
    t.w11 = a * b;
    t.corner = false;
    t.face = cc.face;
    t.i0 = i0;
    t.j0 = j0;
    t.interior = i0 >= 0 && j0 >= 0 && i0 + 1 < S && j0 + 1 < S;
    if (t.interior) {
        const int base = (cc.face * S + j0) * S + i0;
        t.idx[0] = base;
        t.idx[1] = base + 1;
        t.idx[2] = base + S;
        t.idx[3] = base + S + 1;
    } else {
        const int di[4] = {0, 1, 0, 1}, dj[4] = {0, 0, 1, 1};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int id = cube_tap_index(S, cc.face, i0 + di[k], j0 + dj[k]);
            t.corner = t.corner || id < 0;
            t.idx[k] = id < 0 ? 0 : id;
        }
    }
    return t;
}

// The bilinear sum of loaded taps q (sample_cube's expression); a sample
// with a corner tap is resampled with sample_cube.
__device__ __forceinline__ f3 cube_combine(const float4* __restrict__ level, int S, f3 dir, const CubeTaps& t,
                                           const float4 q[4]) {
    if (t.corner) return sample_cube(level, S, dir);
    return {((t.w00 * q[0].x + t.w10 * q[1].x) + t.w01 * q[2].x) + t.w11 * q[3].x,
            ((t.w00 * q[0].y + t.w10 * q[1].y) + t.w01 * q[2].y) + t.w11 * q[3].y,
            ((t.w00 * q[0].z + t.w10 * q[1].z) + t.w01 * q[2].z) + t.w11 * q[3].z};
}

// ------------------------------------------------------------- 2D sampling
__device__ __forceinline__ int wrap_rep(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

// Bilinear REPEAT sample of an RGBA32F image (returns RGB).
__device__ __forceinline__ f3 sample_2d_rgba_repeat(const float4* __restrict__ img, int W, int H, float s, float t) {
    const float u = s * (float)W - 0.5f;
    const float v = t * (float)H - 0.5f;
    const float fu = floorf_(u), fv = floorf_(v);
    const float a = u - fu, b = v - fv;
    const int i0 = wrap_rep((int)fu, W), i1 = wrap_rep((int)fu + 1, W);
    const int j0 = wrap_rep((int)fv, H), j1 = wrap_rep((int)fv + 1, H);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
    const float w01 = (1.0f - a) * b, w11 = a * b;
    const float4 q00 = img[j0 * W + i0], q10 = img[j0 * W + i1], q01 = img[j1 * W + i0], q11 = img[j1 * W + i1];
    return {((w00 * q00.x + w10 * q10.x) + w01 * q01.x) + w11 * q11.x,
            ((w00 * q00.y + w10 * q10.y) + w01 * q01.y) + w11 * q11.y,
            ((w00 * q00.z + w10 * q10.z) + w01 * q01.z) + w11 * q11.z};
}

// Bilinear REPEAT sample of an RG32F image.
__device__ __forceinline__ float2 sample_2d_rg_repeat(const float2* __restrict__ img, int W, int H, float s, float t) {
    const float u = s * (float)W - 0.5f;
    const float v = t * (float)H - 0.5f;
    const float fu = floorf_(u), fv = floorf_(v);
    const float a = u - fu, b = v - fv;
    const int i0 = wrap_rep((int)fu, W), i1 = wrap_rep((int)fu + 1, W);
    const int j0 = wrap_rep((int)fv, H), j1 = wrap_rep((int)fv + 1, H);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
    const float w01 = (1.0f - a) * b, w11 = a * b;
    const float2 q00 = img[j0 * W + i0], q10 = img[j0 * W + i1], q01 = img[j1 * W + i0], q11 = img[j1 * W + i1];
    return {((w00 * q00.x + w10 * q10.x) + w01 * q01.x) + w11 * q11.x,
            ((w00 * q00.y + w10 * q10.y) + w01 * q01.y) + w11 * q11.y};
}

// IBL output-texel direction: texCoord = (x + .5)/D, uv = texCoord*2 - 1,
// normalize(UVToXYZ(face, uv)), then y = -y.
__device__ __forceinline__ f3 texel_dir(int face, int x, int y, int D) {
    const float u = (((float)x + 0.5f) / (float)D) * 2.0f - 1.0f;
    const float v = (((float)y + 0.5f) / (float)D) * 2.0f - 1.0f;
    f3 p;
    switch (face) {
        case 0: p = {1.0f, v, -u}; break;
        case 1: p = {-1.0f, v, u}; break;
        case 2: p = {u, -1.0f, v}; break;
        case 3: p = {u, 1.0f, -v}; break;
        case 4: p = {u, v, 1.0f}; break;
        default: p = {-u, v, -1.0f}; break;
    }
    f3 d = normalize3(p);
    d.y = -d.y;
    return d;
}

__host__ __device__ inline size_t cube_level_offset_floats(uint32_t S, uint32_t level) {
    size_t off = 0;
    for (uint32_t m = 0; m < level; ++m) {
        const size_t s = (S >> m) > 0 ? (S >> m) : 1;
        off += 6 * s * s * 4;
    }
    return off;
}

}  // namespace hvx
