// contract.hip — the dense-contraction cube filter on FP16 MFMA
// (hvx_ibl_contract_* and hvx_ibl_filter_contract of include/hvx.h).
//
// The operator is defined in hvx_contract.h. A plan holds, per output mip,
// the precomputed weight matrices of the canonical face (shared by all six
// faces), built on the GPU in FP64: the row normalisation sums are FP64
// reductions, and each row is stored in FP16 scaled by a power of two that
// puts its largest weight just under 2^15. A filter call stages the input
// radiance as FP16 scaled per face by a power of two (largest value under
// 2^15), then runs one MFMA GEMM per weight matrix:
//
//   C[rows, 6 faces x 4] = W[rows, K] * R[K, 6 faces x 4]    (FP32 accumulate)
//
// where column block f of R holds the input texels seen from output face f
// (a permutation of the level for full-sphere sums; face f's padded plane
// window for the near field). Rows are canonical output texels, so one
// weight matrix serves the six faces as six column blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "hvx_contract.h"
#include "hvx_device.h"
#include "hvx_internal.h"
#include "hvx_mfma.h"

namespace {

using hvx::ContractLevel;
using hvx::f32x4;
using hvx::f3;
using hvx::bf16x8;
using hvx::half8;
using hvx::normalize3;
using hvx::sample_cube;
using hvx::texel_dir;

enum Piece { PIECE_WHOLE = 0, PIECE_NEAR = 1, PIECE_FAR = 2 };

// Dense (full-sphere) weight matrix: rows are the canonical texels of an
// So x So face, columns every texel of a 6 x Sin x Sin level.
struct DenseGeom {
    int type, piece;
    double alpha2, t0, t1;
    uint32_t So, Sin;
    uint32_t rows, rt, ks;  // rows = So^2, row tiles of 16, k-steps of 32
};

// Near-field window of one 16 x 16 output tile, in (unpadded) texel indices
// of the face plane at the near level; woff = weight offset in half8 units.
// The tile's weights are [k-step][row tile (16)][lane (64)][8].
// The probes of one filter call: input chains and output chains (one probe:
// the by-value pointers; several: device arrays of pointers), and where probe
// p's copy of the per-call scratch lives: a scratch pointer of the plan moves
// by delta + p * pstride bytes (0 and 0 for the plan's own scratch).
struct Probes {
    const float4* chain0;
    float4* out0;
    const float4* const* chains;
    float4* const* outs;
    uint64_t delta, pstride;
    uint32_t probe_fast;  // GEMMs: grid x = probe group, y = workgroup (groups sharing weights dispatch together)
    uint32_t n;           // probes in the call
};
// (workgroup, probe) of a GEMM launch
__device__ __forceinline__ uint2 gemm_block(const Probes& Q) {
    return Q.probe_fast ? make_uint2(blockIdx.y, blockIdx.x) : make_uint2(blockIdx.x, blockIdx.y);
}
__device__ __forceinline__ const float4* probe_chain(const Probes& Q, uint32_t p) { return Q.chains ? Q.chains[p] : Q.chain0; }
__device__ __forceinline__ float4* probe_out(const Probes& Q, uint32_t p) { return Q.outs ? Q.outs[p] : Q.out0; }
template <class T> __device__ __forceinline__ T* scr(const Probes& Q, T* ptr, uint32_t p) {
    return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(ptr) + Q.delta + (uint64_t)p * Q.pstride);
}

struct NearTile {
    int32_t x0, y0;
    uint32_t w, h;
    uint64_t woff;
};

struct NearGeom {
    int type;
    double alpha2, t0, t1;
    uint32_t So, Si, te;  // output face, near level face, tiles per edge
};

// ---------------------------------------------------------- device helpers
__device__ __forceinline__ void local_dir(double u, double v, double d[3]) {
    const double r = 1.0 / sqrt(1.0 + u * u + v * v);
    d[0] = u * r;
    d[1] = v * r;
    d[2] = r;
}

// Canonical input texel k of a 6 x S x S level: direction and solid angle.
__device__ __forceinline__ void canon_input(uint32_t k, uint32_t S, double l[3], double* dw) {
    const uint32_t S2 = S * S, g = k / S2, rem = k - g * S2, y = rem / S, x = rem - y * S;
    const double u = hvx::contract_coord((int)x, S), v = hvx::contract_coord((int)y, S);
    double p[3];
    hvx::contract_face_point((int)g, u, v, p);
    const double r = 1.0 / sqrt(1.0 + u * u + v * v);
    l[0] = p[0] * r;
    l[1] = p[1] * r;
    l[2] = p[2] * r;
    *dw = hvx::contract_dw(u, v, S);
}

__device__ __forceinline__ double piece_weight(int type, int piece, double alpha2, double t0, double t1, double c,
                                               double dw) {
    double w = hvx::contract_weight(type, alpha2, c);
    if (w == 0.0) return 0.0;
    w *= dw;
    if (piece == PIECE_NEAR) w *= hvx::contract_phi(t0, t1, c);
    else if (piece == PIECE_FAR) w *= 1.0 - hvx::contract_phi(t0, t1, c);
    return w;
}

__device__ __forceinline__ void dense_row_dir(const DenseGeom& G, uint32_t row, double n[3]) {
    const uint32_t y = row / G.So, x = row - y * G.So;
    local_dir(hvx::contract_coord((int)x, G.So), hvx::contract_coord((int)y, G.So), n);
}

__device__ __forceinline__ double dense_w(const DenseGeom& G, const double n[3], uint32_t k) {
    double l[3], dw;
    canon_input(k, G.Sin, l, &dw);
    const double c = (n[0] * l[0] + n[1] * l[1]) + n[2] * l[2];
    return piece_weight(G.type, G.piece, G.alpha2, G.t0, G.t1, c, dw);
}

__device__ __forceinline__ double near_w(const NearGeom& G, const double n[3], int px, int py) {
    const double u = hvx::contract_coord(px, G.Si), v = hvx::contract_coord(py, G.Si);
    double l[3];
    local_dir(u, v, l);
    const double c = (n[0] * l[0] + n[1] * l[1]) + n[2] * l[2];
    return piece_weight(G.type, PIECE_NEAR, G.alpha2, G.t0, G.t1, c, hvx::contract_dw(u, v, G.Si));
}

// Power-of-two exponent s with w_max * 2^s in (2^14, 2^15].
__host__ __device__ __forceinline__ int row_exponent(double wmax) {
    if (!(wmax > 0.0)) return 0;
    int e;
    (void)frexp(wmax, &e);  // wmax = m 2^e, m in [0.5, 1)
    return 15 - e;
}

// Deterministic block sum and max (256 threads) in double.
__device__ void block_sum_max(double s, double m, double* sh_s, double* sh_m, double* out_s, double* out_m) {
    const uint32_t t = threadIdx.x;
    sh_s[t] = s;
    sh_m[t] = m;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (t < h) {
            sh_s[t] = sh_s[t] + sh_s[t + h];
            sh_m[t] = sh_m[t] < sh_m[t + h] ? sh_m[t + h] : sh_m[t];
        }
        __syncthreads();
    }
    if (t == 0) {
        *out_s = sh_s[0];
        *out_m = sh_m[0];
    }
}

// ------------------------------------------------------ plan: dense weights
__global__ __launch_bounds__(256) void dense_row_kernel(DenseGeom G, double* __restrict__ Z,
                                                        double* __restrict__ wmax) {
    __shared__ double sh_s[256], sh_m[256];
    const uint32_t row = blockIdx.x;
    double n[3];
    dense_row_dir(G, row, n);
    const uint32_t K = 6 * G.Sin * G.Sin;
    double s = 0.0, m = 0.0;
    for (uint32_t k = threadIdx.x; k < K; k += 256) {
        const double w = dense_w(G, n, k);
        s += w;
        m = m < w ? w : m;
    }
    block_sum_max(s, m, sh_s, sh_m, &Z[row], &wmax[row]);
}

// SPLIT: BF16 hi at W, lo = bf16(v - hi) at W + lo_off (16-byte units)
template <bool SPLIT>
__global__ __launch_bounds__(256) void dense_fill_kernel(DenseGeom G, const double* __restrict__ wmax,
                                                         half8* __restrict__ W, uint64_t lo_off) {
    const uint64_t id = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= (uint64_t)G.rt * G.ks * 64) return;
    const uint32_t lane = (uint32_t)(id & 63);
    const uint64_t t = id >> 6;
    const uint32_t ks = (uint32_t)(t % G.ks), rt = (uint32_t)(t / G.ks);
    const uint32_t row = rt * 16 + (lane & 15), k0 = ks * 32 + 8 * (lane >> 4);
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (row < G.rows) {
        double n[3];
        dense_row_dir(G, row, n);
        const double sc = ldexp(1.0, row_exponent(wmax[row]));
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = dense_w(G, n, k0 + q) * sc;
    }
    if constexpr (SPLIT) {
        bf16x8 hi, lo;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            hi[q] = (__bf16)(float)v[q];
            lo[q] = (__bf16)(float)(v[q] - (double)(float)hi[q]);
        }
        reinterpret_cast<bf16x8*>(W)[id] = hi;
        reinterpret_cast<bf16x8*>(W)[id + lo_off] = lo;
    } else {
        half8 out;
#pragma unroll
        for (int q = 0; q < 8; ++q) out[q] = (_Float16)(float)v[q];
        W[id] = out;
    }
}

// rowscale = 2^-s / Z (normalised rows) or 2^-s (raw far rows; then the
// far normaliser goes to a six-face map, identical on every face).
__global__ __launch_bounds__(256) void dense_scale_kernel(uint32_t rows, const double* __restrict__ Z,
                                                          const double* __restrict__ wmax, int raw,
                                                          float* __restrict__ rowscale, float4* __restrict__ zmap) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    const double p = ldexp(1.0, -row_exponent(wmax[r]));
    rowscale[r] = raw ? (float)p : (float)(p / Z[r]);
    if (raw)
        for (uint32_t f = 0; f < 6; ++f) zmap[f * rows + r] = make_float4((float)Z[r], 0.0f, 0.0f, 0.0f);
}

// -------------------------------------------------------- plan: near weights
__global__ __launch_bounds__(256) void near_row_kernel(NearGeom G, const NearTile* __restrict__ tiles,
                                                       double* __restrict__ Z, double* __restrict__ wmax) {
    __shared__ double sh_s[256], sh_m[256];
    const uint32_t o = blockIdx.x, Y = o / G.So, X = o - Y * G.So;
    const NearTile T = tiles[(Y / 16) * G.te + X / 16];
    double n[3];
    local_dir(hvx::contract_coord((int)X, G.So), hvx::contract_coord((int)Y, G.So), n);
    const uint32_t K = T.w * T.h;
    double s = 0.0, m = 0.0;
    for (uint32_t k = threadIdx.x; k < K; k += 256) {
        const uint32_t r = k / T.w, c = k - r * T.w;
        const double w = near_w(G, n, T.x0 + (int)c, T.y0 + (int)r);
        s += w;
        m = m < w ? w : m;
    }
    block_sum_max(s, m, sh_s, sh_m, &Z[o], &wmax[o]);
}

// grid: tiles x 16 row tiles; each block writes one row tile's k-steps.
// SPLIT: the scaled weight v is stored as BF16 hi = bf16(v) at W and
// lo = bf16(v - hi) at W + lo_off (hi + lo carries ~16 significant bits).
template <bool SPLIT>
__global__ __launch_bounds__(256) void near_fill_kernel(NearGeom G, const NearTile* __restrict__ tiles,
                                                        const double* __restrict__ wmax, half8* __restrict__ W,
                                                        uint64_t lo_off) {
    const uint32_t tile = blockIdx.x / 16, rt = blockIdx.x % 16;
    const NearTile T = tiles[tile];
    const uint32_t KS = T.h * (T.w / 32);
    const uint32_t X0 = (tile % G.te) * 16, Y = (tile / G.te) * 16 + rt;
    for (uint32_t idx = threadIdx.x; idx < KS * 64; idx += 256) {
        const uint32_t ks = idx >> 6, lane = idx & 63;
        const uint32_t X = X0 + (lane & 15), o = Y * G.So + X;
        double n[3];
        local_dir(hvx::contract_coord((int)X, G.So), hvx::contract_coord((int)Y, G.So), n);
        const double sc = ldexp(1.0, row_exponent(wmax[o]));
        const uint32_t k0 = ks * 32 + 8 * (lane >> 4), r = k0 / T.w, c0 = k0 - r * T.w;
        const uint64_t at = T.woff + ((uint64_t)ks * 16 + rt) * 64 + lane;  // K-step major
        if constexpr (SPLIT) {
            bf16x8 hi, lo;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double v = near_w(G, n, T.x0 + (int)(c0 + q), T.y0 + (int)r) * sc;
                hi[q] = (__bf16)(float)v;
                lo[q] = (__bf16)(float)(v - (double)(float)hi[q]);
            }
            reinterpret_cast<bf16x8*>(W)[at] = hi;
            reinterpret_cast<bf16x8*>(W)[at + lo_off] = lo;
        } else {
            half8 out;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                out[q] = (_Float16)(float)(near_w(G, n, T.x0 + (int)(c0 + q), T.y0 + (int)r) * sc);
            W[at] = out;
        }
    }
}

// Z(o) = Z_near(o) + far normaliser interpolated at o (any face: the far
// normaliser map is the same on every face).
__global__ __launch_bounds__(256) void near_scale_kernel(NearGeom G, uint32_t Sf, const double* __restrict__ Z,
                                                         const double* __restrict__ wmax,
                                                         const float4* __restrict__ zmap,
                                                         float* __restrict__ rowscale, float* __restrict__ invZ) {
    const uint32_t o = blockIdx.x * 256 + threadIdx.x;
    if (o >= G.So * G.So) return;
    const uint32_t Y = o / G.So, X = o - Y * G.So;
    const f3 zf = sample_cube(zmap, (int)Sf, texel_dir(0, (int)X, (int)Y, (int)G.So));
    const double Zt = Z[o] + (double)zf.x;
    rowscale[o] = (float)(ldexp(1.0, -row_exponent(wmax[o])) / Zt);
    invZ[o] = (float)(1.0 / Zt);
}

// ------------------------------------------------------ per-call staging
// A filter call is eight launches: face maxima, exponent scales, all staging
// (every dense level and padded plane), all dense GEMMs, all their
// epilogues, all near-field tiles, all far-field additions, and the
// roughness-0 mip. Each fused launch finds its job by running offsets in a
// table passed by value.
constexpr int kMaxJobs = 2 * hvx::kContractMaxMips;
constexpr uint32_t kFaceMaxBlocks = 32;
constexpr uint32_t kMaxProbes = 4096;  // probes of one multi-probe call (grid dimension y)
#ifndef HVX_PROBE_GROUP
#define HVX_PROBE_GROUP 4
#endif
constexpr int kProbeGroup = HVX_PROBE_GROUP;  // probes per wave in multi-probe GEMMs

// partial[f][b]: largest RGB magnitude of block b's share of face f
__global__ __launch_bounds__(256) void facemax_kernel(Probes Q, size_t level_off, uint32_t S, float* partial) {
    __shared__ float sh[256];
    const uint32_t f = blockIdx.y, S2 = S * S, pr = blockIdx.z;
    const float4* level = probe_chain(Q, pr) + level_off;
    partial = scr(Q, partial, pr);
    float m = 0.0f;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < S2; i += gridDim.x * 256) {
        const float4 v = level[(size_t)f * S2 + i];
        const float a = fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fabsf(v.z));
        m = a > m ? a : m;  // NaN never replaces m
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[f * gridDim.x + blockIdx.x] = sh[0];
}

// sc[f] = 2^-e_f, sc[6+f] = 2^e_f for the padded plane of face f (bounded by
// face f and its four neighbours); sc[12] = 2^-e, sc[13] = 2^e over all faces.
__global__ __launch_bounds__(64) void scales_kernel(Probes Q, const float* partial, uint32_t nb, float* sc) {
    __shared__ float fm[6];
    const uint32_t t = threadIdx.x;
    partial = scr(Q, partial, blockIdx.x);
    sc = scr(Q, sc, blockIdx.x);
    if (t < 6) {
        float m = 0.0f;
        for (uint32_t b = 0; b < nb; ++b) m = fmaxf(m, partial[t * nb + b]);
        fm[t] = m;
    }
    __syncthreads();
    if (t > 6) return;
    float m = 0.0f;
    for (uint32_t g = 0; g < 6; ++g)
        if (t == 6 || g != (t ^ 1u)) m = fmaxf(m, fm[g]);
    int e = 0;
    if (m > 0.0f && m <= 3.0e38f) {
        int ex;
        (void)frexpf(m, &ex);
        e = ex - 15;
    }
    if (t < 6) {
        sc[t] = ldexpf(1.0f, -e);
        sc[6 + t] = ldexpf(1.0f, e);
    } else {
        sc[12] = ldexpf(1.0f, -e);
        sc[13] = ldexpf(1.0f, e);
    }
}

struct StageItem {
    int kind;  // 0: a dense level (permuted), 1: padded face planes; 3, 2: the same as BF16 hi + lo
    uint32_t S, B, Pw;
    size_t level_off;  // float4 offset of the level in the input chain
    const int32_t* perm;
    half8* dst;
    uint32_t begin;  // first thread of the item (kind 2: lo parts at dst + the item's thread count)
};
struct StageTable {
    StageItem it[kMaxJobs];
    uint32_t n, total;
};

// Writes channel values v[3] of element q of a staged 8-element operand at
// unit `at` (and the next two units) as FP16, or as BF16 hi + lo with the lo
// parts `lo_off` units further.
template <bool SPLIT>
__device__ __forceinline__ void stage_store3(half8* __restrict__ D, size_t at, uint32_t q, const float v[3],
                                             size_t lo_off) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t e = (at + c) * 8 + q;
        if constexpr (SPLIT) {
            const __bf16 hi = (__bf16)v[c];
            reinterpret_cast<__bf16*>(D)[e] = hi;
            reinterpret_cast<__bf16*>(D)[lo_off * 8 + e] = (__bf16)(v[c] - (float)hi);
        } else {
            reinterpret_cast<_Float16*>(D)[e] = (_Float16)v[c];
        }
    }
}

// R[k/8][col][8]: col = 4 f + c holds channel c (< 3) of the level texel
// that output face f sees as canonical input k; other columns are zero (from
// plan creation). One thread per (k, f); lanes 8 apart share k / 8.
template <bool SPLIT>
__device__ __forceinline__ void stage_dense_body(const float4* __restrict__ level, uint32_t Sin,
                                                 const int32_t* __restrict__ perm, float s, half8* __restrict__ R,
                                                 uint32_t id) {
    const uint32_t K = 6 * Sin * Sin;
    const uint32_t q = id & 7, rest = id >> 3, f = rest % 6, k8 = rest / 6;
    const float4 t = level[perm[(size_t)f * K + k8 * 8 + q]];
    const float v[3] = {t.x * s, t.y * s, t.z * s};
    stage_store3<SPLIT>(R, (size_t)k8 * 32 + 4 * f, q, v, (size_t)K / 8 * 32);
}

// P[py][px/8][col][8] over the padded plane (Pw x Pw, texel index -B .. Si+B)
// of each face: inside the face the texel itself, outside the bilinear
// seamless sample of the level at the plane point's direction. One thread
// per (texel, face); columns 4 f + 3 and 24..31 are zero from plan creation.
template <bool SPLIT>
__device__ __forceinline__ void stage_pad_body(const float4* __restrict__ level, uint32_t Si, uint32_t B, uint32_t Pw,
                                               const float* __restrict__ sc, half8* __restrict__ P, uint32_t id) {
    const uint32_t Pw8 = Pw / 8;
    const uint32_t q = id & 7, rest = id >> 3, f = rest % 6, cell = rest / 6, px8 = cell % Pw8, py = cell / Pw8;
    const int x = (int)(px8 * 8 + q) - (int)B, y = (int)py - (int)B;
    f3 t;
    if (x >= 0 && y >= 0 && x < (int)Si && y < (int)Si) {
        const float4 v = level[((size_t)f * Si + y) * Si + x];
        t = f3{v.x, v.y, v.z};
    } else {
        double p[3];
        hvx::contract_face_point((int)f, hvx::contract_coord(x, Si), hvx::contract_coord(y, Si), p);
        t = sample_cube(level, (int)Si, normalize3(f3{(float)p[0], (float)p[1], (float)p[2]}));
    }
    const float sf = sc[f];
    const float v[3] = {t.x * sf, t.y * sf, t.z * sf};
    stage_store3<SPLIT>(P, (size_t)cell * 32 + 4 * f, q, v, (size_t)Pw * Pw8 * 32);
}

__global__ __launch_bounds__(256) void stage_all_kernel(Probes Q, StageTable T, const float* sc) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x, pr = blockIdx.y;
    if (id >= T.total) return;
    uint32_t j = 0;
    while (j + 1 < T.n && id >= T.it[j + 1].begin) ++j;
    const StageItem& I = T.it[j];
    const float4* level = probe_chain(Q, pr) + I.level_off;
    sc = scr(Q, sc, pr);
    half8* dst = scr(Q, I.dst, pr);
    if (I.kind == 0) stage_dense_body<false>(level, I.S, I.perm, sc[12], dst, id - I.begin);
    else if (I.kind == 3) stage_dense_body<true>(level, I.S, I.perm, sc[12], dst, id - I.begin);
    else if (I.kind == 1) stage_pad_body<false>(level, I.S, I.B, I.Pw, sc, dst, id - I.begin);
    else stage_pad_body<true>(level, I.S, I.B, I.Pw, sc, dst, id - I.begin);
}

// ----------------------------------------------------------- MFMA GEMMs
struct DenseJob {
    const half8* W;
    const half8* R;
    uint32_t rt, ks, splitk, rows;
    float* part;            // [splitk][rt * 16][32]
    const float* rowscale;
    float4* out;            // plan-owned far map, or nullptr: the output chain at out_off
    size_t out_off;
    float alpha;
    uint32_t wave_begin, epi_begin;
    uint64_t w_lo, r_lo;  // split BF16 (w_lo != 0): offsets (16-byte units) of the lo parts of W and R
};
struct DenseTable {
    DenseJob j[kMaxJobs];
    uint32_t n, waves, epi;
};

// One K step of a wave's RT (1 or 2) row tiles x 32 columns: the weight
// fragments of the row tiles (x[0..RT-1]; split: hi, then lo in
// x[RT..2RT-1]) and PG probes' two 16-column B fragments (b[q][0..1]; split:
// hi, then lo in b[q][2..3]).
template <bool SPLIT, int PG, int RT = 2>
struct KStep {
    using V = std::conditional_t<SPLIT, bf16x8, half8>;
    static constexpr int kW = SPLIT ? 2 * RT : RT, kB = SPLIT ? 4 : 2;
    V x[kW];
    V b[PG][kB];
};

// acc[q][2 * h + nb] += row tile h x column block nb; split: hi*hi + hi*lo +
// lo*hi (the lo*lo term is below the FP32 accumulation's resolution)
template <bool SPLIT, int PG, int RT>
__device__ __forceinline__ void kstep_mac(f32x4 (&acc)[PG][2 * RT], const KStep<SPLIT, PG, RT>& o) {
#pragma unroll
    for (int q = 0; q < PG; ++q) {
#pragma unroll
        for (int h = 0; h < RT; ++h)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                f32x4& a = acc[q][2 * h + nb];
                a = hvx::mfma16x16x32(o.x[h], o.b[q][nb], a);
                if constexpr (SPLIT) {
                    a = hvx::mfma16x16x32(o.x[h], o.b[q][2 + nb], a);
                    a = hvx::mfma16x16x32(o.x[RT + h], o.b[q][nb], a);
                }
            }
    }
}

// Runs n >= 1 K steps as a software pipeline over groups of U steps: a
// group's loads are issued before the MFMAs of the group loaded before it.
// load_next(KStep&) loads the next step and advances a cursor that stops at
// the last step (the group loaded ahead of the final one repeats it, unused).
// The accumulation order is that of a plain loop over the steps.
template <uint32_t U, bool SPLIT, int PG, int RT, class LoadNext>
__device__ __forceinline__ void kstep_pipeline(f32x4 (&acc)[PG][2 * RT], uint32_t n, LoadNext load_next) {
    KStep<SPLIT, PG, RT> cur[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) load_next(cur[u]);
    uint32_t s = 0;
    for (; s + U <= n; s += U) {
        KStep<SPLIT, PG, RT> nx[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) load_next(nx[u]);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) kstep_mac<SPLIT, PG, RT>(acc, cur[u]);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) cur[u] = nx[u];
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u)
        if (s + u < n) kstep_mac<SPLIT, PG, RT>(acc, cur[u]);
}

#ifndef HVX_KSTEP_UNROLL
#define HVX_KSTEP_UNROLL 8
#endif
// FP16 K steps whose loads are in flight together (half as many split-BF16
// steps, which load twice the fragments): the near field's largest mip runs
// about two waves per SIMD, too few to stream the weights otherwise
constexpr uint32_t kStepUnroll = HVX_KSTEP_UNROLL;
template <bool SPLIT>
constexpr uint32_t kstep_unroll() { return SPLIT ? (kStepUnroll + 1) / 2 : kStepUnroll; }
// The single-probe near GEMM's own depths (FP16, split BF16). Its waves
// (one row tile each) outnumber what the chip holds, so it gains more from
// loads in flight per wave than from waves per SIMD. Whole FP16 call on the
// 512^2 GGX chain: 4 steps (6 waves per SIMD) 0.605 ms, 8 (4 waves) 0.58,
// 10 0.552, 12 (3 waves) 0.536 - 0.546, 14 0.541, 16 (2 waves) 0.549, 20
// spills; split BF16: 4 steps 0.943 - 0.999, 5 0.913 - 0.927, 6 0.93, 8 0.908.
#ifndef HVX_NEAR_UNROLL
#define HVX_NEAR_UNROLL 12
#endif
#ifndef HVX_NEAR_UNROLL_SPLIT
#define HVX_NEAR_UNROLL_SPLIT 5
#endif
template <bool SPLIT>
constexpr uint32_t near_unroll() { return SPLIT ? HVX_NEAR_UNROLL_SPLIT : HVX_NEAR_UNROLL; }

// Dense: a wave owns two row tiles (32 rows) x 32 columns over one K slice;
// partial sums go to part[slice][row][32] and are reduced in slice order.
// PG probes per wave (grid dimension: probe group): each weight fragment is
// loaded once and multiplied into every probe of the group; each probe's
// accumulation is the same sequence as with PG = 1.
// Wave w of a job: four consecutive waves take four row-tile pairs of one K
// slice (a job's pairs are padded to a multiple of four), so a workgroup's
// waves share the slice's B operand.
__device__ __forceinline__ void dense_wave(const DenseJob& J, uint32_t w, uint32_t& pair, uint32_t& slice) {
    const uint32_t per = 4 * J.splitk, grp = w / per, r = w - grp * per;
    slice = r / 4;
    pair = grp * 4 + (r & 3);
}

template <int PG>
__global__ __launch_bounds__(256) void dense_gemm_kernel(Probes Q, DenseTable T) {
    const uint2 wb = gemm_block(Q);
    const uint32_t gw = wb.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63, p0 = wb.y * PG;
    if (gw >= T.waves) return;  // uniform per wave; the kernel has no barriers
    const uint32_t np = min((uint32_t)PG, Q.n - p0);
    uint32_t jj = 0;
    while (jj + 1 < T.n && gw >= T.j[jj + 1].wave_begin) ++jj;
    const DenseJob& J = T.j[jj];
    const uint32_t rt = J.rt, ks = J.ks, splitk = J.splitk;
    uint32_t pair, slice;
    dense_wave(J, gw - J.wave_begin, pair, slice);
    const uint32_t rt0 = 2 * pair;
    if (rt0 >= rt) return;  // padding wave (uniform per wave)
    const bool has1 = rt0 + 1 < rt;
    const uint32_t kb = (uint32_t)((uint64_t)slice * ks / splitk), ke = (uint32_t)((uint64_t)(slice + 1) * ks / splitk);
    const uint32_t i = lane & 15, g = lane >> 4;
    const half8* w0 = J.W + (size_t)rt0 * ks * 64 + lane;
    const half8* w1 = has1 ? w0 + (size_t)ks * 64 : w0;
    const half8* rb[PG];
#pragma unroll
    for (int q = 0; q < PG; ++q) rb[q] = scr(Q, J.R, p0 + ((uint32_t)q < np ? q : 0)) + g * 32 + i;
    f32x4 acc[PG][4];
#pragma unroll
    for (int q = 0; q < PG; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[q][k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // steps [kb, ke) of the slice; an empty slice (ke == kb) loads step kb,
    // which exists (kb < ks), and accumulates nothing
    auto run = [&](auto split_tag) {
        constexpr bool SPLIT = decltype(split_tag)::value;
        using Op = KStep<SPLIT, PG>;
        using V = typename Op::V;
        using GV = const __attribute__((address_space(1))) V*;
        const GV v0 = (GV)reinterpret_cast<const V*>(w0), v1 = (GV)reinterpret_cast<const V*>(w1);
        GV rv[PG];
#pragma unroll
        for (int q = 0; q < PG; ++q) rv[q] = (GV)reinterpret_cast<const V*>(rb[q]);
        uint32_t ls = kb;
        kstep_pipeline<kstep_unroll<SPLIT>(), SPLIT, PG, 2>(acc, ke - kb, [&](Op& o) {
            const size_t wo = (size_t)ls * 64, qo = (size_t)ls * 128;
            o.x[0] = v0[wo];
            o.x[1] = v1[wo];
            if constexpr (SPLIT) {
                o.x[2] = v0[wo + J.w_lo];
                o.x[3] = v1[wo + J.w_lo];
            }
#pragma unroll
            for (int q = 0; q < PG; ++q) {
                o.b[q][0] = rv[q][qo];
                o.b[q][1] = rv[q][qo + 16];
                if constexpr (SPLIT) {
                    o.b[q][2] = rv[q][qo + J.r_lo];
                    o.b[q][3] = rv[q][qo + 16 + J.r_lo];
                }
            }
            ls += ls + 1 < ke ? 1u : 0u;
        });
    };
    if (J.w_lo)  // split BF16
        run(std::true_type{});
    else
        run(std::false_type{});
#pragma unroll
    for (int q = 0; q < PG; ++q) {
        if ((uint32_t)q >= np) break;
        float* p = scr(Q, J.part, p0 + q) + (size_t)slice * rt * 16 * 32;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t row0 = rt0 * 16 + 4 * g + r;
            p[(size_t)row0 * 32 + i] = acc[q][0][r];
            p[(size_t)row0 * 32 + 16 + i] = acc[q][1][r];
            if (has1) {
                p[(size_t)(row0 + 16) * 32 + i] = acc[q][2][r];
                p[(size_t)(row0 + 16) * 32 + 16 + i] = acc[q][3][r];
            }
        }
    }
}

// The grouped dense GEMM with each K step's B operand (PG probes' 4 x 32
// half8, plus the lo parts when split) staged once per workgroup in LDS,
// double buffered, as near_gemm_staged_kernel does: the workgroup's four
// waves share the K slice. Padding waves stage and synchronise but compute
// nothing.
template <int PG>
__global__ __launch_bounds__(256) void dense_gemm_staged_kernel(Probes Q, DenseTable T) {
    const uint2 wb = gemm_block(Q);
    const uint32_t gw0 = wb.x * 4;
    if (gw0 >= T.waves) return;  // uniform per workgroup: jobs hold multiples of four waves
    const uint32_t t = threadIdx.x, gw = gw0 + (t >> 6), lane = t & 63, p0 = wb.y * PG;
    const uint32_t np = min((uint32_t)PG, Q.n - p0);
    uint32_t jj = 0;
    while (jj + 1 < T.n && gw0 >= T.j[jj + 1].wave_begin) ++jj;
    const DenseJob& J = T.j[jj];
    const uint32_t rt = J.rt, ks = J.ks, splitk = J.splitk;
    uint32_t pair, slice;
    dense_wave(J, gw - J.wave_begin, pair, slice);
    const uint32_t rt0 = 2 * pair;
    const bool active = rt0 < rt, has1 = rt0 + 1 < rt;
    const uint32_t kb = (uint32_t)((uint64_t)slice * ks / splitk), ke = (uint32_t)((uint64_t)(slice + 1) * ks / splitk);
    const uint32_t i = lane & 15, g = lane >> 4;
    const half8* w0 = J.W + (active ? (size_t)rt0 * ks * 64 + lane : 0);
    const half8* w1 = has1 ? w0 + (size_t)ks * 64 : w0;
    const half8* Rp[PG];
#pragma unroll
    for (int q = 0; q < PG; ++q) Rp[q] = scr(Q, J.R, p0 + ((uint32_t)q < np ? q : 0));
    constexpr uint32_t kMaxB = PG * 128 * 2;  // half8 per step, split
    __shared__ half8 sB[2][kMaxB];
    // the B operand is read through global (not generic) pointers: generic
    // loads count against the LDS counter too, so each step's LDS reads
    // would wait for the next step's B loads
    using GH = const __attribute__((address_space(1))) half8*;
    GH Rg[PG];
#pragma unroll
    for (int q = 0; q < PG; ++q) Rg[q] = (GH)Rp[q];
    f32x4 acc[PG][4];
#pragma unroll
    for (int q = 0; q < PG; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[q][k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // the job's precision is uniform per workgroup; each variant has its own
    // loop so that its loads are unconditional (conditional loads make the
    // waits conservative)
    auto run = [&](auto split_tag) {
        constexpr bool SPLIT = decltype(split_tag)::value;
        constexpr uint32_t nB = PG * 128 * (SPLIT ? 2u : 1u), kPer = (nB + 255) / 256;
        // element e = t + 256 k of a step's block: lo part e / (PG * 128),
        // probe (e / 128) % PG, unit e % 128; with whole 256-element passes
        // and an even PG the part is the same for the whole pass
        constexpr bool kWhole = nB % 256 == 0 && (PG * 128) % 256 == 0;
        auto b_load = [&](uint32_t s, half8 (&v)[kPer]) {
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k) {
                const uint32_t e = t + 256u * k, q = (e / 128) % PG, o = e % 128;
                if constexpr (kWhole) {
                    const uint32_t lo = (256u * k) / (PG * 128);
                    v[k] = Rg[q][(size_t)s * 128 + o + (lo ? J.r_lo : 0)];
                } else {
                    const uint32_t lo = e / (PG * 128);
                    if (e < nB) v[k] = Rg[q][(size_t)s * 128 + o + (lo ? J.r_lo : 0)];
                }
            }
        };
        auto b_store = [&](uint32_t buf, const half8 (&v)[kPer]) {
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k)
                if (kWhole || t + 256u * k < nB) sB[buf][t + 256u * k] = v[k];
        };
        half8 nx[kPer];
        if (kb < ke) {
            b_load(kb, nx);
            b_store(0, nx);
        }
        __syncthreads();
        for (uint32_t s = kb; s < ke; ++s) {
            const bool more = s + 1 < ke;
            // this step's weights are loaded (every wave; an inactive wave's
            // pointers are the job's first row) before the next step's B
            // operand: the load counter completes in order, so the MFMAs'
            // wait for the weights leaves the B loads in flight
            const size_t o = (size_t)s * 64;
            half8 xw[SPLIT ? 4 : 2];
            xw[0] = w0[o];
            xw[1] = w1[o];
            if constexpr (SPLIT) {
                xw[2] = w0[o + J.w_lo];
                xw[3] = w1[o + J.w_lo];
            }
            b_load(more ? s + 1 : s, nx);       // the last step reloads itself
            __asm__ volatile("" ::: "memory");  // keeps the B loads ahead of the MFMAs
            const half8* sb = sB[(s - kb) & 1];
            if (active) {
                if constexpr (SPLIT) {
                    const bf16x8 x0 = __builtin_bit_cast(bf16x8, xw[0]), x1 = __builtin_bit_cast(bf16x8, xw[1]);
                    const bf16x8 y0 = __builtin_bit_cast(bf16x8, xw[2]), y1 = __builtin_bit_cast(bf16x8, xw[3]);
#pragma unroll
                    for (int q = 0; q < PG; ++q) {
                        const bf16x8* rv = reinterpret_cast<const bf16x8*>(sb + q * 128 + g * 32 + i);
                        const bf16x8 b0 = rv[0], b1 = rv[16], c0 = rv[PG * 128], c1 = rv[PG * 128 + 16];
                        acc[q][0] = hvx::mfma16x16x32(x0, b0, acc[q][0]);
                        acc[q][0] = hvx::mfma16x16x32(x0, c0, acc[q][0]);
                        acc[q][0] = hvx::mfma16x16x32(y0, b0, acc[q][0]);
                        acc[q][1] = hvx::mfma16x16x32(x0, b1, acc[q][1]);
                        acc[q][1] = hvx::mfma16x16x32(x0, c1, acc[q][1]);
                        acc[q][1] = hvx::mfma16x16x32(y0, b1, acc[q][1]);
                        acc[q][2] = hvx::mfma16x16x32(x1, b0, acc[q][2]);
                        acc[q][2] = hvx::mfma16x16x32(x1, c0, acc[q][2]);
                        acc[q][2] = hvx::mfma16x16x32(y1, b0, acc[q][2]);
                        acc[q][3] = hvx::mfma16x16x32(x1, b1, acc[q][3]);
                        acc[q][3] = hvx::mfma16x16x32(x1, c1, acc[q][3]);
                        acc[q][3] = hvx::mfma16x16x32(y1, b1, acc[q][3]);
                    }
                } else {
                    const half8 x0 = xw[0], x1 = xw[1];
#pragma unroll
                    for (int q = 0; q < PG; ++q) {
                        const half8* rv = sb + q * 128 + g * 32 + i;
                        const half8 b0 = rv[0], b1 = rv[16];
                        acc[q][0] = hvx::mfma16x16x32(x0, b0, acc[q][0]);
                        acc[q][1] = hvx::mfma16x16x32(x0, b1, acc[q][1]);
                        acc[q][2] = hvx::mfma16x16x32(x1, b0, acc[q][2]);
                        acc[q][3] = hvx::mfma16x16x32(x1, b1, acc[q][3]);
                    }
                }
            }
            // stored unconditionally: after the last step the buffer is no
            // longer read (its readers passed the previous barrier), and a
            // store under a branch would draw the B loads into the branch,
            // after the MFMAs
            b_store((s + 1 - kb) & 1, nx);
            __syncthreads();
        }
    };
    if (J.w_lo != 0)  // split BF16
        run(std::true_type{});
    else
        run(std::false_type{});
    if (!active) return;
#pragma unroll
    for (int q = 0; q < PG; ++q) {
        if ((uint32_t)q >= np) break;
        float* p = scr(Q, J.part, p0 + q) + (size_t)slice * rt * 16 * 32;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t row0 = rt0 * 16 + 4 * g + r;
            p[(size_t)row0 * 32 + i] = acc[q][0][r];
            p[(size_t)row0 * 32 + 16 + i] = acc[q][1][r];
            if (has1) {
                p[(size_t)(row0 + 16) * 32 + i] = acc[q][2][r];
                p[(size_t)(row0 + 16) * 32 + 16 + i] = acc[q][3][r];
            }
        }
    }
}

// The dense GEMM with the K slice's B blocks staged in LDS: the four waves
// of a workgroup (four row-tile pairs of one K slice) share each step's B
// block (128 half8 per probe, twice that split), loaded once per workgroup,
// GS steps at a time (GS / 2 split), double buffered, one barrier per group;
// weights a group ahead. PG probes per wave (grouped multi-probe calls):
// each weight fragment is loaded once for the group. Each output's
// accumulation is the same sequence of MFMAs as in dense_gemm_kernel, so the
// partial sums are identical. Single probe: its B block is a third of a
// step's loads, not two thirds as in the near field, so the gain is small:
// whole FP16 call on the 512^2 GGX chain 0.469 ms with dense_gemm_kernel,
// GS = 4: 0.466, 8: 0.461 (split BF16 0.767, 0.758, 0.759). Four probes per
// wave (54 probes 512^2 -> 256^2 x 9, FP16 / split): dense_gemm_staged_kernel
// (a barrier per step) 4.36 / 6.16 ms; GS = 2: 3.96 / 6.00; 4: 4.56 / 6.50
// (one wave per SIMD).
#ifndef HVX_DENSE_LDS_GROUP
#define HVX_DENSE_LDS_GROUP 8
#endif
#ifndef HVX_DENSE_LDS_GROUP_MULTI
#define HVX_DENSE_LDS_GROUP_MULTI 2
#endif
// non-temporal weight loads in the dense GEMM: no gain for one probe (0.451
// -> 0.449 ms), and a loss when probe groups share each weight block through
// the caches (54 probes 512^2: 3.98 -> 4.27 ms)
#ifndef HVX_DENSE_NT
#define HVX_DENSE_NT 0
#endif
constexpr int kDenseLdsGroup = HVX_DENSE_LDS_GROUP;             // 0: dense_gemm_kernel
constexpr int kDenseLdsGroupMulti = HVX_DENSE_LDS_GROUP_MULTI;  // 0: dense_gemm_staged_kernel
template <int PG, int GS>
__global__ __launch_bounds__(256) void dense_gemm_lds_kernel(Probes Q, DenseTable T) {
    static_assert(GS % 4 == 0 || (GS % 2 == 0 && PG % 2 == 0), "whole B passes per group, split groups of GS / 2");
    const uint2 wb = gemm_block(Q);
    const uint32_t gw0 = wb.x * 4;
    if (gw0 >= T.waves) return;  // whole workgroups (T.waves is a multiple of 4)
    const uint32_t t = threadIdx.x, gw = gw0 + (t >> 6), lane = t & 63, p0 = wb.y * PG;
    const uint32_t np = min((uint32_t)PG, Q.n - p0);
    uint32_t jj = 0;
    while (jj + 1 < T.n && gw0 >= T.j[jj + 1].wave_begin) ++jj;  // wave_begin: multiples of 4
    const DenseJob& J = T.j[jj];
    const uint32_t rt = J.rt, ks = J.ks, splitk = J.splitk;
    uint32_t pair, slice;
    dense_wave(J, gw - J.wave_begin, pair, slice);  // slice: the same for the workgroup's waves
    const uint32_t rt0 = 2 * pair;
    const bool live = rt0 < rt;  // padding waves load and wait with the others
    const bool has1 = rt0 + 1 < rt;
    const uint32_t kb = (uint32_t)((uint64_t)slice * ks / splitk), ke = (uint32_t)((uint64_t)(slice + 1) * ks / splitk);
    const uint32_t i = lane & 15, g = lane >> 4;
    using GH = const __attribute__((address_space(1))) half8*;
    const GH w0 = (GH)(J.W + (size_t)(live ? rt0 : 0) * ks * 64 + lane);
    const GH w1 = has1 ? w0 + (size_t)ks * 64 : w0;
    __shared__ half8 sB[2][GS * PG * 128];  // a group's B blocks: [step][probe][unit (hi, then lo split)]
    f32x4 acc[PG][4];
#pragma unroll
    for (int q = 0; q < PG; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[q][k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    auto run = [&](auto split_tag) {
        constexpr bool SPLIT = decltype(split_tag)::value;
        constexpr uint32_t G = SPLIT ? GS / 2 : GS;  // steps per group
        constexpr uint32_t kBw = SPLIT ? 256 : 128;  // half8 per probe and step
        constexpr int kWs = SPLIT ? 4 : 2;            // weight fragments per step
        constexpr int kPer = G * PG * kBw / 256;
        static_assert(G * PG * kBw % 256 == 0, "whole 256-lane passes");
        const size_t r_lo = J.r_lo, w_lo = J.w_lo;
        const uint32_t n = ke - kb, groups = (n + G - 1) / G;
        if (groups == 0) return;  // uniform over the workgroup
        auto step_of = [&](uint32_t s) { return kb + (s < n ? s : n - 1); };  // steps past the last repeat it
        auto b_load = [&](uint32_t s0, half8 (&v)[kPer]) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const uint32_t e = t + 256u * k, st = e / (PG * kBw), rem = e % (PG * kBw), q = rem / kBw,
                               o = rem % kBw;
                const GH rq = (GH)scr(Q, J.R, p0 + (q < np ? q : 0));
                v[k] = rq[(size_t)step_of(s0 + st) * 128 + o % 128 + (o >= 128 ? r_lo : 0)];
            }
        };
        auto b_store = [&](uint32_t buf, const half8 (&v)[kPer]) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) sB[buf][t + 256u * k] = v[k];
        };
        auto w_load = [&](uint32_t s0, half8 (&x)[G][kWs]) {
#pragma unroll
            for (uint32_t k = 0; k < G; ++k) {
                const size_t o = (size_t)step_of(s0 + k) * 64;
#if HVX_DENSE_NT
                x[k][0] = __builtin_nontemporal_load(&w0[o]);
                x[k][1] = __builtin_nontemporal_load(&w1[o]);
                if constexpr (SPLIT) {
                    x[k][2] = __builtin_nontemporal_load(&w0[o + w_lo]);
                    x[k][3] = __builtin_nontemporal_load(&w1[o + w_lo]);
                }
#else
                x[k][0] = w0[o];
                x[k][1] = w1[o];
                if constexpr (SPLIT) {
                    x[k][2] = w0[o + w_lo];
                    x[k][3] = w1[o + w_lo];
                }
#endif
            }
        };
        half8 nb[kPer], xw[G][kWs], xn[G][kWs];
        b_load(0, nb);
        w_load(0, xw);
        b_store(0, nb);
        __syncthreads();
        for (uint32_t gi = 0; gi < groups; ++gi) {
            const uint32_t s0 = gi * G, sn = gi + 1 < groups ? s0 + G : s0;
            w_load(sn, xn);
            b_load(sn, nb);
            __asm__ volatile("" ::: "memory");
            const half8* sb = sB[gi & 1];
#pragma unroll
            for (uint32_t k = 0; k < G; ++k) {
                if (live && s0 + k < n) {
#pragma unroll
                    for (int q = 0; q < PG; ++q) {
                        const half8* bk = sb + (k * PG + q) * kBw + g * 32 + i;
#pragma unroll
                        for (int h = 0; h < 2; ++h)
#pragma unroll
                            for (int nbk = 0; nbk < 2; ++nbk) {
                                f32x4& a = acc[q][2 * h + nbk];
                                if constexpr (SPLIT) {
                                    const bf16x8 x = __builtin_bit_cast(bf16x8, xw[k][h]);
                                    const bf16x8 y = __builtin_bit_cast(bf16x8, xw[k][2 + h]);
                                    const bf16x8 b = __builtin_bit_cast(bf16x8, bk[16 * nbk]);
                                    const bf16x8 c = __builtin_bit_cast(bf16x8, bk[128 + 16 * nbk]);
                                    a = hvx::mfma16x16x32(x, b, a);
                                    a = hvx::mfma16x16x32(x, c, a);
                                    a = hvx::mfma16x16x32(y, b, a);
                                } else {
                                    a = hvx::mfma16x16x32(xw[k][h], bk[16 * nbk], a);
                                }
                            }
                    }
                }
            }
            b_store((gi & 1) ^ 1, nb);
#pragma unroll
            for (uint32_t k = 0; k < G; ++k)
#pragma unroll
                for (int u = 0; u < kWs; ++u) xw[k][u] = xn[k][u];
            __syncthreads();
        }
    };
    if (J.w_lo)  // split BF16 (uniform over the workgroup: one job per workgroup)
        run(std::true_type{});
    else
        run(std::false_type{});
    if (!live) return;
#pragma unroll
    for (int q = 0; q < PG; ++q) {
        if ((uint32_t)q >= np) break;
        float* p = scr(Q, J.part, p0 + q) + (size_t)slice * rt * 16 * 32;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t row0 = rt0 * 16 + 4 * g + r;
            p[(size_t)row0 * 32 + i] = acc[q][0][r];
            p[(size_t)row0 * 32 + 16 + i] = acc[q][1][r];
            if (has1) {
                p[(size_t)(row0 + 16) * 32 + i] = acc[q][2][r];
                p[(size_t)(row0 + 16) * 32 + 16 + i] = acc[q][3][r];
            }
        }
    }
}

// out[f][row] = (sum over slices) * rowscale[row] * 2^e  (alpha 1 in the
// chain, 0 in a far map)
__global__ __launch_bounds__(256) void dense_epilogue_kernel(Probes Q, DenseTable T, const float* sc) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x, pr = blockIdx.y;
    if (id >= T.epi) return;
    sc = scr(Q, sc, pr);
    uint32_t jj = 0;
    while (jj + 1 < T.n && id >= T.j[jj + 1].epi_begin) ++jj;
    const DenseJob& J = T.j[jj];
    const uint32_t local = id - J.epi_begin, rows = J.rows, f = local / rows, row = local - f * rows;
    const uint32_t rows_pad = J.rt * 16;
    float v[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t s = 0; s < J.splitk; ++s) {
        const float* p = scr(Q, J.part, pr) + ((size_t)s * rows_pad + row) * 32 + f * 4;
        v[0] += p[0];
        v[1] += p[1];
        v[2] += p[2];
    }
    const float k = J.rowscale[row] * sc[13];
    float4* out = J.out ? scr(Q, J.out, pr) : probe_out(Q, pr) + J.out_off;
    out[(size_t)f * rows + row] = make_float4(v[0] * k, v[1] * k, v[2] * k, J.alpha);
}

struct NearJob {
    const half8* W;
    const NearTile* tiles;
    const half8* P;
    uint32_t Pw8, B, te, So, Sf;
    const float* rowscale;
    const float* invZ;
    const float4* far;
    size_t out_off;  // float4 offset of the mip in the output chain
    uint32_t tile_begin, texel_begin;
    uint64_t w_lo, p_lo;  // split BF16: offsets (16-byte units) of the lo parts of W and P
};
struct NearTable {
    NearJob j[kMaxJobs];
    uint32_t n, tiles, texels;
};

// the tables travel as kernel arguments (at most 4 KB)
static_assert(sizeof(StageTable) < 4000 && sizeof(DenseTable) < 4000 && sizeof(NearTable) < 4000,
              "launch tables must fit the kernel argument space");

// Near field: one workgroup per 16 x 16 output tile, eight waves of two
// output rows each; B rows come from the face planes' window. SPLIT:
// weights and radiance are BF16 hi + lo pairs and each product is
// hi*hi + hi*lo + lo*hi (the lo*lo term is below the FP32 accumulation's
// resolution), three MFMAs per FP16 one.
constexpr uint32_t kNearWaves = 8;  // waves per 16x16 output tile (two row tiles each)
constexpr uint32_t kNearParts = 2;  // workgroups per tile (fixed by the 256-lane LDS staging below)
static_assert(kNearWaves % kNearParts == 0, "a tile splits into whole waves");
static_assert(64 * kNearWaves / kNearParts == 256, "the near GEMM's LDS staging assumes 256-lane workgroups");
#ifndef HVX_NEAR_ROWS
#define HVX_NEAR_ROWS 1
#endif
// The single-probe near GEMM: row tiles per wave (a tile's 16 row tiles take
// 16 / kNearRows waves, in 256-lane workgroups). One row tile per wave
// doubles the waves that stream the weights but loads each B fragment twice
// as often (from L2); four row tiles per wave halve the B loads and the
// waves. Measured on the 512^2 GGX chain, whole FP16 call: one 0.595 ms, two
// 0.625 ms (the smaller work units balance mip 2, whose corner tiles have
// windows 3.4x the mean). Each output's accumulation order is the same for all.
constexpr uint32_t kNearRows = HVX_NEAR_ROWS;
constexpr uint32_t kNearSingleParts = 16 / kNearRows / 4;
static_assert(kNearRows == 1 || kNearRows == 2 || kNearRows == 4, "1, 2 or 4 row tiles per wave");
// XCD-aware order of the single-probe near GEMM: workgroups are dealt to the
// 8 XCDs round-robin, so consecutive workgroups (the parts of one tile)
// would each fetch the tile's radiance window into a different L2. Instead
// hardware workgroup b (XCD b % 8, its k = b / 8-th workgroup there) takes
// part k % P of tile ((k / (P G)) * 8 + b % 8) * G + (k % (P G)) / P: all
// parts of a tile, and runs of G neighbouring tiles, share one XCD, and the
// tiles still start in the order of the table (largest windows first).
// Measured on the 512^2 GGX chain (whole call, FP16 / split BF16): hardware
// order 0.597 / 0.998 ms; G = 1: 0.585 / 0.974; 4: 0.581 / 0.932; 8: 0.590 /
// 0.965; 16: 0.625 / 1.151 (too few XCD turns to balance the mips).
#ifndef HVX_NEAR_XCD_GROUP
#define HVX_NEAR_XCD_GROUP 4
#endif
constexpr uint32_t kNearXcdGroup = HVX_NEAR_XCD_GROUP;  // 0: hardware order
constexpr uint32_t kXcdCount = 8;
// workgroups of the single-probe near GEMM over `tiles` tiles (padding
// workgroups of the XCD-aware order return at once)
__host__ __device__ constexpr uint32_t near_single_wgs(uint32_t tiles) {
    return kNearXcdGroup ? (tiles + kXcdCount * kNearXcdGroup - 1) / (kXcdCount * kNearXcdGroup) * kXcdCount *
                               kNearXcdGroup * kNearSingleParts
                         : tiles * kNearSingleParts;
}
// logical workgroup (tile * P + part) of hardware workgroup b
__device__ __forceinline__ uint32_t near_single_block(uint32_t b) {
    if constexpr (kNearXcdGroup == 0) return b;
    constexpr uint32_t P = kNearSingleParts, PG = P * kNearXcdGroup;
    const uint32_t x = b % kXcdCount, k = b / kXcdCount;
    const uint32_t tile = ((k / PG) * kXcdCount + x) * kNearXcdGroup + (k % PG) / P;
    return tile * P + k % P;
}

template <bool SPLIT, int PG>
__global__ __launch_bounds__(512) void near_gemm_kernel(Probes Q, NearTable T, const float* sc) {
    const uint2 wb = gemm_block(Q);
    const uint32_t p0 = wb.y * PG, np = min((uint32_t)PG, Q.n - p0);
    uint32_t jj = 0;
    // logical workgroup lb owns part lb % kNearSingleParts (four waves' row
    // tiles) of tile lb / kNearSingleParts: finer work units than whole tiles
    // balance the unequal mips across the CUs
    const uint32_t lb = near_single_block(wb.x);
    const uint32_t gt = lb / kNearSingleParts;
    if (gt >= T.tiles) return;  // padding of the XCD-aware order (uniform; no barriers below)
    while (jj + 1 < T.n && gt >= T.j[jj + 1].tile_begin) ++jj;
    const NearJob& J = T.j[jj];
    const uint32_t tile = gt - J.tile_begin, lane = threadIdx.x & 63;
    const uint32_t wave = (threadIdx.x >> 6) + (lb % kNearSingleParts) * 4;
    const NearTile Tl = J.tiles[tile];
    // PG probes per wave: each weight fragment is loaded once for the group;
    // each probe's accumulation is the same sequence as with PG = 1
    const half8* Pp[PG];
#pragma unroll
    for (int q = 0; q < PG; ++q) Pp[q] = scr(Q, J.P, p0 + ((uint32_t)q < np ? q : 0));
    const uint32_t wc = Tl.w / 32, KS = Tl.h * wc;
    const uint32_t i = lane & 15, g = lane >> 4, rt0 = kNearRows * wave;
    // weights of a tile stored K-step major: step s of all 16 row tiles is one
    // contiguous 16 KB block
    constexpr size_t kStep = 16 * 64, kRow = 64;
    const half8* w0 = J.W + Tl.woff + (size_t)rt0 * kRow + lane;
    const uint32_t xb = (uint32_t)(Tl.x0 + (int)J.B) / 8 + g;
    f32x4 acc[PG][2 * kNearRows];
#pragma unroll
    for (int q = 0; q < PG; ++q)
#pragma unroll
        for (uint32_t k = 0; k < 2 * kNearRows; ++k) acc[q][k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // operands read through global (not generic) pointers, so each MFMA
    // waits only for its own loads
    using Op = KStep<SPLIT, PG, kNearRows>;
    using V = typename Op::V;
    using GV = const __attribute__((address_space(1))) V*;
    const GV wv = (GV)reinterpret_cast<const V*>(w0);
    GV pv[PG];
#pragma unroll
    for (int q = 0; q < PG; ++q) pv[q] = (GV)reinterpret_cast<const V*>(Pp[q]);
    // load cursor: step ls at window row lr, 32-column block lcc. A tile's
    // window spans at least 3 rows and 32 columns (the host pads the
    // footprint by a texel on each side), so KS >= 3.
    uint32_t ls = 0, lr = 0, lcc = 0;
    kstep_pipeline<near_unroll<SPLIT>(), SPLIT, PG, kNearRows>(acc, KS, [&](Op& o) {
        const size_t pidx = (((size_t)(Tl.y0 + (int)J.B + (int)lr) * J.Pw8 + xb + lcc * 4) * 32) + i;
        const size_t o0 = (size_t)ls * kStep;
#pragma unroll
        for (uint32_t h = 0; h < kNearRows; ++h) {
            o.x[h] = wv[o0 + h * kRow];
            if constexpr (SPLIT) o.x[kNearRows + h] = wv[o0 + h * kRow + J.w_lo];
        }
#pragma unroll
        for (int q = 0; q < PG; ++q) {
            o.b[q][0] = pv[q][pidx];
            o.b[q][1] = pv[q][pidx + 16];
            if constexpr (SPLIT) {
                o.b[q][2] = pv[q][pidx + J.p_lo];
                o.b[q][3] = pv[q][pidx + 16 + J.p_lo];
            }
        }
        const uint32_t step = ls + 1 < KS ? 1u : 0u;  // selects, not branches, between the loads
        ls += step;
        lcc += step;
        const bool wrap = lcc == wc;
        lr += wrap ? 1u : 0u;
        lcc = wrap ? 0u : lcc;
    });
    const uint32_t So = J.So, X0 = (tile % J.te) * 16, Y0 = (tile / J.te) * 16;
    const size_t So2 = (size_t)So * So;
#pragma unroll
    for (int q = 0; q < PG; ++q) {
        if ((uint32_t)q >= np) break;
        float* out = reinterpret_cast<float*>(probe_out(Q, p0 + q) + J.out_off);
        const float* sq = scr(Q, sc, p0 + q);
#pragma unroll
        for (uint32_t h = 0; h < kNearRows; ++h) {
            const uint32_t Y = Y0 + rt0 + h;
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const f32x4 a = acc[q][2 * h + nb];
                const uint32_t col = nb * 16 + i, f = col >> 2, c = col & 3;
                if (col >= 24) continue;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t X = X0 + 4 * g + e, o = Y * So + X;
                    out[((size_t)f * So2 + o) * 4 + c] = c < 3 ? a[e] * J.rowscale[o] * sq[6 + f] : 1.0f;
                }
            }
        }
    }
}

// The single-probe near GEMM with the radiance window staged in LDS: a
// workgroup's four waves (four row tiles of one tile) share each K step's B
// block (128 half8, twice that split), so the workgroup loads it once into
// LDS, GS steps at a time, double buffered, with one barrier per GS steps;
// each wave then issues one global load per step for its weights (two
// split) instead of also loading the block itself. Weights are loaded a
// group ahead. Each output's accumulation is the same sequence of MFMAs as
// in near_gemm_kernel, so the outputs are identical. Measured on the 512^2
// GGX chain (whole call): FP16 with near_gemm_kernel 0.54 ms; GS = 4: 0.555,
// 8: 0.496, 12: 0.476, 16: 0.464 (2 waves per SIMD), 20: 0.475, 24 spills;
// split BF16 with near_gemm_kernel 0.915 ms; GS = 4: 0.837, 6: 0.794, 8:
// 0.762 (2 waves per SIMD).
#ifndef HVX_NEAR_LDS_GROUP
#define HVX_NEAR_LDS_GROUP 16
#endif
// weights are read once per call: non-temporal loads (whole FP16 / split
// call on the 512^2 GGX chain: 0.460 / 0.754 ms -> 0.449 / 0.723)
#ifndef HVX_NEAR_NT
#define HVX_NEAR_NT 1
#endif
#ifndef HVX_NEAR_LDS_GROUP_SPLIT
#define HVX_NEAR_LDS_GROUP_SPLIT 8
#endif
constexpr int kNearLdsGroup = HVX_NEAR_LDS_GROUP;             // 0: near_gemm_kernel
constexpr int kNearLdsGroupSplit = HVX_NEAR_LDS_GROUP_SPLIT;  // 0: near_gemm_kernel
template <bool SPLIT, int GS>
__global__ __launch_bounds__(256) void near_gemm_lds_kernel(Probes Q, NearTable T, const float* sc) {
    static_assert(kNearRows == 1 && GS % 2 == 0, "one row tile per wave, whole B passes per group");
    const uint32_t lb = near_single_block(blockIdx.x);
    const uint32_t gt = lb / kNearSingleParts;
    if (gt >= T.tiles) return;  // padding of the XCD-aware order (uniform, before any barrier)
    uint32_t jj = 0;
    while (jj + 1 < T.n && gt >= T.j[jj + 1].tile_begin) ++jj;
    const NearJob& J = T.j[jj];
    const uint32_t tile = gt - J.tile_begin, t = threadIdx.x, lane = t & 63;
    const uint32_t wave = (t >> 6) + (lb % kNearSingleParts) * 4;
    const NearTile Tl = J.tiles[tile];
    const uint32_t wc = Tl.w / 32, KS = Tl.h * wc;
    const uint32_t i = lane & 15, g = lane >> 4, rt0 = wave;
    constexpr size_t kStep = 16 * 64;
    constexpr uint32_t kBw = SPLIT ? 256 : 128;  // half8 of a step's B block (hi, then lo)
    constexpr int kWs = SPLIT ? 2 : 1;            // weight fragments per step (hi, lo)
    using GH = const __attribute__((address_space(1))) half8*;
    const GH wv = (GH)(J.W + Tl.woff + (size_t)rt0 * 64 + lane);
    const GH pv = (GH)scr(Q, J.P, 0);
    const uint32_t xb0 = (uint32_t)(Tl.x0 + (int)J.B) / 8;
    const size_t p_lo = J.p_lo, w_lo = J.w_lo;
    // B block of step s: 128 consecutive half8 at ((row * Pw8 + xb0 + col * 4) * 32)
    // (its lo part p_lo further), row = s / wc, col = s % wc; steps past the
    // last repeat it
    auto b_base = [&](uint32_t s) {
        s = s < KS ? s : KS - 1;
        const uint32_t r = s / wc, c = s - r * wc;
        return ((size_t)(Tl.y0 + (int)J.B + (int)r) * J.Pw8 + xb0 + c * 4) * 32;
    };
    constexpr int kPer = GS * kBw / 256;  // half8 of a group's B per thread
    __shared__ half8 sB[2][GS][kBw];
    auto b_load = [&](uint32_t s0, half8 (&v)[kPer]) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t e = t + 256u * k, o = e % kBw;
            v[k] = pv[b_base(s0 + e / kBw) + o % 128 + (o >= 128 ? p_lo : 0)];
        }
    };
    auto b_store = [&](uint32_t buf, const half8 (&v)[kPer]) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t e = t + 256u * k;
            sB[buf][e / kBw][e % kBw] = v[k];
        }
    };
    auto w_load = [&](uint32_t s0, half8 (&x)[GS][kWs]) {
#pragma unroll
        for (int k = 0; k < GS; ++k) {
            const size_t o = (size_t)(s0 + k < KS ? s0 + k : KS - 1) * kStep;
#if HVX_NEAR_NT
            x[k][0] = __builtin_nontemporal_load(&wv[o]);
            if constexpr (SPLIT) x[k][1] = __builtin_nontemporal_load(&wv[o + w_lo]);
#else
            x[k][0] = wv[o];
            if constexpr (SPLIT) x[k][1] = wv[o + w_lo];
#endif
        }
    };
    f32x4 acc[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
    half8 nb[kPer], xw[GS][kWs], xn[GS][kWs];
    b_load(0, nb);
    w_load(0, xw);
    b_store(0, nb);
    __syncthreads();
    const uint32_t groups = (KS + GS - 1) / GS;
    for (uint32_t gi = 0; gi < groups; ++gi) {
        const uint32_t s0 = gi * GS;
        // the next group's weights, then its B block, are in flight during
        // this group's MFMAs (the last group reloads itself, unused)
        const uint32_t sn = gi + 1 < groups ? s0 + GS : s0;
        w_load(sn, xn);
        b_load(sn, nb);
        __asm__ volatile("" ::: "memory");
        const uint32_t buf = gi & 1;
#pragma unroll
        for (int k = 0; k < GS; ++k) {
            if (s0 + k < KS) {  // uniform over the workgroup
                const half8* sb = sB[buf][k];
                if constexpr (SPLIT) {
                    const bf16x8 x = __builtin_bit_cast(bf16x8, xw[k][0]), y = __builtin_bit_cast(bf16x8, xw[k][1]);
                    const bf16x8 b0 = __builtin_bit_cast(bf16x8, sb[g * 32 + i]);
                    const bf16x8 b1 = __builtin_bit_cast(bf16x8, sb[g * 32 + 16 + i]);
                    const bf16x8 c0 = __builtin_bit_cast(bf16x8, sb[128 + g * 32 + i]);
                    const bf16x8 c1 = __builtin_bit_cast(bf16x8, sb[128 + g * 32 + 16 + i]);
                    acc[0] = hvx::mfma16x16x32(x, b0, acc[0]);
                    acc[0] = hvx::mfma16x16x32(x, c0, acc[0]);
                    acc[0] = hvx::mfma16x16x32(y, b0, acc[0]);
                    acc[1] = hvx::mfma16x16x32(x, b1, acc[1]);
                    acc[1] = hvx::mfma16x16x32(x, c1, acc[1]);
                    acc[1] = hvx::mfma16x16x32(y, b1, acc[1]);
                } else {
                    const half8 b0 = sb[g * 32 + i], b1 = sb[g * 32 + 16 + i];
                    acc[0] = hvx::mfma16x16x32(xw[k][0], b0, acc[0]);
                    acc[1] = hvx::mfma16x16x32(xw[k][0], b1, acc[1]);
                }
            }
        }
        // the other buffer's readers passed the previous barrier
        b_store(buf ^ 1, nb);
#pragma unroll
        for (int k = 0; k < GS; ++k)
#pragma unroll
            for (int u = 0; u < kWs; ++u) xw[k][u] = xn[k][u];
        __syncthreads();
    }
    const uint32_t So = J.So, X0 = (tile % J.te) * 16, Y0 = (tile / J.te) * 16;
    const size_t So2 = (size_t)So * So;
    float* out = reinterpret_cast<float*>(probe_out(Q, 0) + J.out_off);
    const float* sq = scr(Q, sc, 0);
    const uint32_t Y = Y0 + rt0;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        const f32x4 a = acc[n];
        const uint32_t col = n * 16 + i, f = col >> 2, c = col & 3;
        if (col >= 24) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t X = X0 + 4 * g + e, o = Y * So + X;
            out[((size_t)f * So2 + o) * 4 + c] = c < 3 ? a[e] * J.rowscale[o] * sq[6 + f] : 1.0f;
        }
    }
}

template <bool SPLIT, int PG>
__global__ __launch_bounds__(256) void near_gemm_staged_kernel(Probes Q, NearTable T, const float* sc) {
    const uint2 wb = gemm_block(Q);
    const uint32_t p0 = wb.y * PG, np = min((uint32_t)PG, Q.n - p0);
    uint32_t jj = 0;
    // workgroup b owns half b & 1 (eight row tiles) of tile b >> 1: finer
    // work units than whole tiles balance the unequal mips across the CUs
    const uint32_t gt = wb.x / kNearParts;
    while (jj + 1 < T.n && gt >= T.j[jj + 1].tile_begin) ++jj;
    const NearJob& J = T.j[jj];
    const uint32_t tile = gt - J.tile_begin, lane = threadIdx.x & 63;
    const uint32_t wave = (threadIdx.x >> 6) + (wb.x % kNearParts) * (kNearWaves / kNearParts);
    const NearTile Tl = J.tiles[tile];
    // PG probes per wave: each weight fragment is loaded once for the group;
    // each probe's accumulation is the same sequence as with PG = 1
    const half8* Pp[PG];
#pragma unroll
    for (int q = 0; q < PG; ++q) Pp[q] = scr(Q, J.P, p0 + ((uint32_t)q < np ? q : 0));
    const uint32_t wc = Tl.w / 32, KS = Tl.h * wc;
    const uint32_t i = lane & 15, g = lane >> 4, rt0 = 2 * wave;
    // weights of a tile stored K-step major: step s of all 16 row tiles is one
    // contiguous 16 KB block
    constexpr size_t kStep = 16 * 64, kRow = 64;
    const half8* w0 = J.W + Tl.woff + (size_t)rt0 * kRow + lane;
    const uint32_t xb0 = (uint32_t)(Tl.x0 + (int)J.B) / 8;
    // The B operand of a K step (for each probe: 4 x 32 consecutive half8 of
    // the staged plane, plus as many lo parts when split) is the same for
    // every wave of the workgroup: it is staged once per step in LDS, double
    // buffered, the next step's copy loaded while this step's MFMAs run. With
    // PG probes per wave this cuts the workgroup's operand reads PG-fold per
    // step, which pays for the barrier (256-probe batch, contraction mode: 8.8 -> 8.4 ms).
    // Single-probe calls stage several steps per barrier instead
    // (near_gemm_lds_kernel): with a barrier per step one probe lost (512^2 x 10
    // GGX chain: 0.366 -> 0.392 ms).
    constexpr int kB = PG * 128 * (SPLIT ? 2 : 1);  // half8 per step
    constexpr int kPer = (kB + 255) / 256;           // per thread (256 lanes)
    static_assert(kB % 256 == 0 || kB < 256, "the B block splits evenly over the workgroup");
    __shared__ half8 sB[2][kB];
    const uint32_t t = threadIdx.x;
    // the B operand is read through global (not generic) pointers: generic
    // loads count against the LDS counter too, so each step's LDS reads
    // would wait for the next step's B loads
    using GH = const __attribute__((address_space(1))) half8*;
    GH Pg[PG];
#pragma unroll
    for (int q = 0; q < PG; ++q) Pg[q] = (GH)Pp[q];
    // element e = t + 256 k of the step's block: lo part e / (PG * 128),
    // probe (e / 128) % PG, unit e % 128. With whole 256-element passes and
    // an even PG the part is the same for the whole pass, so every lane
    // loads unconditionally (conditional loads make the waits conservative).
    const uint32_t p_lo = J.p_lo;
    auto b_load = [&](uint32_t r, uint32_t cc, half8 (&v)[kPer]) {
        const size_t base = ((size_t)(Tl.y0 + (int)J.B + (int)r) * J.Pw8 + xb0 + cc * 4) * 32;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t e = t + 256u * k, q = (e / 128) % PG, o = e % 128;
            if constexpr (kB % 256 == 0 && (PG * 128) % 256 == 0) {
                const uint32_t lo = (256u * k) / (PG * 128);
                v[k] = Pg[q][base + o + (lo ? p_lo : 0)];
            } else {
                const uint32_t lo = e / (PG * 128);
                if (e < (uint32_t)kB) v[k] = Pg[q][base + o + (lo ? p_lo : 0)];
            }
        }
    };
    half8 nx[kPer];
    b_load(0, 0, nx);
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (t + 256u * k < (uint32_t)kB) sB[0][t + 256u * k] = nx[k];
    __syncthreads();
    f32x4 acc[PG][4];
#pragma unroll
    for (int q = 0; q < PG; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[q][k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t r = 0, cc = 0;
    for (uint32_t s = 0; s < KS; ++s) {
        uint32_t rn = r, cn = cc + 1;
        if (cn == wc) {
            cn = 0;
            ++rn;
        }
        const bool more = s + 1 < KS;
        // this step's weights are loaded before the next step's B operand:
        // the load counter completes in order, so the MFMAs' wait for the
        // weights leaves the B loads in flight
        using V = std::conditional_t<SPLIT, bf16x8, half8>;
        const V* wv = reinterpret_cast<const V*>(w0);
        const size_t o0 = (size_t)s * kStep, o1 = o0 + kRow;
        V xw[SPLIT ? 4 : 2];
        xw[0] = wv[o0];
        xw[1] = wv[o1];
        if constexpr (SPLIT) {
            xw[2] = wv[o0 + J.w_lo];
            xw[3] = wv[o1 + J.w_lo];
        }
        b_load(more ? rn : r, more ? cn : cc, nx);  // the last step reloads itself
        __asm__ volatile("" ::: "memory");          // keeps the B loads ahead of the MFMAs
        const half8* sb = sB[s & 1];
        if constexpr (SPLIT) {
            const bf16x8 x0 = xw[0], x1 = xw[1], y0 = xw[2], y1 = xw[3];
#pragma unroll
            for (int q = 0; q < PG; ++q) {
                const bf16x8* pv = reinterpret_cast<const bf16x8*>(sb + q * 128 + g * 32 + i);
                const bf16x8 b0 = pv[0], b1 = pv[16], c0 = pv[PG * 128], c1 = pv[PG * 128 + 16];
                acc[q][0] = hvx::mfma16x16x32(x0, b0, acc[q][0]);
                acc[q][0] = hvx::mfma16x16x32(x0, c0, acc[q][0]);
                acc[q][0] = hvx::mfma16x16x32(y0, b0, acc[q][0]);
                acc[q][1] = hvx::mfma16x16x32(x0, b1, acc[q][1]);
                acc[q][1] = hvx::mfma16x16x32(x0, c1, acc[q][1]);
                acc[q][1] = hvx::mfma16x16x32(y0, b1, acc[q][1]);
                acc[q][2] = hvx::mfma16x16x32(x1, b0, acc[q][2]);
                acc[q][2] = hvx::mfma16x16x32(x1, c0, acc[q][2]);
                acc[q][2] = hvx::mfma16x16x32(y1, b0, acc[q][2]);
                acc[q][3] = hvx::mfma16x16x32(x1, b1, acc[q][3]);
                acc[q][3] = hvx::mfma16x16x32(x1, c1, acc[q][3]);
                acc[q][3] = hvx::mfma16x16x32(y1, b1, acc[q][3]);
            }
        } else {
            const half8 x0 = xw[0], x1 = xw[1];
#pragma unroll
            for (int q = 0; q < PG; ++q) {
                const half8* pv = sb + q * 128 + g * 32 + i;
                const half8 b0 = pv[0], b1 = pv[16];
                acc[q][0] = hvx::mfma16x16x32(x0, b0, acc[q][0]);
                acc[q][1] = hvx::mfma16x16x32(x0, b1, acc[q][1]);
                acc[q][2] = hvx::mfma16x16x32(x1, b0, acc[q][2]);
                acc[q][3] = hvx::mfma16x16x32(x1, b1, acc[q][3]);
            }
        }
        // stored unconditionally: after the last step the buffer is no longer
        // read (its readers passed the previous barrier), and a store under a
        // branch would draw the B loads into the branch, after the MFMAs
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (t + 256u * k < (uint32_t)kB) sB[(s + 1) & 1][t + 256u * k] = nx[k];
        __syncthreads();
        r = rn;
        cc = cn;
    }
    const uint32_t So = J.So, X0 = (tile % J.te) * 16, Y0 = (tile / J.te) * 16;
    const size_t So2 = (size_t)So * So;
#pragma unroll
    for (int q = 0; q < PG; ++q) {
        if ((uint32_t)q >= np) break;
        float* out = reinterpret_cast<float*>(probe_out(Q, p0 + q) + J.out_off);
        const float* sq = scr(Q, sc, p0 + q);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t Y = Y0 + rt0 + h;
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const f32x4 a = acc[q][2 * h + nb];
                const uint32_t col = nb * 16 + i, f = col >> 2, c = col & 3;
                if (col >= 24) continue;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t X = X0 + 4 * g + e, o = Y * So + X;
                    out[((size_t)f * So2 + o) * 4 + c] = c < 3 ? a[e] * J.rowscale[o] * sq[6 + f] : 1.0f;
                }
            }
        }
    }
}

// out(o) += far field interpolated at o / Z(o)
__global__ __launch_bounds__(256) void far_add_kernel(Probes Q, NearTable T) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x, pr = blockIdx.y;
    if (id >= T.texels) return;
    uint32_t jj = 0;
    while (jj + 1 < T.n && id >= T.j[jj + 1].texel_begin) ++jj;
    const NearJob& J = T.j[jj];
    const uint32_t So = J.So, So2 = So * So, local = id - J.texel_begin;
    const uint32_t f = local / So2, o = local - f * So2, Y = o / So, X = o - Y * So;
    const f3 F = sample_cube(scr(Q, J.far, pr), (int)J.Sf, texel_dir((int)f, (int)X, (int)Y, (int)So));
    const float z = J.invZ[o];
    float4* out = probe_out(Q, pr) + J.out_off;
    float4 v = out[local];
    v.x += F.x * z;
    v.y += F.y * z;
    v.z += F.z * z;
    out[local] = v;
}

// Roughness 0: the bilinear seamless sample of level 0 at the texel direction.
__global__ __launch_bounds__(256) void delta_kernel(Probes Q, uint32_t S, uint32_t So, size_t out_off) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x, pr = blockIdx.y;
    const float4* level0 = probe_chain(Q, pr);
    float4* out = probe_out(Q, pr) + out_off;
    const uint32_t So2 = So * So;
    if (id >= 6 * So2) return;
    const uint32_t f = id / So2, o = id - f * So2, Y = o / So, X = o - Y * So;
    const f3 c = sample_cube(level0, (int)S, texel_dir((int)f, (int)X, (int)Y, (int)So));
    out[id] = make_float4(c.x, c.y, c.z, 1.0f);
}

// --------------------------------------------------- bright-texel split
// (hvx_contract.h): threshold from a histogram of the largest channel's float
// pattern, clamped level 0, the hot texels in texel order, and their exact
// level-0 sum added to every non-delta output texel.
struct HotArgs {
    const float4* level0;  // the caller's level 0
    float4* clamped;       // level 0 of the clamped chain
    uint32_t S, K;
    uint32_t* hist;        // kHotBins counts, zero between calls
    uint32_t* cnt;         // [0] hot texels found, [1] hot texels listed, [2] blocks
    float* T;
    uint32_t* idx;         // kHotMax texel indices; after the sort, the member range of each block
    float4* list;          // per listed texel: (direction, 0), (solid angle * excess, 0)
    float4* groups;        // per 4x4 block: (mean direction, radius), (summed excess, member range)
};

__device__ __forceinline__ float hot_largest(float4 t) { return fmaxf(fmaxf(t.x, t.y), t.z); }

__global__ __launch_bounds__(256) void hot_hist_kernel(HotArgs H) {
    __shared__ uint32_t sh[hvx::kHotBins];
    for (uint32_t b = threadIdx.x; b < hvx::kHotBins; b += 256) sh[b] = 0u;
    __syncthreads();
    const size_t n = 6ull * H.S * H.S, stride = (size_t)gridDim.x * 256;
    const size_t i0 = (size_t)blockIdx.x * 256 + threadIdx.x;
    // four texels in flight per thread; the bins are then counted one by one
    for (size_t i = i0; i < n; i += 4 * stride) {
        float m4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const size_t k = i + q * stride;
            m4[q] = k < n ? hot_largest(H.level0[k]) : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float m = m4[q];
            const uint32_t bin = m > 0.0f ? hvx::contract_hot_bin(__float_as_uint(m)) : 0xffffffffu;
            // neighbouring texels mostly share a bin: one LDS atomic per wave then
            const uint32_t b0 = __builtin_amdgcn_readfirstlane(bin);
            if (__all(bin == b0)) {
                const unsigned long long act = __ballot(1);
                if (b0 != 0xffffffffu && (threadIdx.x & 63) == (uint32_t)__ffsll(act) - 1)
                    atomicAdd(&sh[b0], (uint32_t)__popcll(act));
            } else if (bin != 0xffffffffu) {
                atomicAdd(&sh[bin], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < hvx::kHotBins; b += 256)
        if (sh[b]) atomicAdd(&H.hist[b], sh[b]);
}

// One workgroup of 1024: c(b) = texels in bins >= b (a scan over the bins
// from the top, two per thread), T from the number of bins with c(b) > K;
// clears the histogram and the counters for the next call.
__global__ __launch_bounds__(1024) void hot_select_kernel(HotArgs H) {
    __shared__ uint32_t s_wave[16];
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const uint32_t b_hi = hvx::kHotBins - 1 - 2 * t, b_lo = b_hi - 1;  // thread t: bins 2047-2t, 2046-2t
    const uint32_t h_hi = H.hist[b_hi], h_lo = H.hist[b_lo];
    H.hist[b_hi] = 0u;
    H.hist[b_lo] = 0u;
    uint32_t inc = h_hi + h_lo;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(inc, d, 64);
        if ((int)lane >= d) inc += u;
    }
    if (lane == 63) s_wave[wv] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wv; ++w) before += s_wave[w];
    const uint32_t c_lo = before + inc, c_hi = c_lo - h_lo;  // c(b_lo), c(b_hi)
    const int over = __syncthreads_count(c_hi > H.K) + __syncthreads_count(c_lo > H.K);
    if (t == 0) {
        H.T[0] = __uint_as_float(hvx::contract_hot_threshold_bits((uint32_t)over));
        H.cnt[0] = 0u;
        H.cnt[1] = 0u;
        H.cnt[2] = 0u;
    }
}

__global__ __launch_bounds__(256) void hot_compact_kernel(HotArgs H) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= 6ull * H.S * H.S) return;
    const float T = H.T[0];
    float4 v = H.level0[i];
    const float m = hot_largest(v);
    if (m > 0.0f && m >= T) {
        const uint32_t k = atomicAdd(&H.cnt[0], 1u);
        if (k < H.K) H.idx[k] = (uint32_t)i;
        v.x = v.x > T ? T : v.x;
        v.y = v.y > T ? T : v.y;
        v.z = v.z > T ? T : v.z;
    }
    H.clamped[i] = v;
}

// One workgroup of 1024: sorts the found texels by (face, 4x4 block, texel
// in block) so that the sum order is fixed and each 4x4 block's texels are
// contiguous, then writes each texel's direction and solid angle times
// excess (FP64 geometry, as the oracle's level-0 sum) and the member range
// of every block.
__device__ __forceinline__ uint32_t hot_key(uint32_t i, uint32_t S) {
    const uint32_t S2 = S * S, g = i / S2, r = i - g * S2, y = r / S, x = r - y * S, Q = S >> 2;
    return (((g * Q + (y >> 2)) * Q + (x >> 2)) << 4) | ((y & 3u) << 2) | (x & 3u);
}
__device__ __forceinline__ uint32_t hot_index(uint32_t key, uint32_t S) {
    const uint32_t Q = S >> 2, blk = key >> 4, bx = blk % Q, by = (blk / Q) % Q, g = blk / (Q * Q);
    const uint32_t y = (by << 2) | ((key >> 2) & 3u), x = (bx << 2) | (key & 3u);
    return (g * S + y) * S + x;
}

__global__ __launch_bounds__(1024) void hot_list_kernel(HotArgs H) {
    __shared__ uint32_t s[hvx::kHotMax];
    __shared__ uint32_t s_wave[16];
    const uint32_t n = min(H.cnt[0], H.K), t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (uint32_t j = t; j < hvx::kHotMax; j += 1024) s[j] = j < n ? hot_key(H.idx[j], H.S) : 0xffffffffu;  // S >= 4
    __syncthreads();
    for (uint32_t size = 2; size <= hvx::kHotMax; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t j = t; j < hvx::kHotMax / 2; j += 1024) {
                const uint32_t lo = 2 * j - (j & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint32_t a = s[lo], b = s[hi];
                if ((a > b) == up) {
                    s[lo] = b;
                    s[hi] = a;
                }
            }
            __syncthreads();
        }
    const float T = H.T[0];
    const uint32_t S2 = H.S * H.S;
    for (uint32_t k = t; k < n; k += 1024) {
        const uint32_t i = hot_index(s[k], H.S), g = i / S2, r = i - g * S2, y = r / H.S, x = r - y * H.S;
        const double u = hvx::contract_coord((int)x, H.S), v = hvx::contract_coord((int)y, H.S);
        double p[3];
        hvx::contract_face_point((int)g, u, v, p);
        const double rl = 1.0 / sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]), dw = hvx::contract_dw(u, v, H.S);
        const float4 c = H.level0[i];
        const float ex = c.x > T ? c.x - T : 0.0f, ey = c.y > T ? c.y - T : 0.0f, ez = c.z > T ? c.z - T : 0.0f;
        H.list[2 * k] = make_float4((float)(p[0] * rl), (float)(p[1] * rl), (float)(p[2] * rl), 0.0f);
        H.list[2 * k + 1] = make_float4((float)(dw * ex), (float)(dw * ey), (float)(dw * ez), 0.0f);
    }
    // block starts: thread t owns entries 4t .. 4t+3; exclusive scan of the counts
    uint32_t starts = 0;
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t k = 4 * t + q;
        starts += (k < n && (k == 0 || (s[k] >> 4) != (s[k - 1] >> 4))) ? 1u : 0u;
    }
    uint32_t inc = starts;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(inc, d, 64);
        if ((int)lane >= d) inc += u;
    }
    if (lane == 63) s_wave[wv] = inc;
    __syncthreads();
    uint32_t gi = inc - starts;
    for (uint32_t w = 0; w < wv; ++w) gi += s_wave[w];
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t k = 4 * t + q;
        if (!(k < n && (k == 0 || (s[k] >> 4) != (s[k - 1] >> 4)))) continue;
        uint32_t e = k + 1;
        while (e < n && (s[e] >> 4) == (s[k] >> 4)) ++e;
        H.idx[gi++] = k | ((e - k) << 16);  // the found indices are consumed: idx now holds block ranges
    }
    if (t == 1023) {
        H.cnt[1] = n;
        H.cnt[2] = gi;
    }
}

// One thread per block of hot texels: excess-weighted mean direction, the
// angular radius of the members around it, the summed excess.
__global__ __launch_bounds__(256) void hot_group_kernel(HotArgs H) {
    const uint32_t gi = blockIdx.x * 256 + threadIdx.x;
    if (gi >= H.cnt[2]) return;
    const uint32_t pk = H.idx[gi], k = pk & 0xffffu, e = k + (pk >> 16);
    double E[3] = {0.0, 0.0, 0.0}, m[3] = {0.0, 0.0, 0.0};
    for (uint32_t j = k; j < e; ++j) {
        const float4 L = H.list[2 * j], X = H.list[2 * j + 1];
        const double wgt = (double)X.x + (double)X.y + (double)X.z;
        E[0] += X.x;
        E[1] += X.y;
        E[2] += X.z;
        m[0] += wgt * L.x;
        m[1] += wgt * L.y;
        m[2] += wgt * L.z;
    }
    double ml = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    if (!(ml > 0.0)) {  // no excess in the block: any member direction
        const float4 L = H.list[2 * k];
        m[0] = L.x;
        m[1] = L.y;
        m[2] = L.z;
        ml = 1.0;
    }
    double rad = 0.0;
    for (uint32_t j = k; j < e; ++j) {
        const float4 L = H.list[2 * j];
        const double c = (m[0] * L.x + m[1] * L.y + m[2] * L.z) / ml;
        rad = fmax(rad, acos(fmin(1.0, fmax(-1.0, c))));
    }
    H.groups[2 * gi] = make_float4((float)(m[0] / ml), (float)(m[1] / ml), (float)(m[2] / ml),
                                   (float)(rad * (1.0 + 1e-6) + 1e-7));
    H.groups[2 * gi + 1] = make_float4((float)E[0], (float)E[1], (float)E[2], __uint_as_float(pk));
}

struct HotMip {
    uint32_t So, first;  // output face; first output of the mip in the launch (a multiple of 64: one mip per workgroup)
    float alpha2, h1;    // GGX a^2 and (1 - a^2) / 2
    float scale;         // a^2 / pi / (weight integral); diffuse: 1 / pi
    uint32_t _pad;
    size_t out_off;      // the mip in the output chain, float4 units
};
struct HotAdd {
    HotMip m[hvx::kContractMaxMips];
    uint32_t nm, blocks;
    const float4* list;
    const float4* groups;
    const uint32_t* cnt;
    float4* out;
    float theta_near;  // blocks farther than this (radians) from every output direction may enter as one entry
};

// w'(n . l) E: w' = c / d^2 (GGX, d = a^2 + (1 - c)(1 - a^2)/2 = N.H^2 (a^2 - 1)
// + 1) or c (cosine), 0 for c <= 0; 1 - c is formed as |n - l|^2 / 2, which
// keeps its relative accuracy for nearby directions.
template <bool DIFFUSE>
__device__ __forceinline__ void hot_term(const f3& n, const float4 L, const float4 E, float a2, float h1, float& ax,
                                         float& ay, float& az) {
    const float dx = n.x - L.x, dy = n.y - L.y, dz = n.z - L.z;
    const float q = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));  // 2 (1 - c)
    const float c = __builtin_fmaf(q, -0.5f, 1.0f);
    float w;
    if constexpr (DIFFUSE) {
        w = fmaxf(c, 0.0f);
    } else {
        const float d = __builtin_fmaf(q, h1, a2);  // h1 holds (1 - a^2) / 4 here
        w = fmaxf(c, 0.0f) * __builtin_amdgcn_rcpf(d * d);
    }
    ax = __builtin_fmaf(w, E.x, ax);
    ay = __builtin_fmaf(w, E.y, ay);
    az = __builtin_fmaf(w, E.z, az);
}

// out(o) += scale * sum over the hot texels of w'(n_o . l_k) E_k. A workgroup
// holds 64 output texels of one mip. A 4x4 block of hot texels whose cone
// lies more than theta_near outside the workgroup's cone of output
// directions, and wholly in front of every one of them (c > 0: clear of the
// kink of w at c = 0), enters as one entry at its excess-weighted mean
// direction (the terms linear in the offsets cancel; the rest is of relative
// order 10 (block extent / distance)^2); a block wholly behind every output
// direction adds exactly 0 and is skipped; every other block is summed texel
// by texel.
template <bool DIFFUSE>
__global__ __launch_bounds__(256) void hot_add_kernel(HotAdd A) {
    constexpr uint32_t kG = 32, kE = kG * 16;  // blocks per round; entries they can expand to
    __shared__ float4 se[2 * kE];
    __shared__ uint32_t s_ne;
    __shared__ float s_acc[3][3][64];
    // 64 output texels per workgroup (an 8 x 8 tile of one face when the face
    // has at least 8 texels), one per lane; the four waves split the entries
    const uint32_t b0 = blockIdx.x * 64, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t mi = 0;
    while (mi + 1 < A.nm && b0 >= A.m[mi + 1].first) ++mi;  // uniform over the workgroup
    const HotMip M = A.m[mi];
    const uint32_t So2 = M.So * M.So, loc0 = b0 + lane - M.first;
    const bool live = loc0 < 6 * So2;
    uint32_t f, X, Y;
    {
        const uint32_t l = live ? loc0 : 0u;
        f = l / So2;
        const uint32_t w = l - f * So2;
        if (M.So >= 8) {  // So^2 is a multiple of 64
            const uint32_t te = M.So >> 3, tile = w >> 6, e = w & 63u;
            X = (tile % te) * 8 + (e & 7u);
            Y = (tile / te) * 8 + (e >> 3);
        } else {
            Y = w / M.So;
            X = w - Y * M.So;
        }
    }
    const uint32_t loc = (f * M.So + Y) * M.So + X;
    const f3 n = texel_dir((int)f, (int)X, (int)Y, (int)M.So);
    // the workgroup's cone (every wave holds the same 64 directions): mean
    // direction of the live texels, smallest cosine to it
    float cx = live ? n.x : 0.0f, cy = live ? n.y : 0.0f, cz = live ? n.z : 0.0f;
    for (int d = 32; d > 0; d >>= 1) {
        cx += __shfl_xor(cx, d, 64);
        cy += __shfl_xor(cy, d, 64);
        cz += __shfl_xor(cz, d, 64);
    }
    const float cl = sqrtf(cx * cx + cy * cy + cz * cz);
    const bool cone = cl > 1e-3f;  // else: directions all round the sphere, no cone
    if (cone) {
        cx /= cl;
        cy /= cl;
        cz /= cl;
    }
    float cmin = live ? n.x * cx + n.y * cy + n.z * cz : 1.0f;
    for (int d = 32; d > 0; d >>= 1) cmin = fminf(cmin, __shfl_xor(cmin, d, 64));
    // radius of the output cone, rounded up (a few 1e-6 rad)
    const float rad_o = acosf(fmaxf(-1.0f, fminf(1.0f, cmin))) * 1.0001f + 1e-5f;
    const float kHalfPi = 1.57079633f;
    const float h1 = DIFFUSE ? 0.0f : 0.5f * M.h1;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    const uint32_t ng = A.cnt[2];
    for (uint32_t base = 0; base < ng; base += kG) {
        const uint32_t cg = min(kG, ng - base);
        __syncthreads();  // the previous round's entries are consumed
        if (wv == 0) {
            // the first wave sorts this round's blocks into skipped (0 entries),
            // aggregated (1) and summed texel by texel (its members), and
            // expands them into the entry list at their prefix offsets
            uint32_t mode = 2u, cnt = 0u, pk = 0u;
            float4 G0 = make_float4(0.f, 0.f, 0.f, 0.f), G1 = G0;
            if (lane < cg) {
                G0 = A.groups[2 * (base + lane)];
                G1 = A.groups[2 * (base + lane) + 1];
                pk = __float_as_uint(G1.w);
                const float ang = acosf(fmaxf(-1.0f, fminf(1.0f, G0.x * cx + G0.y * cy + G0.z * cz)));
                const float lo = ang * 0.9999f - 1e-5f - rad_o - G0.w, hi = ang * 1.0001f + 1e-5f + rad_o + G0.w;
                mode = !cone ? 0u : lo > kHalfPi + 1e-4f ? 2u : (lo > A.theta_near && hi < kHalfPi - 1e-4f) ? 1u : 0u;
                cnt = mode == 2u ? 0u : mode == 1u ? 1u : (pk >> 16);
            }
            uint32_t inc = cnt;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t u = __shfl_up(inc, d, 64);
                if ((int)lane >= d) inc += u;
            }
            const uint32_t off = inc - cnt;
            if (lane == 63) s_ne = inc;
            if (mode == 1u) {
                se[2 * off] = G0;
                se[2 * off + 1] = G1;
            } else if (mode == 0u) {
                const uint32_t k0 = pk & 0xffffu;
                for (uint32_t j = 0; j < cnt; ++j) {
                    se[2 * (off + j)] = A.list[2 * (k0 + j)];
                    se[2 * (off + j) + 1] = A.list[2 * (k0 + j) + 1];
                }
            }
        }
        __syncthreads();
        const uint32_t ne = s_ne;
        for (uint32_t j = wv; j < ne; j += 4) hot_term<DIFFUSE>(n, se[2 * j], se[2 * j + 1], M.alpha2, h1, ax, ay, az);
    }
    // the waves' partial sums, added in wave order
    if (wv > 0) {
        s_acc[0][wv - 1][lane] = ax;
        s_acc[1][wv - 1][lane] = ay;
        s_acc[2][wv - 1][lane] = az;
    }
    __syncthreads();
    if (wv != 0 || !live) return;
    ax = ((ax + s_acc[0][0][lane]) + s_acc[0][1][lane]) + s_acc[0][2][lane];
    ay = ((ay + s_acc[1][0][lane]) + s_acc[1][1][lane]) + s_acc[1][2][lane];
    az = ((az + s_acc[2][0][lane]) + s_acc[2][1][lane]) + s_acc[2][2][lane];
    float4& r = A.out[M.out_off + loc];
    r.x = __builtin_fmaf(ax, M.scale, r.x);
    r.y = __builtin_fmaf(ay, M.scale, r.y);
    r.z = __builtin_fmaf(az, M.scale, r.z);
}

// -------------------------------------------------------------- host side
uint32_t ilog2(uint32_t v) {
    uint32_t l = 0;
    while ((1u << (l + 1)) <= v) ++l;
    return l;
}

// Integer frame of face f: world point of face-plane (u, v) = u E0 + v E1 + E2.
void face_frame(int f, int E[3][3]) {
    double z[3], a[3], b[3];
    hvx::contract_face_point(f, 0.0, 0.0, z);
    hvx::contract_face_point(f, 1.0, 0.0, a);
    hvx::contract_face_point(f, 0.0, 1.0, b);
    for (int j = 0; j < 3; ++j) {
        E[0][j] = (int)(a[j] - z[j]);
        E[1][j] = (int)(b[j] - z[j]);
        E[2][j] = (int)z[j];
    }
}

// World texel index (face-major) of the texel whose direction is output
// face f's frame applied to canonical input k: exact integer arithmetic in
// doubled texel units (a texel centre of face g is a E0 + b E1 + S E2 with
// odd a, b in (-S, S)).
int32_t perm_index(int f, uint32_t k, uint32_t S) {
    const uint32_t S2 = S * S, g = k / S2, rem = k % S2, y = rem / S, x = rem % S;
    const int s = (int)S, a = 2 * (int)x + 1 - s, b = 2 * (int)y + 1 - s;
    int Eg[3][3], Ef[3][3], p[3], q[3];
    face_frame((int)g, Eg);
    face_frame(f, Ef);
    for (int j = 0; j < 3; ++j) p[j] = a * Eg[0][j] + b * Eg[1][j] + s * Eg[2][j];
    for (int j = 0; j < 3; ++j) q[j] = p[0] * Ef[0][j] + p[1] * Ef[1][j] + p[2] * Ef[2][j];
    const int ax = std::abs(q[0]), ay = std::abs(q[1]), az = std::abs(q[2]);
    int nf, na, nb;
    if (ax >= ay && ax >= az) {
        if (q[0] > 0) { nf = 0; na = -q[2]; nb = -q[1]; }
        else { nf = 1; na = q[2]; nb = -q[1]; }
    } else if (ay >= az) {
        if (q[1] > 0) { nf = 2; na = q[0]; nb = q[2]; }
        else { nf = 3; na = q[0]; nb = -q[2]; }
    } else {
        if (q[2] > 0) { nf = 4; na = q[0]; nb = -q[1]; }
        else { nf = 5; na = -q[0]; nb = -q[1]; }
    }
    const int ni = (na + s - 1) / 2, nj = (nb + s - 1) / 2;
    return (int32_t)(((size_t)nf * S + nj) * S + ni);
}

void orthonormal(const double n[3], double e1[3], double e2[3]) {
    const double a[3] = {std::fabs(n[0]) < 0.9 ? 1.0 : 0.0, std::fabs(n[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    e1[0] = a[1] * n[2] - a[2] * n[1];
    e1[1] = a[2] * n[0] - a[0] * n[2];
    e1[2] = a[0] * n[1] - a[1] * n[0];
    const double l = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (int j = 0; j < 3; ++j) e1[j] /= l;
    e2[0] = n[1] * e1[2] - n[2] * e1[1];
    e2[1] = n[2] * e1[0] - n[0] * e1[2];
    e2[2] = n[0] * e1[1] - n[1] * e1[0];
}

// Windows (bounding boxes of the angular caps of radius t1 around every
// output of a tile, on the face plane at the near level), aligned for the
// GEMM's B loads: x0 to 8 texels, width to 32.
std::vector<NearTile> near_tiles(uint32_t So, uint32_t Si, double t1, uint32_t* B_out) {
    const uint32_t te = So / 16;
    std::vector<NearTile> tiles(te * te);
    int B = 0;
    uint64_t woff = 0;
    const double ct = std::cos(t1), st = std::sin(t1);
    for (uint32_t ty = 0; ty < te; ++ty)
        for (uint32_t tx = 0; tx < te; ++tx) {
            double xmin = 1e30, xmax = -1e30, ymin = 1e30, ymax = -1e30;
            for (uint32_t yy = 0; yy < 16; ++yy)
                for (uint32_t xx = 0; xx < 16; ++xx) {
                    const double u = hvx::contract_coord((int)(tx * 16 + xx), So);
                    const double v = hvx::contract_coord((int)(ty * 16 + yy), So);
                    const double r = 1.0 / std::sqrt(1.0 + u * u + v * v);
                    const double n[3] = {u * r, v * r, r};
                    double e1[3], e2[3];
                    orthonormal(n, e1, e2);
                    for (int j = 0; j < 64; ++j) {
                        const double a = 2.0 * hvx::kContractPi * j / 64.0, ca = std::cos(a), sa = std::sin(a);
                        double l[3];
                        for (int q = 0; q < 3; ++q) l[q] = ct * n[q] + st * (ca * e1[q] + sa * e2[q]);
                        const double pu = l[0] / l[2], pv = l[1] / l[2];
                        const double px = ((pu + 1.0) * Si - 1.0) / 2.0, py = ((pv + 1.0) * Si - 1.0) / 2.0;
                        xmin = std::min(xmin, px);
                        xmax = std::max(xmax, px);
                        ymin = std::min(ymin, py);
                        ymax = std::max(ymax, py);
                    }
                }
            const int x_lo = (int)std::floor(xmin) - 1, x_hi = (int)std::ceil(xmax) + 1;
            const int y_lo = (int)std::floor(ymin) - 1, y_hi = (int)std::ceil(ymax) + 1;
            NearTile& T = tiles[ty * te + tx];
            T.x0 = x_lo >= 0 ? (x_lo / 8) * 8 : -(((-x_lo) + 7) / 8) * 8;
            T.w = (uint32_t)(((x_hi + 1 - T.x0) + 31) / 32 * 32);
            T.y0 = y_lo;
            T.h = (uint32_t)(y_hi + 1 - y_lo);
            T.woff = woff;
            woff += (uint64_t)16 * T.h * (T.w / 32) * 64;
            B = std::max({B, -T.x0, T.x0 + (int)T.w - (int)Si, -T.y0, T.y0 + (int)T.h - (int)Si});
        }
    *B_out = (uint32_t)((B + 7) / 8 * 8);
    return tiles;
}

struct DenseMat {
    DenseGeom G;
    int stage;  // index into the staged dense levels
    bool split = false;  // split-BF16 weights (hi, then lo after rt * ks * 64 units)
    uint32_t splitk;
    half8* W = nullptr;
    float* rowscale = nullptr;
};

struct StageDense {
    uint32_t Sin, level;
    bool split = false;  // BF16 hi + lo (lo after K / 8 * 32 units)
    int32_t* perm = nullptr;
    half8* R = nullptr;  // in the per-call scratch, at off
    size_t off = 0;
};

struct StagePad {
    uint32_t Si, level, B, Pw;
    half8* P = nullptr;  // in the per-call scratch, at off
    size_t off = 0;
};

// Reserves `bytes` (256-byte aligned) at the end of a scratch layout.
size_t scratch_take(size_t* total, size_t bytes) {
    const size_t at = (*total + 255) & ~(size_t)255;
    *total = at + bytes;
    return at;
}

struct MipPlan {
    ContractLevel L{};
    int dense = -1;  // FULL: the mip's matrix; SPLIT: the far matrix
    int pad = -1;    // SPLIT: staged padded planes
    NearGeom NG{};
    std::vector<NearTile> tiles;
    NearTile* d_tiles = nullptr;
    half8* Wn = nullptr;
    float* rowscale_n = nullptr;
    float* invZ = nullptr;
    float4* far_map = nullptr;  // in the per-call scratch, at far_off
    size_t far_off = 0;
    float4* zfar_map = nullptr;
    uint64_t weight_bytes = 0;
    uint64_t w_lo = 0;  // split BF16: offset of the lo weights in Wn (16-byte units)
};

}  // namespace

struct hvx_contract_plan {
    hvx_ctx* ctx = nullptr;
    int type = 0, precision = 0;
    uint32_t S = 0, So = 0, mips = 0;
    std::vector<MipPlan> mp;
    std::vector<DenseMat> dm;
    std::vector<StageDense> sd;
    std::vector<StagePad> sp;
    // Per-call scratch, carved from one block of scratch_bytes: far maps,
    // staged levels and planes, face maxima, scales, K-slice partial sums.
    // Multi-probe calls use mcap further copies of the block (mscratch).
    char* scratch = nullptr;
    size_t scratch_bytes = 0;
    char* mscratch = nullptr;
    uint32_t mcap = 0;
    float* fmax = nullptr;  // per-face partial maxima [6][kFaceMaxBlocks]
    float* sc = nullptr;
    float* part = nullptr;   // every dense job's K-slice partial sums
    uint32_t min_level = 0;  // deepest input chain level read
    uint32_t scale_level = ~0u;  // finest staged level: its maxima bound every staged value
    uint32_t delta_mip = ~0u;
    StageTable stage{};
    DenseTable dense{};
    NearTable near{};
    // bright-texel split (hvx_ibl_contract_plan_set_hot): off while hot_k == 0
    uint32_t hot_k = 0;
    float* hot_chain = nullptr;  // the clamped chain, min_level + 1 levels
    uint32_t* hot_hist = nullptr;
    uint32_t* hot_cnt = nullptr;
    float* hot_T = nullptr;
    uint32_t* hot_idx = nullptr;
    float4* hot_list = nullptr;
    float4* hot_groups = nullptr;
    std::vector<void*> allocs;

    template <class T> hipError_t alloc(T** p, size_t count) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) allocs.push_back(q);
        *p = static_cast<T*>(q);
        return e;
    }
};

namespace {

int pfail(hvx_ctx* ctx, int st, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int pfail(hvx_ctx* ctx, int st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    return st;
}

int stage_dense_index(hvx_contract_plan* P, uint32_t Sin, bool split) {
    for (size_t i = 0; i < P->sd.size(); ++i)
        if (P->sd[i].Sin == Sin && P->sd[i].split == split) return (int)i;
    StageDense s;
    s.Sin = Sin;
    s.split = split;
    s.level = ilog2(P->S / Sin);
    P->sd.push_back(s);
    return (int)P->sd.size() - 1;
}

hipError_t build_dense(hvx_contract_plan* P, DenseMat& D, int num_cus, float4* zmap) {
    const DenseGeom& G = D.G;
    hipError_t e;
    double *Z = nullptr, *wm = nullptr;
    if ((e = hipMalloc(&Z, G.rows * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&wm, G.rows * sizeof(double))) != hipSuccess) {
        (void)hipFree(Z);
        return e;
    }
    const uint64_t n = (uint64_t)G.rt * G.ks * 64;
    if ((e = P->alloc(&D.W, D.split ? 2 * n : n)) == hipSuccess && (e = P->alloc(&D.rowscale, G.rows)) == hipSuccess) {
        hipLaunchKernelGGL(dense_row_kernel, dim3(G.rows), dim3(256), 0, nullptr, G, Z, wm);
        if (D.split)
            hipLaunchKernelGGL(dense_fill_kernel<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, G,
                               wm, D.W, n);
        else
            hipLaunchKernelGGL(dense_fill_kernel<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, G,
                               wm, D.W, (uint64_t)0);
        hipLaunchKernelGGL(dense_scale_kernel, dim3((G.rows + 255) / 256), dim3(256), 0, nullptr, G.rows, Z, wm,
                           G.piece == PIECE_FAR ? 1 : 0, D.rowscale, zmap);
        e = hipDeviceSynchronize();
    }
    (void)hipFree(Z);
    (void)hipFree(wm);
    // K slices: enough waves to keep the weight stream in flight
    const uint32_t pairs = (G.rt + 1) / 2;
    const uint32_t want = (uint32_t)std::max(1, num_cus * 8) / std::max(pairs, 1u);
    D.splitk = std::max(1u, std::min(want, std::max(1u, G.ks / 16)));
    return e;
}

hipError_t build_near(hvx_contract_plan* P, MipPlan& M, bool split) {
    hipError_t e;
    const uint32_t So = M.L.So, rows = So * So;
    uint64_t halfs8 = 0;
    for (const NearTile& T : M.tiles) halfs8 += (uint64_t)16 * T.h * (T.w / 32) * 64;
    M.w_lo = split ? halfs8 : 0;
    double *Z = nullptr, *wm = nullptr;
    if ((e = hipMalloc(&Z, rows * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&wm, rows * sizeof(double))) != hipSuccess) {
        (void)hipFree(Z);
        return e;
    }
    if ((e = P->alloc(&M.d_tiles, M.tiles.size())) == hipSuccess &&
        (e = hipMemcpy(M.d_tiles, M.tiles.data(), M.tiles.size() * sizeof(NearTile), hipMemcpyHostToDevice)) ==
            hipSuccess &&
        (e = P->alloc(&M.Wn, split ? 2 * halfs8 : halfs8)) == hipSuccess &&
        (e = P->alloc(&M.rowscale_n, rows)) == hipSuccess && (e = P->alloc(&M.invZ, rows)) == hipSuccess) {
        hipLaunchKernelGGL(near_row_kernel, dim3(rows), dim3(256), 0, nullptr, M.NG, M.d_tiles, Z, wm);
        if (split)
            hipLaunchKernelGGL(near_fill_kernel<true>, dim3((unsigned)M.tiles.size() * 16), dim3(256), 0, nullptr,
                               M.NG, M.d_tiles, wm, M.Wn, halfs8);
        else
            hipLaunchKernelGGL(near_fill_kernel<false>, dim3((unsigned)M.tiles.size() * 16), dim3(256), 0, nullptr,
                               M.NG, M.d_tiles, wm, M.Wn, (uint64_t)0);
        hipLaunchKernelGGL(near_scale_kernel, dim3((rows + 255) / 256), dim3(256), 0, nullptr, M.NG, M.L.Sf, Z, wm,
                           M.zfar_map, M.rowscale_n, M.invZ);
        e = hipDeviceSynchronize();
    }
    M.weight_bytes += halfs8 * 16 * (split ? 2 : 1);
    (void)hipFree(Z);
    (void)hipFree(wm);
    return e;
}

}  // namespace

namespace {

// The eight launches of a filter call over n probes (grid dimension y, z for
// the face maxima: probe). delta_chain (single probe, nullable): the chain the
// delta mip samples instead of the probe's chain.
int launch_call(hvx_contract_plan* P, const Probes& Q, uint32_t n, uint32_t chain_levels, hipStream_t s,
                const char* what, const float4* delta_chain = nullptr) {
    hvx_ctx* ctx = P->ctx;
    if (chain_levels <= P->min_level)
        return pfail(ctx, HVX_ERR_INVALID_ARG, "the input chain needs %u levels (has %u); build it with "
                     "hvx_ibl_cube_mips", P->min_level + 1, chain_levels);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess))
        return pfail(ctx, HVX_ERR_HIP, "cannot bind device %d", ctx->device);
    // exponent scales from the finest staged level: coarser levels are its box
    // means and padded texels its bilinear samples, so its maxima bound them
    if (P->stage.total) {
        const uint32_t Sl = std::max(P->S >> P->scale_level, 1u), S2 = Sl * Sl;
        const uint32_t nb = std::min(kFaceMaxBlocks, (S2 + 255) / 256);
        hipLaunchKernelGGL(facemax_kernel, dim3(nb, 6, n), dim3(256), 0, s, Q,
                           hvx::cube_level_offset_floats(P->S, P->scale_level) / 4, Sl, P->fmax);
        hipLaunchKernelGGL(scales_kernel, dim3(n), dim3(64), 0, s, Q, P->fmax, nb, P->sc);
        hipLaunchKernelGGL(stage_all_kernel, dim3((P->stage.total + 255) / 256, n), dim3(256), 0, s, Q, P->stage,
                           P->sc);
    }
    // GEMMs with several probes: groups of kProbeGroup probes per wave share
    // each weight fragment, and the group index is the fastest grid dimension,
    // so groups reading one weight block run together and share it in the caches
    Probes G = Q;
    // single-probe calls stage the B operand in LDS several K steps per
    // barrier (near_gemm_lds_kernel, dense_gemm_lds_kernel<1, ...>): with one
    // barrier per step it paid only when several probes share it (512^2 x 10
    // GGX chain: FP16 near GEMM 203 us unstaged, 246 us staged per step)
    const bool grouped = n > 1;
    const uint32_t dense_wgs = (P->dense.waves + 3) / 4,
                   near_wgs = grouped ? P->near.tiles * kNearParts : near_single_wgs(P->near.tiles);
    const uint32_t groups = grouped ? (n + kProbeGroup - 1) / kProbeGroup : 1;
    G.probe_fast = grouped && std::max(dense_wgs, near_wgs) <= 65535u;
    if (P->dense.n) {
        const dim3 grid = G.probe_fast ? dim3(groups, dense_wgs) : dim3(dense_wgs, groups);
        if (grouped && kDenseLdsGroupMulti > 0)
            hipLaunchKernelGGL((dense_gemm_lds_kernel<kProbeGroup, (kDenseLdsGroupMulti > 0 ? kDenseLdsGroupMulti : 4)>),
                               grid, dim3(256), 0, s, G, P->dense);
        else if (grouped)
            hipLaunchKernelGGL(dense_gemm_staged_kernel<kProbeGroup>, grid, dim3(256), 0, s, G, P->dense);
        else if (kDenseLdsGroup > 0)
            hipLaunchKernelGGL((dense_gemm_lds_kernel<1, (kDenseLdsGroup > 0 ? kDenseLdsGroup : 4)>), grid, dim3(256), 0,
                               s, G, P->dense);
        else
            hipLaunchKernelGGL(dense_gemm_kernel<1>, grid, dim3(256), 0, s, G, P->dense);
        hipLaunchKernelGGL(dense_epilogue_kernel, dim3((P->dense.epi + 255) / 256, n), dim3(256), 0, s, Q, P->dense,
                           P->sc);
    }
    if (P->near.n) {
        const dim3 grid = G.probe_fast ? dim3(groups, near_wgs) : dim3(near_wgs, groups),
                   block(64 * kNearWaves / kNearParts);
        const bool split = P->precision == HVX_PRECISION_BF16_SPLIT;
        if (grouped && split)
            hipLaunchKernelGGL((near_gemm_staged_kernel<true, kProbeGroup>), grid, block, 0, s, G, P->near, P->sc);
        else if (grouped)
            hipLaunchKernelGGL((near_gemm_staged_kernel<false, kProbeGroup>), grid, block, 0, s, G, P->near, P->sc);
        else if (split && kNearLdsGroupSplit > 0)
            hipLaunchKernelGGL((near_gemm_lds_kernel<true, (kNearLdsGroupSplit > 0 ? kNearLdsGroupSplit : 2)>), grid,
                               block, 0, s, G, P->near, P->sc);
        else if (split)
            hipLaunchKernelGGL((near_gemm_kernel<true, 1>), grid, block, 0, s, G, P->near, P->sc);
        else if (kNearLdsGroup > 0)
            hipLaunchKernelGGL((near_gemm_lds_kernel<false, (kNearLdsGroup > 0 ? kNearLdsGroup : 2)>), grid, block, 0,
                               s, G, P->near, P->sc);
        else
            hipLaunchKernelGGL((near_gemm_kernel<false, 1>), grid, block, 0, s, G, P->near, P->sc);
        hipLaunchKernelGGL(far_add_kernel, dim3((P->near.texels + 255) / 256, n), dim3(256), 0, s, Q, P->near);
    }
    if (P->delta_mip != ~0u) {
        const uint32_t So = P->mp[P->delta_mip].L.So;
        Probes D = Q;
        if (delta_chain) D.chain0 = delta_chain;
        hipLaunchKernelGGL(delta_kernel, dim3((6 * So * So + 255) / 256, n), dim3(256), 0, s, D, P->S, So,
                           hvx::cube_level_offset_floats(P->So, P->delta_mip) / 4);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? HVX_OK : pfail(ctx, HVX_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// Hot-texel blocks at least max(0.5, 300 level-0 texel widths) radians
// outside a workgroup's cone of output directions enter its sum as one
// entry: the relative error of that entry is about 10 (block extent /
// distance)^2, under 2e-4 for 4x4 level-0 blocks (rms extent about 1.3
// texel widths); faces below 1024 therefore aggregate little or nothing.
float hot_near(uint32_t S) { return std::max(0.5f, 300.0f * 1.5707963f / (float)S); }

// A single-probe call with the bright-texel split: threshold, clamped level 0
// and its box levels, the hot list, the contraction of the clamped chain (the
// delta mip samples the caller's chain), then the hot texels' exact sum.
int launch_hot_call(hvx_contract_plan* P, const float* d_chain, uint32_t chain_levels, float* d_out, hipStream_t s) {
    hvx_ctx* ctx = P->ctx;
    if (chain_levels <= P->min_level)
        return pfail(ctx, HVX_ERR_INVALID_ARG, "the input chain needs %u levels (has %u); build it with "
                     "hvx_ibl_cube_mips", P->min_level + 1, chain_levels);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess))
        return pfail(ctx, HVX_ERR_HIP, "cannot bind device %d", ctx->device);
    const HotArgs H{reinterpret_cast<const float4*>(d_chain), reinterpret_cast<float4*>(P->hot_chain), P->S,
                    P->hot_k, P->hot_hist, P->hot_cnt, P->hot_T, P->hot_idx, P->hot_list, P->hot_groups};
    const size_t texels = 6ull * P->S * P->S;
    hipLaunchKernelGGL(hot_hist_kernel, dim3((uint32_t)std::min<size_t>(1024, (texels + 1023) / 1024)), dim3(256), 0,
                       s, H);
    hipLaunchKernelGGL(hot_select_kernel, dim3(1), dim3(1024), 0, s, H);
    hipLaunchKernelGGL(hot_compact_kernel, dim3((uint32_t)((texels + 255) / 256)), dim3(256), 0, s, H);
    hipLaunchKernelGGL(hot_list_kernel, dim3(1), dim3(1024), 0, s, H);
    hipLaunchKernelGGL(hot_group_kernel, dim3((hvx::kHotMax + 255) / 256), dim3(256), 0, s, H);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pfail(ctx, HVX_ERR_HIP, "bright-texel split: %s", hipGetErrorString(e));
    if (P->min_level > 0 && (e = hvx::launch_cube_mips(P->hot_chain, P->S, P->min_level + 1, s)) != hipSuccess)
        return pfail(ctx, HVX_ERR_HIP, "clamped chain levels: %s", hipGetErrorString(e));
    const Probes Q{reinterpret_cast<const float4*>(P->hot_chain), reinterpret_cast<float4*>(d_out), nullptr, nullptr,
                   0, 0, 0, 1};
    int st = launch_call(P, Q, 1, chain_levels, s, "hvx_ibl_filter_contract", reinterpret_cast<const float4*>(d_chain));
    if (st) return st;
    HotAdd A{};
    uint32_t first = 0;
    const int ctype = P->type == HVX_FILTER_DIFFUSE ? hvx::CONTRACT_DIFFUSE : hvx::CONTRACT_SPECULAR;
    for (uint32_t m = 0; m < P->mips; ++m) {
        const ContractLevel& L = P->mp[m].L;
        if (L.kind == hvx::CONTRACT_DELTA) continue;
        HotMip& M = A.m[A.nm++];
        M.So = L.So;
        M.first = first;
        M.alpha2 = (float)L.alpha2;
        M.h1 = (float)(0.5 * (1.0 - L.alpha2));
        const double z = hvx::contract_weight_integral(ctype, L.alpha2);
        M.scale = (float)((ctype == hvx::CONTRACT_DIFFUSE ? 1.0 : L.alpha2 / hvx::kContractPi) / z);
        M.out_off = hvx::cube_level_offset_floats(P->So, m) / 4;
        first += (6 * L.So * L.So + 63) / 64 * 64;
    }
    if (A.nm == 0) return HVX_OK;
    A.blocks = first / 64;
    A.list = P->hot_list;
    A.groups = P->hot_groups;
    A.cnt = P->hot_cnt;
    A.theta_near = hot_near(P->S);
    A.out = reinterpret_cast<float4*>(d_out);
    if (ctype == hvx::CONTRACT_DIFFUSE)
        hipLaunchKernelGGL(hot_add_kernel<true>, dim3(A.blocks), dim3(256), 0, s, A);
    else
        hipLaunchKernelGGL(hot_add_kernel<false>, dim3(A.blocks), dim3(256), 0, s, A);
    e = hipGetLastError();
    return e == hipSuccess ? HVX_OK : pfail(ctx, HVX_ERR_HIP, "bright-texel sum: %s", hipGetErrorString(e));
}

}  // namespace

extern "C" {

int hvx_ibl_contract_plan_create(hvx_ctx* ctx, int type, uint32_t face, uint32_t out_face, uint32_t out_mips,
                                 int precision, hvx_contract_plan** out) {
    if (!ctx || !out) return HVX_ERR_NULL;
    *out = nullptr;
    if (type != HVX_FILTER_DIFFUSE && type != HVX_FILTER_SPECULAR)
        return pfail(ctx, HVX_ERR_INVALID_ARG, "unknown filter type %d", type);
    if (precision != HVX_PRECISION_FP16_SCALED && precision != HVX_PRECISION_BF16_SPLIT)
        return pfail(ctx, HVX_ERR_INVALID_ARG, "unknown precision %d", precision);
    const bool split = precision == HVX_PRECISION_BF16_SPLIT;
    if (type == HVX_FILTER_DIFFUSE && out_mips != 1)
        return pfail(ctx, HVX_ERR_INVALID_ARG, "the diffuse filter writes one mip");
    if (face < 4 || face > 8192 || (face & (face - 1)) || out_face < 1 || out_face > 4096 ||
        (out_face & (out_face - 1)) || out_mips < 1 || out_mips > hvx_mip_count(out_face) ||
        out_mips > hvx::kContractMaxMips)
        return pfail(ctx, HVX_ERR_INVALID_ARG, "sizes: face %u and out_face %u must be powers of two (face >= 4), "
                     "out_mips in [1, mips(out_face)]", face, out_face);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess))
        return pfail(ctx, HVX_ERR_HIP, "cannot bind device %d", ctx->device);
    hvx_contract_plan* P = new (std::nothrow) hvx_contract_plan();
    if (!P) return HVX_ERR_HIP;
    P->ctx = ctx;
    P->type = type;
    P->precision = precision;
    P->S = face;
    P->So = out_face;
    P->mips = out_mips;
    ContractLevel lv[hvx::kContractMaxMips];
    const int ctype = type == HVX_FILTER_DIFFUSE ? hvx::CONTRACT_DIFFUSE : hvx::CONTRACT_SPECULAR;
    hvx::contract_levels(ctype, face, out_face, out_mips, lv);
    auto fail_destroy = [&](hipError_t e, const char* where) {
        hvx_ibl_contract_plan_destroy(P);
        return pfail(ctx, HVX_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
    };
    hipError_t e;
    size_t part_floats = 0, sbytes = 0;
    P->mp.resize(out_mips);
    for (uint32_t m = 0; m < out_mips; ++m) {
        MipPlan& M = P->mp[m];
        M.L = lv[m];
        if (M.L.kind == hvx::CONTRACT_DELTA) continue;
        DenseMat D;
        D.G.type = ctype;
        D.G.alpha2 = M.L.alpha2;
        D.G.t0 = M.L.t0;
        D.G.t1 = M.L.t1;
        if (M.L.kind == hvx::CONTRACT_FULL) {
            D.G.piece = PIECE_WHOLE;
            D.G.So = M.L.So;
            D.G.Sin = M.L.Si;
        } else {
            D.G.piece = PIECE_FAR;
            D.G.So = M.L.Sf;
            D.G.Sin = M.L.Sf;
            M.far_off = scratch_take(&sbytes, 6 * (size_t)M.L.Sf * M.L.Sf * sizeof(float4));
            if ((e = P->alloc(&M.zfar_map, 6 * (size_t)M.L.Sf * M.L.Sf)) != hipSuccess)
                return fail_destroy(e, "far normaliser map");
        }
        D.G.rows = D.G.So * D.G.So;
        D.G.rt = (D.G.rows + 15) / 16;
        D.G.ks = 6 * D.G.Sin * D.G.Sin / 32;
        D.split = split && M.L.kind == hvx::CONTRACT_SPLIT;  // low-roughness mips: near and far
        D.stage = stage_dense_index(P, D.G.Sin, D.split);
        if ((e = build_dense(P, D, ctx->num_cus, M.zfar_map)) != hipSuccess) return fail_destroy(e, "dense weights");
        M.weight_bytes += (uint64_t)D.G.rt * D.G.ks * 64 * 16 * (D.split ? 2 : 1);
        part_floats += (size_t)D.splitk * D.G.rt * 16 * 32;
        P->dm.push_back(D);
        M.dense = (int)P->dm.size() - 1;
        if (M.L.kind == hvx::CONTRACT_SPLIT) {
            M.NG = NearGeom{ctype, M.L.alpha2, M.L.t0, M.L.t1, M.L.So, M.L.Si, M.L.So / 16};
            uint32_t B = 0;
            M.tiles = near_tiles(M.L.So, M.L.Si, M.L.t1, &B);
            StagePad sp;
            sp.Si = M.L.Si;
            sp.level = ilog2(face / M.L.Si);
            sp.B = B;
            sp.Pw = M.L.Si + 2 * B;
            P->sp.push_back(sp);
            M.pad = (int)P->sp.size() - 1;
            if ((e = build_near(P, M, split)) != hipSuccess) return fail_destroy(e, "near weights");
        }
    }
    // staged inputs and shared scratch
    for (StageDense& s : P->sd) {
        const uint32_t K = 6 * s.Sin * s.Sin;
        std::vector<int32_t> perm((size_t)6 * K);
        for (int f = 0; f < 6; ++f)
            for (uint32_t k = 0; k < K; ++k) perm[(size_t)f * K + k] = perm_index(f, k, s.Sin);
        if ((e = P->alloc(&s.perm, perm.size())) != hipSuccess ||
            (e = hipMemcpy(s.perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
            return fail_destroy(e, "dense staging");
        s.off = scratch_take(&sbytes, (size_t)K / 8 * 32 * (s.split ? 2 : 1) * sizeof(half8));
        P->min_level = std::max(P->min_level, s.level);
    }
    for (StagePad& s : P->sp) {
        s.off = scratch_take(&sbytes, (size_t)s.Pw * (s.Pw / 8) * 32 * (split ? 2 : 1) * sizeof(half8));
        P->min_level = std::max(P->min_level, s.level);
    }
    const size_t fmax_off = scratch_take(&sbytes, 6 * kFaceMaxBlocks * sizeof(float));
    const size_t sc_off = scratch_take(&sbytes, 16 * sizeof(float));
    const size_t part_off = scratch_take(&sbytes, part_floats * sizeof(float));
    // one block; zeroed once (the pad columns of staged operands stay zero)
    P->scratch_bytes = (sbytes + 255) & ~(size_t)255;
    if ((e = P->alloc(&P->scratch, P->scratch_bytes)) != hipSuccess ||
        (e = hipMemset(P->scratch, 0, P->scratch_bytes)) != hipSuccess)
        return fail_destroy(e, "scratch");
    for (MipPlan& M : P->mp)
        if (M.L.kind == hvx::CONTRACT_SPLIT) M.far_map = reinterpret_cast<float4*>(P->scratch + M.far_off);
    for (StageDense& s : P->sd) s.R = reinterpret_cast<half8*>(P->scratch + s.off);
    for (StagePad& s : P->sp) s.P = reinterpret_cast<half8*>(P->scratch + s.off);
    P->fmax = reinterpret_cast<float*>(P->scratch + fmax_off);
    P->sc = reinterpret_cast<float*>(P->scratch + sc_off);
    P->part = reinterpret_cast<float*>(P->scratch + part_off);
    // launch tables of a filter call: every staged level, dense job and near job
    StageTable& ST = P->stage;
    for (const StageDense& sd : P->sd) P->scale_level = std::min(P->scale_level, sd.level);
    for (const StagePad& sp : P->sp) P->scale_level = std::min(P->scale_level, sp.level);
    for (const StageDense& sd : P->sd) {
        StageItem& I = ST.it[ST.n++];
        I = StageItem{sd.split ? 3 : 0, sd.Sin, 0, 0, hvx::cube_level_offset_floats(face, sd.level) / 4, sd.perm,
                      sd.R, ST.total};
        ST.total += 6 * sd.Sin * sd.Sin * 6;
    }
    for (const StagePad& sp : P->sp) {
        StageItem& I = ST.it[ST.n++];
        I = StageItem{split ? 2 : 1, sp.Si, sp.B, sp.Pw, hvx::cube_level_offset_floats(face, sp.level) / 4, nullptr,
                      sp.P, ST.total};
        ST.total += sp.Pw * sp.Pw * 6;
    }
    DenseTable& DT = P->dense;
    NearTable& NT = P->near;
    double near_k[kMaxJobs] = {};  // mean window (K) of each near job
    float* part = P->part;
    for (uint32_t m = 0; m < out_mips; ++m) {
        const MipPlan& M = P->mp[m];
        if (M.L.kind == hvx::CONTRACT_DELTA) {
            P->delta_mip = m;
            continue;
        }
        const DenseMat& D = P->dm[M.dense];
        DenseJob& J = DT.j[DT.n++];
        J.W = D.W;
        J.R = P->sd[D.stage].R;
        J.rt = D.G.rt;
        J.ks = D.G.ks;
        J.splitk = D.splitk;
        J.rows = D.G.rows;
        J.part = part;
        part += (size_t)D.splitk * D.G.rt * 16 * 32;
        J.rowscale = D.rowscale;
        J.out = M.L.kind == hvx::CONTRACT_FULL ? nullptr : M.far_map;
        J.out_off = hvx::cube_level_offset_floats(out_face, m) / 4;
        J.alpha = M.L.kind == hvx::CONTRACT_FULL ? 1.0f : 0.0f;
        J.wave_begin = DT.waves;
        J.epi_begin = DT.epi;
        J.w_lo = D.split ? (uint64_t)D.G.rt * D.G.ks * 64 : 0;
        J.r_lo = D.split ? (uint64_t)6 * D.G.Sin * D.G.Sin / 8 * 32 : 0;
        DT.waves += ((D.G.rt + 1) / 2 + 3) / 4 * 4 * D.splitk;  // pairs padded to a multiple of four
        DT.epi += 6 * D.G.rows;
        if (M.L.kind == hvx::CONTRACT_SPLIT) {
            const StagePad& sp = P->sp[M.pad];
            NearJob& N = NT.j[NT.n++];
            N = NearJob{M.Wn, M.d_tiles, sp.P, sp.Pw / 8, sp.B, M.NG.te, M.L.So, M.L.Sf, M.rowscale_n, M.invZ,
                        M.far_map, hvx::cube_level_offset_floats(out_face, m) / 4, NT.tiles, NT.texels, M.w_lo,
                        split ? (uint64_t)sp.Pw * (sp.Pw / 8) * 32 : 0};
            NT.tiles += (uint32_t)M.tiles.size();
            NT.texels += 6 * M.L.So * M.L.So;
            double k = 0.0;
            for (const NearTile& t : M.tiles) k += (double)t.w * t.h;
            near_k[NT.n - 1] = k / (double)M.tiles.size();
        }
    }
    // near jobs in decreasing mean window size: the workgroups with the most
    // K steps start first and the many shorter ones fill in behind them (the
    // per-output accumulation does not depend on the order)
    {
        std::vector<uint32_t> ord(NT.n);
        for (uint32_t k = 0; k < NT.n; ++k) ord[k] = k;
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return near_k[a] > near_k[b]; });
        NearTable sorted = NT;
        uint32_t tb = 0, xb = 0;
        for (uint32_t k = 0; k < NT.n; ++k) {
            sorted.j[k] = NT.j[ord[k]];
            sorted.j[k].tile_begin = tb;
            sorted.j[k].texel_begin = xb;
            tb += sorted.j[k].te * sorted.j[k].te;
            xb += 6 * sorted.j[k].So * sorted.j[k].So;
        }
        NT = sorted;
    }
    if ((e = hipDeviceSynchronize()) != hipSuccess) return fail_destroy(e, "plan set-up");
    *out = P;
    return HVX_OK;
}

int hvx_ibl_contract_plan_destroy(hvx_contract_plan* P) {
    if (!P) return HVX_ERR_NULL;
    (void)hipDeviceSynchronize();
    for (void* p : P->allocs) (void)hipFree(p);
    if (P->mscratch) (void)hipFree(P->mscratch);
    delete P;
    return HVX_OK;
}

int hvx_ibl_contract_plan_query(const hvx_contract_plan* P, uint32_t mip, hvx_contract_mip_info* info) {
    if (!P || !info) return HVX_ERR_NULL;
    if (mip >= P->mips) return HVX_ERR_INVALID_ARG;
    const MipPlan& M = P->mp[mip];
    info->kind = M.L.kind;
    info->out_face = M.L.So;
    info->in_face = M.L.Si;
    info->far_face = M.L.Sf;
    info->window_t0 = (float)M.L.t0;
    info->window_t1 = (float)M.L.t1;
    info->weight_bytes = M.weight_bytes;
    info->min_input_level = P->min_level;
    return HVX_OK;
}

int hvx_ibl_filter_contract(hvx_contract_plan* P, const float* d_chain, uint32_t chain_levels, float* d_out,
                            void* stream) {
    if (!P) return HVX_ERR_NULL;
    hvx_ctx* ctx = P->ctx;
    if (!d_chain || !d_out) return pfail(ctx, HVX_ERR_NULL, "NULL chain or output");
    if (((uintptr_t)d_chain | (uintptr_t)d_out) % 16) return pfail(ctx, HVX_ERR_ALIGN, "chains need 16-byte alignment");
    if (P->hot_k) return launch_hot_call(P, d_chain, chain_levels, d_out, (hipStream_t)stream);
    const Probes Q{reinterpret_cast<const float4*>(d_chain), reinterpret_cast<float4*>(d_out), nullptr, nullptr, 0, 0, 0, 1};
    return launch_call(P, Q, 1, chain_levels, (hipStream_t)stream, "hvx_ibl_filter_contract");
}

int hvx_ibl_contract_plan_set_hot(hvx_contract_plan* P, uint32_t max_hot) {
    if (!P) return HVX_ERR_NULL;
    if (max_hot > hvx::kHotMax)
        return pfail(P->ctx, HVX_ERR_INVALID_ARG, "max_hot must be in [0, %u]", hvx::kHotMax);
    if (max_hot && !P->hot_chain) {
        hipError_t e;
        if ((e = P->alloc(&P->hot_chain, hvx::cube_level_offset_floats(P->S, P->min_level + 1))) != hipSuccess ||
            (e = P->alloc(&P->hot_hist, hvx::kHotBins)) != hipSuccess ||
            (e = P->alloc(&P->hot_cnt, 3)) != hipSuccess || (e = P->alloc(&P->hot_T, 1)) != hipSuccess ||
            (e = P->alloc(&P->hot_idx, hvx::kHotMax)) != hipSuccess ||
            (e = P->alloc(&P->hot_list, 2 * hvx::kHotMax)) != hipSuccess ||
            (e = P->alloc(&P->hot_groups, 2 * hvx::kHotMax)) != hipSuccess ||
            (e = hipMemset(P->hot_hist, 0, hvx::kHotBins * sizeof(uint32_t))) != hipSuccess ||
            (e = hipMemset(P->hot_cnt, 0, 3 * sizeof(uint32_t))) != hipSuccess ||
            (e = hipDeviceSynchronize()) != hipSuccess) {
            P->hot_chain = nullptr;  // the partial allocations are released with the plan
            return pfail(P->ctx, HVX_ERR_HIP, "bright-texel split buffers: %s", hipGetErrorString(e));
        }
    }
    P->hot_k = max_hot;
    return HVX_OK;
}

int hvx_ibl_contract_hot_info(hvx_contract_plan* P, float* threshold, uint32_t* n_hot) {
    if (!P || !threshold || !n_hot) return HVX_ERR_NULL;
    if (!P->hot_chain) return pfail(P->ctx, HVX_ERR_INVALID_ARG, "the bright-texel split was never enabled");
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(threshold, P->hot_T, sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(n_hot, P->hot_cnt + 1, sizeof(uint32_t), hipMemcpyDeviceToHost);
    return e == hipSuccess ? HVX_OK : pfail(P->ctx, HVX_ERR_HIP, "hvx_ibl_contract_hot_info: %s", hipGetErrorString(e));
}

int hvx_ibl_contract_plan_reserve(hvx_contract_plan* P, uint32_t max_probes) {
    if (!P) return HVX_ERR_NULL;
    if (max_probes == 0 || max_probes > kMaxProbes)
        return pfail(P->ctx, HVX_ERR_INVALID_ARG, "max_probes must be in [1, %u]", kMaxProbes);
    if (max_probes <= P->mcap) return HVX_OK;
    (void)hipDeviceSynchronize();  // the old copies may be in use
    if (P->mscratch) (void)hipFree(P->mscratch);
    P->mscratch = nullptr;
    P->mcap = 0;
    const size_t bytes = P->scratch_bytes * max_probes;
    hipError_t e;
    if ((e = hipMalloc(reinterpret_cast<void**>(&P->mscratch), bytes)) != hipSuccess ||
        (e = hipMemset(P->mscratch, 0, bytes)) != hipSuccess) {
        if (P->mscratch) (void)hipFree(P->mscratch);
        P->mscratch = nullptr;
        return pfail(P->ctx, HVX_ERR_HIP, "multi-probe scratch (%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    P->mcap = max_probes;
    return HVX_OK;
}

int hvx_ibl_filter_contract_multi(hvx_contract_plan* P, uint32_t n, const float* const* d_chains,
                                  uint32_t chain_levels, float* const* d_outs, void* stream) {
    if (!P) return HVX_ERR_NULL;
    if (!d_chains || !d_outs) return pfail(P->ctx, HVX_ERR_NULL, "NULL pointer arrays");
    if (n == 0) return HVX_OK;
    if (n > P->mcap)
        return pfail(P->ctx, HVX_ERR_CAPACITY, "%u probes but scratch reserved for %u (hvx_ibl_contract_plan_reserve)",
                     n, P->mcap);
    if (P->hot_k)
        return pfail(P->ctx, HVX_ERR_UNSUPPORTED, "the bright-texel split serves single-probe calls only");
    const Probes Q{nullptr, nullptr, reinterpret_cast<const float4* const*>(d_chains),
                   reinterpret_cast<float4* const*>(d_outs),
                   (uint64_t)(reinterpret_cast<uintptr_t>(P->mscratch) - reinterpret_cast<uintptr_t>(P->scratch)),
                   (uint64_t)P->scratch_bytes, 0, n};
    return launch_call(P, Q, n, chain_levels, (hipStream_t)stream, "hvx_ibl_filter_contract_multi");
}

}  // extern "C"
