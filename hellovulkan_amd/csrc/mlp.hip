// mlp.hip — the per-environment radiance network (hvx_mlp_* of
// include/hvx.h), trained with hand-written BF16 MFMA forward and backward
// kernels and FP32 master weights (Adam).
//
// One training step over a batch of B samples (B a multiple of 256):
//   pe            X0 = PE(d, lod)                        bf16 [B][128] (+ transpose);
//                 once per training call when the call revisits its samples
//   gemm FWD x3   H_l = relu(H_{l-1} W_l^T + b_l)        bf16 [B][64]  (+ transpose)
//   out           y = H3 W4^T + b4, loss, dY, dZ3 = (dY W4) * [H3 > 0], dW4/db4 partials
//   per layer l = 3, 2, 1:
//     gemm PART   dW_l = dZ_l^T H_{l-1}, db_l = sum_b dZ_l   (K = batch, split into
//                 128-sample slices over waves, reduced in slice order)
//     gemm BWD    dZ_{l-1} = (dZ_l W_l) * [H_{l-1} > 0]   (l > 1)
//   adam          FP32 master update and the BF16 copies of W1..W3 (and W2^T, W3^T)
// Every GEMM is C = A Bt^T with both operands row-major along K, so each MFMA
// operand load is one contiguous 16-byte read; the epilogues also write the
// transposed activation that the weight-gradient GEMMs read that way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "hvx_device.h"
#include "hvx_internal.h"
#include "hvx_contract.h"
#include "hvx_mfma.h"
#include "hvx_mlp.h"
#include "mlp_device.h"

namespace {

using namespace hvx;
typedef __bf16 bf16;

constexpr int kIn = hvx::kMlpIn, kH = hvx::kMlpHidden;
constexpr uint32_t kMaxBatch = 1u << 20;

// ------------------------------------------------------------------- data
__device__ __forceinline__ uint64_t smix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ float u01(uint64_t seed, uint64_t stream, uint64_t i) {
    return (float)((double)(smix(seed + (stream << 40) + i) >> 11) * (1.0 / 9007199254740992.0));
}

// textureLod of the prefiltered chain (trilinear, LOD clamped to [0, mips-1])
__device__ __forceinline__ f3 chain_lod(const float4* chain, uint32_t S, uint32_t mips, f3 dir, float lod) {
    const float q = (float)(mips - 1);
    const float lam = gclamp(lod, 0.0f, q);
    const float fl = floorf_(lam);
    const uint32_t l0 = (uint32_t)fl, l1 = min(l0 + 1, mips - 1);
    const float d = lam - fl;
    const int s0 = (int)max(S >> l0, 1u), s1 = (int)max(S >> l1, 1u);
    const f3 c0 = sample_cube(chain + cube_level_offset_floats(S, l0) / 4, s0, dir);
    const f3 c1 = sample_cube(chain + cube_level_offset_floats(S, l1) / 4, s1, dir);
    return {(1.0f - d) * c0.x + d * c1.x, (1.0f - d) * c0.y + d * c1.y, (1.0f - d) * c0.z + d * c1.z};
}

// n samples: direction uniform on the sphere, lod uniform in [0, lod_max];
// target log(textureLod(chain, d, lod) + eps) per channel
__global__ __launch_bounds__(256) void mlp_data_kernel(const float4* __restrict__ chain, uint32_t S, uint32_t mips,
                                                       float lod_max, uint64_t seed, uint32_t n,
                                                       float4* __restrict__ in, float4* __restrict__ tgt) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float z = 2.0f * u01(seed, 0, i) - 1.0f, ph = 6.2831853071795864f * u01(seed, 1, i);
    const float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    const f3 d = {r * cosf(ph), r * sinf(ph), z};
    const float lod = lod_max * u01(seed, 2, i);
    const f3 c = chain_lod(chain, S, mips, d, lod);
    const float e = (float)hvx::kMlpEps;
    in[i] = make_float4(d.x, d.y, d.z, lod);
    tgt[i] = make_float4(logf(fmaxf(c.x, 0.0f) + e), logf(fmaxf(c.y, 0.0f) + e), logf(fmaxf(c.z, 0.0f) + e), 0.0f);
}

// ----------------------------------------------------------------- encoding
// thread per (sample, input j, frequency k); rows >= n are the encoding of 0
// One thread per sample: its 128 features (row of X, sixteen 16-byte
// stores) and their transposed copies (XT, coalesced across the samples).
__global__ __launch_bounds__(64) void mlp_pe_kernel(const float4* __restrict__ in, uint32_t n, uint32_t B,
                                                    float lod_max, bf16* __restrict__ X, bf16* __restrict__ XT) {
    const uint32_t b = blockIdx.x * 64 + threadIdx.x;
    if (b >= B) return;
    float xin[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (b < n) {
        const float4 v = in[b];
        xin[0] = v.x;
        xin[1] = v.y;
        xin[2] = v.z;
        xin[3] = 2.0f * v.w / lod_max - 1.0f;
    }
    bf16x8* xrow = reinterpret_cast<bf16x8*>(X + (size_t)b * kIn);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            // frequencies 4 g .. 4 g + 3 of input j, as mlp_sincos evaluates them
            float sv[4], cv[4];
            hvx::mlp_sincos_direct(xin[j], 4 * g, sv[0], cv[0]);
            hvx::mlp_sincos_direct(xin[j], 4 * g + 1, sv[1], cv[1]);
            hvx::mlp_sincos_double(sv[0], cv[0], sv[2], cv[2]);
            hvx::mlp_sincos_double(sv[1], cv[1], sv[3], cv[3]);
            bf16x8 pk;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                pk[2 * q] = (bf16)sv[q];
                pk[2 * q + 1] = (bf16)cv[q];
            }
            xrow[4 * j + g] = pk;
            const uint32_t c0 = 32 * j + 8 * g;
#pragma unroll
            for (int q = 0; q < 8; ++q) XT[(size_t)(c0 + q) * B + b] = pk[q];
        }
}

// -------------------------------------------------------------------- GEMM
enum Epi { EPI_FWD = 0, EPI_BWD = 1, EPI_PART = 2 };

struct GemmArgs {
    const bf16* A;   // [M][Kd]
    const bf16* Bt;  // [Nc][Kd]
    uint32_t M, Nc, Kd, splitk;
    uint32_t lda, ldb;  // row strides of A and Bt in elements (0: Kd)
    const float* bias;  // FWD: [Nc]
    const bf16* mask;   // BWD: [M][Nc], zero where <= 0
    bf16* C;            // FWD/BWD: [M][Nc]
    bf16* CT;           // FWD/BWD: [Nc][M]
    float* part;        // PART: per slice [M][Nc] then (colsum) [M], slices pstride apart
    uint32_t pstride;   // PART: floats per slice
    uint32_t colsum;    // PART: also sum A's rows over K (the bias gradient)
};

// A wave computes a 16 x 64 block of C over one K slice.
template <int EPI>
__device__ __forceinline__ void gemm_abt_wave(const GemmArgs& g, uint32_t wave, uint32_t lane) {
    const uint32_t rts = g.M / 16, cgs = g.Nc / 64;
    if (wave >= rts * cgs * g.splitk) return;  // uniform per wave; no barriers below
    const uint32_t slice = wave % g.splitk, t = wave / g.splitk, cg = t % cgs, rt = t / cgs;
    const uint32_t ks = g.Kd / 32;
    const uint32_t kb = slice * ks / g.splitk, ke = (slice + 1) * ks / g.splitk;
    const uint32_t i = lane & 15, gq = lane >> 4;
    const size_t lda = g.lda ? g.lda : g.Kd, ldb = g.ldb ? g.ldb : g.Kd;
    const bf16* ap = g.A + (size_t)(rt * 16 + i) * lda + 8 * gq;
    const bf16* bp = g.Bt + (size_t)(cg * 64 + i) * ldb + 8 * gq;
    const size_t bstep = (size_t)16 * ldb;
    f32x4 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // PART with colsum: the waves of column group 0 also sum the A elements
    // they load (row rt * 16 + i, this lane's 8 of every 32 K)
    const bool rowsum = EPI == EPI_PART && g.colsum && cg == 0;
    float rs = 0.0f;
    for (uint32_t s = kb; s < ke; ++s) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(ap + 32 * s);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(bp + nb * bstep + 32 * s);
            acc[nb] = hvx::mfma16x16x32(a, b, acc[nb]);
        }
        if (rowsum) {
            float t = 0.0f;
#pragma unroll
            for (int q = 0; q < 8; ++q) t += (float)a[q];
            rs += t;
        }
    }
    if (rowsum) {  // the four lanes of row i hold disjoint K subsets
        rs += __shfl_xor(rs, 16, 64);
        rs += __shfl_xor(rs, 32, 64);
        if (gq == 0) g.part[(size_t)slice * g.pstride + g.M * g.Nc + rt * 16 + i] = rs;
    }
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const uint32_t col = cg * 64 + nb * 16 + i;
        bf16 ht[4];  // rows 4 gq .. 4 gq + 3 of column col: one 8-byte store into CT
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t row = rt * 16 + 4 * gq + r;
            float v = acc[nb][r];
            if (EPI == EPI_PART) {
                g.part[(size_t)slice * g.pstride + (size_t)row * g.Nc + col] = v;
            } else {
                if (EPI == EPI_FWD) v = fmaxf(v + g.bias[col], 0.0f);
                else if (!((float)g.mask[(size_t)row * g.Nc + col] > 0.0f)) v = 0.0f;
                const bf16 h = (bf16)v;
                g.C[(size_t)row * g.Nc + col] = h;
                ht[r] = h;
            }
        }
        if (EPI != EPI_PART) {
            typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
            *reinterpret_cast<bf16x4*>(g.CT + (size_t)col * g.M + rt * 16 + 4 * gq) = bf16x4{ht[0], ht[1], ht[2], ht[3]};
        }
    }
}

template <int EPI>
__global__ __launch_bounds__(256) void gemm_abt_kernel(GemmArgs g) {
    gemm_abt_wave<EPI>(g, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

// The three weight-gradient GEMMs of a step in one launch: waves [0, w1)
// compute g[0], [w1, w2) g[1], the rest g[2], each as gemm_abt_kernel does.
struct GemmArgs3 {
    GemmArgs g[3];
    uint32_t w1, w2;
};
__global__ __launch_bounds__(256) void gemm_part3_kernel(GemmArgs3 a) {
    const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (wave < a.w1) gemm_abt_wave<EPI_PART>(a.g[0], wave, lane);
    else if (wave < a.w2) gemm_abt_wave<EPI_PART>(a.g[1], wave - a.w1, lane);
    else gemm_abt_wave<EPI_PART>(a.g[2], wave - a.w2, lane);
}

// ------------------------------------------------------- output and loss
constexpr int kOutPart = kH * 3 + 3 + 1;  // dW4, db4, loss
// samples per output-layer block: the loss and the output layer's gradient
// are summed per block of this many samples, then over the blocks in order
constexpr uint32_t kOutBlock = 128;

// Training: one thread per sample (B a multiple of kOutBlock). Evaluation
// (TRAIN false): writes radiance exp(y) - eps for samples < n.
template <bool TRAIN>
__global__ __launch_bounds__(kOutBlock) void mlp_out_kernel(const bf16* __restrict__ H3, const float* __restrict__ P,
                                                      const float4* __restrict__ tgt, uint32_t n, uint32_t B,
                                                      float inv_norm, float* __restrict__ Y, bf16* __restrict__ dZ,
                                                      bf16* __restrict__ dZT, float* __restrict__ part) {
    __shared__ float sdY[kOutBlock][3];
    __shared__ float sl[kOutBlock];
    __shared__ bf16 sh3[kH][kOutBlock + 8];  // the block's H3 transposed (padded rows)
    const uint32_t t = threadIdx.x, b = blockIdx.x * kOutBlock + t;
    const float* W4 = P + hvx::kMlpW4;
    const float* b4 = P + hvx::kMlpB4;
    float h[kH];
    const bf16x8* hrow = reinterpret_cast<const bf16x8*>(H3 + (size_t)b * kH);
#pragma unroll
    for (int k8 = 0; k8 < kH / 8; ++k8) {
        const bf16x8 v = hrow[k8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            h[8 * k8 + q] = (float)v[q];
            if (TRAIN) sh3[8 * k8 + q][t] = v[q];
        }
    }
    float y[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) y[c] = hvx::mlp_output(h, W4 + c * kH, b4[c]);
    if (!TRAIN) {
        if (b < n)
            for (int c = 0; c < 3; ++c) Y[(size_t)b * 3 + c] = expf(y[c]) - (float)hvx::kMlpEps;
        return;
    }
    const float4 tv = tgt[b];
    const float e0 = y[0] - tv.x, e1 = y[1] - tv.y, e2 = y[2] - tv.z;
    sdY[t][0] = e0 * inv_norm;
    sdY[t][1] = e1 * inv_norm;
    sdY[t][2] = e2 * inv_norm;
    sl[t] = (e0 * e0 + e1 * e1) + e2 * e2;
#pragma unroll
    for (int k8 = 0; k8 < kH / 8; ++k8) {
        bf16x8 pk;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = 8 * k8 + q;
            const float dh = (sdY[t][0] * W4[k] + sdY[t][1] * W4[kH + k]) + sdY[t][2] * W4[2 * kH + k];
            const bf16 v = (bf16)(h[k] > 0.0f ? dh : 0.0f);
            pk[q] = v;
            dZT[(size_t)k * B + b] = v;
        }
        reinterpret_cast<bf16x8*>(dZ + (size_t)b * kH)[k8] = pk;
    }
    __syncthreads();
    float* pp = part + (size_t)blockIdx.x * kOutPart;
    for (uint32_t o = t; o < (uint32_t)kOutPart - 1; o += kOutBlock) {
        float acc = 0.0f;
        if (o < 3 * kH) {
            const uint32_t c = o / kH, k = o % kH;
            for (uint32_t j = 0; j < kOutBlock; ++j) acc += sdY[j][c] * (float)sh3[k][j];
        } else {
            const uint32_t c = o - 3 * kH;
            for (uint32_t j = 0; j < kOutBlock; ++j) acc += sdY[j][c];
        }
        pp[o] = acc;
    }
    for (uint32_t hh = kOutBlock / 2; hh > 0; hh >>= 1) {
        __syncthreads();
        if (t < hh) sl[t] += sl[t + hh];
    }
    if (t == 0) pp[kOutPart - 1] = sl[0];
}


// The output layer's block partials (kOutPart per block: dW4, db4, loss):
// a workgroup takes 64 of them and 16 block ranges (one per wave), each
// summed in block order, then adds the 16 range sums in a fixed tree; dW4
// and db4 into out, the loss (scaled) into *loss when asked for
constexpr uint32_t kOutRedGroups = 16;
__global__ __launch_bounds__(64 * kOutRedGroups) void reduce_out_kernel(const float* __restrict__ part,
                                                                        uint32_t blocks, float* __restrict__ out,
                                                                        float loss_scale, float* __restrict__ loss) {
    __shared__ float sh[kOutRedGroups][64];
    const uint32_t l = threadIdx.x & 63, gq = threadIdx.x >> 6, i = blockIdx.x * 64 + l;
    float acc = 0.0f;
    if (i < (uint32_t)kOutPart) {
        const uint32_t s0 = gq * blocks / kOutRedGroups, s1 = (gq + 1) * blocks / kOutRedGroups;
#pragma unroll 4
        for (uint32_t s = s0; s < s1; ++s) acc += part[(size_t)s * kOutPart + i];
    }
    sh[gq][l] = acc;
    for (uint32_t h = kOutRedGroups / 2; h > 0; h >>= 1) {
        __syncthreads();
        if (gq < h) sh[gq][l] += sh[gq + h][l];
    }
    if (gq != 0 || i >= (uint32_t)kOutPart) return;
    if (i < (uint32_t)kOutPart - 1) out[i] = sh[0][l];
    else if (loss) *loss = sh[0][l] * loss_scale;
}

// ------------------------------------------------ fused per-sample step
// The per-sample part of a training step for a block of kOutBlock samples
// (eight waves, one 16-sample row tile each): the three forward layers, the
// output layer with the loss and the block's output-layer partials, and the
// two backward products dZ2 = (dZ3 W3) [H2 > 0], dZ1 = (dZ2 W2) [H1 > 0],
// with the activations kept in LDS between them. Every product is the MFMA
// sequence of gemm_abt_kernel over the same BF16 operands and the output
// part is mlp_out_kernel's, so the results equal those separate launches
// bit for bit; the activations and gradients the weight-gradient GEMMs read
// (H_l, dZ_l and their transposes) are written as they write them.
struct StepArgs {
    const bf16* X;  // [B][kIn] encodings of the batch
    const float* P;
    const bf16 *W[3], *W2T, *W3T;
    const float4* tgt;
    uint32_t B;
    float inv_norm;
    bf16 *H[3], *HT[3];
    bf16 *dZ[3], *dZT[3];  // dZ1, dZ2, dZ3 and their transposes
    float* part;            // [B / kOutBlock][kOutPart]
};
constexpr int kHp = kH + 8;  // LDS row of 64 units (bf16), padded

// C block (16 rows x 64 columns) of A Bt^T: row i of A at arow (this lane's
// 8 of each 32 K), Bt row i at bt; gemm_abt_kernel's sequence of MFMAs
template <int KD>
__device__ __forceinline__ void step_tile(const bf16* arow, const bf16* bt, f32x4 acc[4]) {
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KD / 32; ++s) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(arow + 32 * s);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(bt + nb * 16 * KD + 32 * s);
            acc[nb] = hvx::mfma16x16x32(a, b, acc[nb]);
        }
    }
}

// Epilogue of one C block (local rows lr0 .. lr0 + 15 of the block): FWD
// (bias and ReLU) or BWD (zero where the mask activation is not positive),
// rounded to BF16 into C, CT (global) and the LDS rows out
template <bool FWD>
__device__ __forceinline__ void step_store(const f32x4 acc[4], uint32_t b0, uint32_t lr0, uint32_t B, const float* bias,
                                           const bf16 (*mask)[kHp], bf16* C, bf16* CT, bf16 (*out)[kHp]) {
    const uint32_t lane = threadIdx.x & 63, i = lane & 15, gq = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const uint32_t col = nb * 16 + i;
        bf16 ht[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t lrow = lr0 + 4 * gq + r;
            float v = acc[nb][r];
            if (FWD) v = fmaxf(v + bias[col], 0.0f);
            else if (!((float)mask[lrow][col] > 0.0f)) v = 0.0f;
            const bf16 h = (bf16)v;
            C[(size_t)(b0 + lrow) * kH + col] = h;
            if (out) out[lrow][col] = h;
            ht[r] = h;
        }
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        *reinterpret_cast<bf16x4*>(CT + (size_t)col * B + b0 + lr0 + 4 * gq) = bf16x4{ht[0], ht[1], ht[2], ht[3]};
    }
}

constexpr uint32_t kStepWaves = 8;                           // waves per block
constexpr uint32_t kStepTiles = kOutBlock / 16 / kStepWaves;  // 16-sample row tiles per wave
static_assert(kStepTiles >= 1 && kStepTiles * 16 * kStepWaves == kOutBlock, "whole row tiles per wave");
__global__ __launch_bounds__(64 * kStepWaves) void mlp_step_kernel(StepArgs g) {
    __shared__ __align__(16) bf16 sH[3][kOutBlock][kHp];  // H1, H2, H3; H3's rows later hold dZ2
    __shared__ __align__(16) bf16 sZ[kOutBlock][kHp];     // dZ3
    __shared__ float sdY[kOutBlock][3];
    __shared__ float sl[kOutBlock];
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63, i = lane & 15, gq = lane >> 4;
    const uint32_t b0 = blockIdx.x * kOutBlock, B = g.B;
    const int bias[3] = {hvx::kMlpB1, hvx::kMlpB2, hvx::kMlpB3};
    // forward: each wave its own rows, its row tiles interleaved
#pragma unroll 1
    for (int l = 0; l < 3; ++l) {
        f32x4 acc[kStepTiles][4];
#pragma unroll
        for (uint32_t rt = 0; rt < kStepTiles; ++rt) {
            const uint32_t lr0 = (wave * kStepTiles + rt) * 16;
            if (l == 0)
                step_tile<kIn>(g.X + (size_t)(b0 + lr0 + i) * kIn + 8 * gq, g.W[0] + (size_t)i * kIn + 8 * gq, acc[rt]);
            else
                step_tile<kH>(&sH[l - 1][lr0 + i][8 * gq], g.W[l] + (size_t)i * kH + 8 * gq, acc[rt]);
        }
#pragma unroll
        for (uint32_t rt = 0; rt < kStepTiles; ++rt)
            step_store<true>(acc[rt], b0, (wave * kStepTiles + rt) * 16, B, g.P + bias[l], nullptr, g.H[l], g.HT[l],
                             sH[l]);
        __syncthreads();
    }
    // output layer, loss, dZ3 (one thread per sample)
    const float* W4 = g.P + hvx::kMlpW4;
    const float* b4 = g.P + hvx::kMlpB4;
    if (t < kOutBlock) {
        float h[kH];
#pragma unroll
        for (int k = 0; k < kH; ++k) h[k] = (float)sH[2][t][k];
        float y[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) y[c] = hvx::mlp_output(h, W4 + c * kH, b4[c]);
        const float4 tv = g.tgt[b0 + t];
        const float e0 = y[0] - tv.x, e1 = y[1] - tv.y, e2 = y[2] - tv.z;
        sdY[t][0] = e0 * g.inv_norm;
        sdY[t][1] = e1 * g.inv_norm;
        sdY[t][2] = e2 * g.inv_norm;
        sl[t] = (e0 * e0 + e1 * e1) + e2 * e2;
#pragma unroll
        for (int k8 = 0; k8 < kH / 8; ++k8) {
            bf16x8 pk;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int k = 8 * k8 + q;
                const float dh = (sdY[t][0] * W4[k] + sdY[t][1] * W4[kH + k]) + sdY[t][2] * W4[2 * kH + k];
                const bf16 v = (bf16)(h[k] > 0.0f ? dh : 0.0f);
                pk[q] = v;
                g.dZT[2][(size_t)k * B + b0 + t] = v;
                sZ[t][k] = v;
            }
            reinterpret_cast<bf16x8*>(g.dZ[2] + (size_t)(b0 + t) * kH)[k8] = pk;
        }
    }
    __syncthreads();
    float* pp = g.part + (size_t)blockIdx.x * kOutPart;
    static_assert(64 * kStepWaves >= (uint32_t)kOutPart - 1, "one thread per output-layer partial");
    if (t < 3 * kH) {
        const uint32_t c = t / kH, k = t % kH;
        float acc = 0.0f;
        for (uint32_t j = 0; j < kOutBlock; ++j) acc += sdY[j][c] * (float)sH[2][j][k];
        pp[t] = acc;
    } else if (t < 3 * kH + 3) {
        const uint32_t c = t - 3 * kH;
        float acc = 0.0f;
        for (uint32_t j = 0; j < kOutBlock; ++j) acc += sdY[j][c];
        pp[t] = acc;
    }
    for (uint32_t hh = kOutBlock / 2; hh > 0; hh >>= 1) {
        __syncthreads();
        if (t < hh) sl[t] += sl[t + hh];
    }
    if (t == 0) pp[kOutPart - 1] = sl[0];
    __syncthreads();  // H3 is read for the last time above
    // dZ2 = (dZ3 W3) [H2 > 0] into H3's rows, then dZ1 = (dZ2 W2) [H1 > 0]
    {
        f32x4 acc[kStepTiles][4];
#pragma unroll
        for (uint32_t rt = 0; rt < kStepTiles; ++rt)
            step_tile<kH>(&sZ[(wave * kStepTiles + rt) * 16 + i][8 * gq], g.W3T + (size_t)i * kH + 8 * gq, acc[rt]);
#pragma unroll
        for (uint32_t rt = 0; rt < kStepTiles; ++rt)
            step_store<false>(acc[rt], b0, (wave * kStepTiles + rt) * 16, B, nullptr, sH[1], g.dZ[1], g.dZT[1], sH[2]);
    }
    __syncthreads();
    {
        f32x4 acc[kStepTiles][4];
#pragma unroll
        for (uint32_t rt = 0; rt < kStepTiles; ++rt)
            step_tile<kH>(&sH[2][(wave * kStepTiles + rt) * 16 + i][8 * gq], g.W2T + (size_t)i * kH + 8 * gq,
                          acc[rt]);
#pragma unroll
        for (uint32_t rt = 0; rt < kStepTiles; ++rt)
            step_store<false>(acc[rt], b0, (wave * kStepTiles + rt) * 16, B, nullptr, sH[0], g.dZ[0], g.dZT[0],
                              nullptr);
    }
}

// out[i] = scale * sum_s part[s * stride + i] for many slices: a workgroup
// takes 64 outputs and 16 slice ranges (one per wave), each summed in slice
// order, then adds the 16 range sums in a fixed tree
constexpr uint32_t kRedGroups = 16;
// out[i] = scale * (sum over slices of part[s * stride + i]) for i < len,
// except that with `last` set element len - 1 goes to *last (times
// last_scale) instead
__device__ __forceinline__ void reduce_slices_block(const float* __restrict__ part, uint32_t slices, uint32_t stride,
                                                    uint32_t len, float scale, float* __restrict__ out,
                                                    uint32_t block, float* __restrict__ last = nullptr,
                                                    float last_scale = 1.0f) {
    __shared__ float sh[kRedGroups][64];
    const uint32_t l = threadIdx.x & 63, gq = threadIdx.x >> 6, i = block * 64 + l;
    float acc = 0.0f;
    if (i < len) {
        const uint32_t s0 = gq * slices / kRedGroups, s1 = (gq + 1) * slices / kRedGroups;
#pragma unroll 8
        for (uint32_t s = s0; s < s1; ++s) acc += part[(size_t)s * stride + i];
    }
    sh[gq][l] = acc;
    for (uint32_t h = kRedGroups / 2; h > 0; h >>= 1) {
        __syncthreads();
        if (gq < h) sh[gq][l] += sh[gq + h][l];
    }
    if (gq != 0 || i >= len) return;
    if (last && i == len - 1) *last = sh[0][l] * last_scale;
    else out[i] = sh[0][l] * scale;
}
__global__ __launch_bounds__(64 * kRedGroups) void reduce_slices_kernel(const float* __restrict__ part,
                                                                         uint32_t slices, uint32_t stride,
                                                                         uint32_t len, float scale,
                                                                         float* __restrict__ out) {
    reduce_slices_block(part, slices, stride, len, scale, out, blockIdx.x);
}
// A step's reductions in one launch: segment k (blocks [b[k], b[k + 1]))
// reduces len[k] outputs over slices[k] partials of stride[k] into out[k]
// (scale 1): the three layers' weight gradients and the output layer's block
// partials, whose last element (the loss, scaled by loss_scale) goes to
// *loss when it is set (otherwise it is not written).
struct ReduceStep {
    const float* part[4];
    float* out[4];
    uint32_t len[4], stride[4], slices[4], b[5];
    float* loss;
    float loss_scale;
};
__global__ __launch_bounds__(64 * kRedGroups) void reduce_step_kernel(ReduceStep r) {
    const uint32_t blk = blockIdx.x;
    uint32_t k = 0;
    while (k < 3 && blk >= r.b[k + 1]) ++k;
    const bool out_layer = k == 3;
    // without a loss pointer the output layer's last element is not written
    const uint32_t len = out_layer && !r.loss ? r.len[k] - 1 : r.len[k];
    reduce_slices_block(r.part[k], r.slices[k], r.stride[k], len, 1.0f, r.out[k], blk - r.b[k],
                        out_layer ? r.loss : nullptr, r.loss_scale);
}


// Adam step of every parameter; the updated weights also refresh their BF16
// working copies (as mlp_cast_kernel does), so a step needs no cast launch.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ P, const float* __restrict__ G,
                                                   float* __restrict__ m, float* __restrict__ v, uint32_t n, float lr,
                                                   float bc1, float bc2, bf16* __restrict__ W1, bf16* __restrict__ W2,
                                                   bf16* __restrict__ W3, bf16* __restrict__ W2T,
                                                   bf16* __restrict__ W3T) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float mi = 0.9f * m[i] + 0.1f * g, vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = P[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
    P[i] = p;
    const bf16 h = (bf16)p;
    if (i - (uint32_t)hvx::kMlpW1 < (uint32_t)(kH * kIn)) W1[i - hvx::kMlpW1] = h;
    if (i - (uint32_t)hvx::kMlpW2 < (uint32_t)(kH * kH)) {
        const uint32_t j = i - hvx::kMlpW2;
        W2[j] = h;
        W2T[(j % kH) * kH + j / kH] = h;
    }
    if (i - (uint32_t)hvx::kMlpW3 < (uint32_t)(kH * kH)) {
        const uint32_t j = i - hvx::kMlpW3;
        W3[j] = h;
        W3T[(j % kH) * kH + j / kH] = h;
    }
}

// BF16 working copies: W1, W2, W3 as [out][in]; W2^T, W3^T as [in][out]
__global__ __launch_bounds__(256) void mlp_cast_kernel(const float* __restrict__ P, bf16* __restrict__ W1,
                                                       bf16* __restrict__ W2, bf16* __restrict__ W3,
                                                       bf16* __restrict__ W2T, bf16* __restrict__ W3T) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < (uint32_t)(kH * kIn)) W1[i] = (bf16)P[hvx::kMlpW1 + i];
    if (i < (uint32_t)(kH * kH)) {
        const uint32_t o = i / kH, k = i % kH;
        W2[i] = (bf16)P[hvx::kMlpW2 + i];
        W3[i] = (bf16)P[hvx::kMlpW3 + i];
        W2T[k * kH + o] = (bf16)P[hvx::kMlpW2 + i];
        W3T[k * kH + o] = (bf16)P[hvx::kMlpW3 + i];
    }
}

// Inference through the in-shading path (mlp_device.h): each workgroup
// stages the weights once, its four waves evaluate 16 samples per call.
__global__ __launch_bounds__(256) void mlp_infer_kernel(MlpView v, const float4* __restrict__ in, uint32_t n,
                                                        float* __restrict__ rgb) {
    __shared__ MlpLds L;
    mlp_stage(L, v);
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    for (uint32_t base = blockIdx.x * 64; base < n; base += gridDim.x * 64) {
        const uint32_t sidx = base + wave * 16 + i;
        float x[4] = {0.0f, 0.0f, 1.0f, 0.0f};
        if (sidx < n) {
            const float4 q = in[sidx];
            x[0] = q.x;
            x[1] = q.y;
            x[2] = q.z;
            x[3] = q.w;
        }
        const float y = mlp_forward_tile(L, x, v.lod_max);
        if (g < 3 && sidx < n) rgb[(size_t)sidx * 3 + g] = y;
    }
}

// Fallback mask cells: (lod bin b, face f, texel y, x) of a mask_face cube.
// Points of a cell: a kMaskGrid x kMaskGrid grid over the texel (corners,
// edge and inner points) at the bin's two ends and middle.
constexpr uint32_t kMaskGrid = 5;
constexpr uint32_t kMaskPoints = kMaskGrid * kMaskGrid * 3;

__device__ __forceinline__ f3 plane_dir_f(int f, double u, double v) {
    double p[3];
    hvx::contract_face_point(f, u, v, p);
    return normalize3(f3{(float)p[0], (float)p[1], (float)p[2]});
}

__global__ __launch_bounds__(256) void mask_points_kernel(uint32_t mf, uint32_t bins, float lod_max,
                                                          float4* __restrict__ pts) {
    const uint32_t cells = bins * 6 * mf * mf;
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= cells * kMaskPoints) return;
    const uint32_t cell = id / kMaskPoints, p = id % kMaskPoints;
    const uint32_t x = cell % mf, y = (cell / mf) % mf, f = (cell / (mf * mf)) % 6, b = cell / (6 * mf * mf);
    const uint32_t gp = p % (kMaskGrid * kMaskGrid), li = p / (kMaskGrid * kMaskGrid);
    const double cx = x + (double)(gp % kMaskGrid) / (kMaskGrid - 1), cy = y + (double)(gp / kMaskGrid) / (kMaskGrid - 1);
    const f3 d = plane_dir_f((int)f, 2.0 * cx / mf - 1.0, 2.0 * cy / mf - 1.0);
    const float lod = lod_max * ((float)b + 0.5f * (float)li) / (float)bins;
    pts[id] = make_float4(d.x, d.y, d.z, lod);
}

// err[cell] = largest relative error (every channel, every point of the
// cell) of the network against the chain's textureLod value; infinity where
// the reference is 0 and the prediction is not. lo[cell] = per channel,
// kCellLowFactor times the smallest prediction over the cell's points: an
// estimate from below of what the network returns anywhere in the cell.
// Shading evaluates the network only where the cell could serve at that
// estimate and decides on the actual output, so the factor trades network
// evaluations against fragments served (3840x2160 frame, 15,000-step
// network: 1.0, 0.75 and 0.5 all serve 24.4 %, learned shading 2.34, 2.36
// and 2.41 ms).
#ifndef HVX_MLP_LO_FACTOR
#define HVX_MLP_LO_FACTOR 1.0f
#endif
constexpr float kCellLowFactor = HVX_MLP_LO_FACTOR;
__global__ __launch_bounds__(256) void cell_error_kernel(const float4* __restrict__ pts, const float* __restrict__ rgb,
                                                         const float4* __restrict__ chain, uint32_t S, uint32_t mips,
                                                         uint32_t cells, float* __restrict__ err,
                                                         float4* __restrict__ lo) {
    const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= cells) return;
    float e = 0.0f, mn[3] = {INFINITY, INFINITY, INFINITY};
    for (uint32_t p = 0; p < kMaskPoints; ++p) {
        const float4 q = pts[(size_t)cell * kMaskPoints + p];
        const f3 ref = chain_lod(chain, S, mips, f3{q.x, q.y, q.z}, q.w);
        const float* m = rgb + ((size_t)cell * kMaskPoints + p) * 3;
        const float r3[3] = {ref.x, ref.y, ref.z};
        for (int c = 0; c < 3; ++c) {
            const float d = fabsf(m[c] - r3[c]);
            float rel = d == 0.0f ? 0.0f : (r3[c] != 0.0f ? d / fabsf(r3[c]) : INFINITY);
            if (!(rel == rel)) rel = INFINITY;  // a NaN prediction fails every bound
            e = fmaxf(e, rel);
            mn[c] = fminf(mn[c], m[c]);
        }
    }
    err[cell] = e;
    float l3[3];
    for (int c = 0; c < 3; ++c) l3[c] = mn[c] > 0.0f ? kCellLowFactor * mn[c] : 0.0f;
    lo[cell] = make_float4(l3[0], l3[1], l3[2], 0.0f);
}

// The error the fallback decides on: the largest measured error of the cell
// times kCellSafety, a margin for the error between the test points, which
// is not measured.
constexpr float kCellSafety = 2.0f;
__global__ __launch_bounds__(256) void cell_margin_kernel(const float* __restrict__ raw, uint32_t cells,
                                                          float* __restrict__ err) {
    const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
    if (cell < cells) err[cell] = raw[cell] * kCellSafety;
}

}  // namespace

struct hvx_mlp {
    hvx_ctx* ctx = nullptr;
    float lod_max = 1.0f;
    float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;
    bf16 *W1 = nullptr, *W2 = nullptr, *W3 = nullptr, *W2T = nullptr, *W3T = nullptr;
    uint32_t cap = 0;  // batch capacity
    bf16 *X0 = nullptr, *X0T = nullptr, *H[3] = {}, *HT[3] = {}, *dZa = nullptr, *dZaT = nullptr, *dZb = nullptr,
         *dZbT = nullptr, *dZc = nullptr, *dZcT = nullptr;
    float *part = nullptr, *opart = nullptr;
    uint32_t splitk = 1;
    uint64_t steps_done = 0;
    // a training call's inputs encoded once: [n][kIn] and [kIn][enc_cap]
    bf16 *enc = nullptr, *encT = nullptr;
    size_t enc_cap = 0;
    // per-cell estimate from below of the network's output, from the last
    // hvx_mlp_cell_errors call (cell grid lo_face x lo_face x lo_bins)
    float4* cell_lo = nullptr;  // RGB and a pad per cell
    uint32_t lo_face = 0, lo_bins = 0;
    std::vector<void*> batch_allocs;
    std::vector<void*> allocs;
};

namespace hvx {
MlpView mlp_view(const hvx_mlp* M) { return MlpView{M->W1, M->W2, M->W3, M->P, M->lod_max}; }
const float4* mlp_cell_lower(const hvx_mlp* M, uint32_t cell_face, uint32_t lod_bins) {
    return M->cell_lo && M->lo_face == cell_face && M->lo_bins == lod_bins ? M->cell_lo : nullptr;
}
}  // namespace hvx

namespace {

int mfail(hvx_ctx* ctx, int st, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int mfail(hvx_ctx* ctx, int st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    return st;
}

template <class T> hipError_t dalloc(std::vector<void*>& list, T** p, size_t count) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) list.push_back(q);
    *p = static_cast<T*>(q);
    return e;
}

hipError_t ensure_batch(hvx_mlp* M, uint32_t B) {
    if (B <= M->cap) return hipSuccess;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    for (void* p : M->batch_allocs) (void)hipFree(p);
    M->batch_allocs.clear();
    M->cap = 0;
    auto& L = M->batch_allocs;
    // weight gradients are [64][<=128] GEMMs over K = B: split K into slices
    // of 128 samples (up to 256) so that the GEMM has 1-2 k waves
    M->splitk = std::max(1u, std::min(256u, B / 128));
    if ((e = dalloc(L, &M->X0, (size_t)B * kIn)) || (e = dalloc(L, &M->X0T, (size_t)B * kIn))) return e;
    for (int l = 0; l < 3; ++l)
        if ((e = dalloc(L, &M->H[l], (size_t)B * kH)) || (e = dalloc(L, &M->HT[l], (size_t)B * kH))) return e;
    if ((e = dalloc(L, &M->dZa, (size_t)B * kH)) || (e = dalloc(L, &M->dZaT, (size_t)B * kH)) ||
        (e = dalloc(L, &M->dZb, (size_t)B * kH)) || (e = dalloc(L, &M->dZbT, (size_t)B * kH)) ||
        (e = dalloc(L, &M->dZc, (size_t)B * kH)) || (e = dalloc(L, &M->dZcT, (size_t)B * kH)) ||
        (e = dalloc(L, &M->part, (size_t)M->splitk * ((kH * kIn + kH) + 2 * (kH * kH + kH)))) ||
        (e = dalloc(L, &M->opart, (size_t)(B / kOutBlock) * kOutPart)))
        return e;
    M->cap = B;
    return hipSuccess;
}

template <int EPI> void gemm(const GemmArgs& g, hipStream_t s) {
    const uint32_t waves = (g.M / 16) * (g.Nc / 64) * g.splitk;
    hipLaunchKernelGGL(gemm_abt_kernel<EPI>, dim3((waves + 3) / 4), dim3(256), 0, s, g);
}

// forward of B (padded) samples, n of them real; x0: their encoding, if
// already made
void forward(hvx_mlp* M, const float4* in, uint32_t n, uint32_t B, hipStream_t s, const bf16* x0 = nullptr) {
    if (!x0)
        hipLaunchKernelGGL(mlp_pe_kernel, dim3((B + 63) / 64), dim3(64), 0, s, in, n, B, M->lod_max, M->X0, M->X0T);
    const bf16* Ws[3] = {M->W1, M->W2, M->W3};
    const int bias[3] = {hvx::kMlpB1, hvx::kMlpB2, hvx::kMlpB3};
    const bf16* prev = x0 ? x0 : M->X0;
    for (int l = 0; l < 3; ++l) {
        GemmArgs g{};
        g.A = prev;
        g.Bt = Ws[l];
        g.M = B;
        g.Nc = kH;
        g.Kd = l == 0 ? kIn : kH;
        g.splitk = 1;
        g.bias = M->P + bias[l];
        g.C = M->H[l];
        g.CT = M->HT[l];
        gemm<EPI_FWD>(g, s);
        prev = M->H[l];
    }
}

// gradient of the mean squared log-radiance error over B samples into G (and
// the loss into *loss); x0 / x0t (row stride ldx): the samples' encoding, if
// already made
void gradient_fused(hvx_mlp* M, const float4* in, const float4* tgt, uint32_t B, float* G, float* loss,
                    hipStream_t s, const bf16* x0, const bf16* x0t, uint32_t ldx);

void gradient(hvx_mlp* M, const float4* in, const float4* tgt, uint32_t B, float* G, float* loss, hipStream_t s,
              const bf16* x0 = nullptr, const bf16* x0t = nullptr, uint32_t ldx = 0) {
    if (!M->ctx->mlp_separate) {
        gradient_fused(M, in, tgt, B, G, loss, s, x0, x0t, ldx);
        return;
    }
    forward(M, in, B, B, s, x0);
    const uint32_t nblk = B / kOutBlock;
    hipLaunchKernelGGL(mlp_out_kernel<true>, dim3(nblk), dim3(kOutBlock), 0, s, M->H[2], M->P, tgt, B, B,
                       2.0f / (3.0f * (float)B), nullptr, M->dZa, M->dZaT, M->opart);
    hipLaunchKernelGGL(reduce_out_kernel, dim3((kOutPart + 63) / 64), dim3(64 * kOutRedGroups), 0, s, M->opart, nblk, G + hvx::kMlpW4,
                       1.0f / (3.0f * (float)B), loss);
    const int wofs[3] = {hvx::kMlpW1, hvx::kMlpW2, hvx::kMlpW3};
    const bf16* WT[3] = {nullptr, M->W2T, M->W3T};
    bf16 *dz = M->dZa, *dzT = M->dZaT, *nz = M->dZb, *nzT = M->dZbT;
    for (int l = 2; l >= 0; --l) {
        const uint32_t K = l == 0 ? kIn : kH;
        GemmArgs g{};
        g.A = dzT;
        g.Bt = l == 0 ? (x0t ? x0t : M->X0T) : M->HT[l - 1];
        g.ldb = l == 0 && x0t ? ldx : 0;
        g.M = kH;
        g.Nc = K;
        g.Kd = B;
        g.splitk = M->splitk;
        g.part = M->part;
        g.pstride = kH * K + kH;
        g.colsum = 1;
        gemm<EPI_PART>(g, s);
        // W_l and b_l are adjacent in the parameter vector: one reduction
        static_assert(hvx::kMlpB1 == hvx::kMlpW1 + kH * kIn && hvx::kMlpB2 == hvx::kMlpW2 + kH * kH &&
                          hvx::kMlpB3 == hvx::kMlpW3 + kH * kH,
                      "each layer's bias follows its weights");
        hipLaunchKernelGGL(reduce_slices_kernel, dim3((kH * K + kH + 63) / 64), dim3(64 * kRedGroups), 0, s,
                           M->part, M->splitk, (uint32_t)(kH * K + kH), (uint32_t)(kH * K + kH), 1.0f, G + wofs[l]);
        if (l > 0) {
            GemmArgs d{};
            d.A = dz;
            d.Bt = WT[l];
            d.M = B;
            d.Nc = kH;
            d.Kd = kH;
            d.splitk = 1;
            d.mask = M->H[l - 1];
            d.C = nz;
            d.CT = nzT;
            gemm<EPI_BWD>(d, s);
            std::swap(dz, nz);
            std::swap(dzT, nzT);
        }
    }
}

// The same gradient with the per-sample work in one launch (mlp_step_kernel),
// then the three weight-gradient GEMMs and their reductions as in gradient().
void gradient_fused(hvx_mlp* M, const float4* in, const float4* tgt, uint32_t B, float* G, float* loss,
                    hipStream_t s, const bf16* x0, const bf16* x0t, uint32_t ldx) {
    if (!x0)
        hipLaunchKernelGGL(mlp_pe_kernel, dim3((B + 63) / 64), dim3(64), 0, s, in, B, B, M->lod_max, M->X0, M->X0T);
    const uint32_t nblk = B / kOutBlock;
    StepArgs a{};
    a.X = x0 ? x0 : M->X0;
    a.P = M->P;
    a.W[0] = M->W1;
    a.W[1] = M->W2;
    a.W[2] = M->W3;
    a.W2T = M->W2T;
    a.W3T = M->W3T;
    a.tgt = tgt;
    a.B = B;
    a.inv_norm = 2.0f / (3.0f * (float)B);
    for (int l = 0; l < 3; ++l) {
        a.H[l] = M->H[l];
        a.HT[l] = M->HT[l];
    }
    a.dZ[0] = M->dZc;
    a.dZT[0] = M->dZcT;
    a.dZ[1] = M->dZb;
    a.dZT[1] = M->dZbT;
    a.dZ[2] = M->dZa;
    a.dZT[2] = M->dZaT;
    a.part = M->opart;
    hipLaunchKernelGGL(mlp_step_kernel, dim3(nblk), dim3(64 * kStepWaves), 0, s, a);
    // the three weight-gradient GEMMs in one launch, each into its own
    // partials, then their reductions in one launch (the same per-wave and
    // per-output work as separate launches)
    const int wofs[3] = {hvx::kMlpW1, hvx::kMlpW2, hvx::kMlpW3};
    GemmArgs3 g3{};
    ReduceStep r3{};
    float* part = M->part;
    uint32_t waves = 0;
    for (int l = 0; l < 3; ++l) {
        const uint32_t K = l == 0 ? kIn : kH, len = kH * K + kH;
        GemmArgs& g = g3.g[l];
        g.A = a.dZT[l];
        g.Bt = l == 0 ? (x0t ? x0t : M->X0T) : M->HT[l - 1];
        g.ldb = l == 0 && x0t ? ldx : 0;
        g.M = kH;
        g.Nc = K;
        g.Kd = B;
        g.splitk = M->splitk;
        g.part = part;
        g.pstride = len;
        g.colsum = 1;
        r3.part[l] = part;
        r3.out[l] = G + wofs[l];
        r3.len[l] = r3.stride[l] = len;
        r3.slices[l] = M->splitk;
        r3.b[l + 1] = r3.b[l] + (len + 63) / 64;
        part += (size_t)M->splitk * len;
        const uint32_t w = (g.M / 16) * (g.Nc / 64) * g.splitk;
        if (l == 0) g3.w1 = w;
        else if (l == 1) g3.w2 = g3.w1 + w;
        waves += w;
    }
    // the output layer: dW4, db4 and the loss from the step kernel's block partials
    r3.part[3] = M->opart;
    r3.out[3] = G + hvx::kMlpW4;
    r3.len[3] = r3.stride[3] = kOutPart;
    r3.slices[3] = nblk;
    r3.b[4] = r3.b[3] + (kOutPart + 63) / 64;
    r3.loss = loss;
    r3.loss_scale = 1.0f / (3.0f * (float)B);
    hipLaunchKernelGGL(gemm_part3_kernel, dim3((waves + 3) / 4), dim3(256), 0, s, g3);
    hipLaunchKernelGGL(reduce_step_kernel, dim3(r3.b[4]), dim3(64 * kRedGroups), 0, s, r3);
}

void cast(hvx_mlp* M, hipStream_t s) {
    hipLaunchKernelGGL(mlp_cast_kernel, dim3(kH * kIn / 256), dim3(256), 0, s, M->P, M->W1, M->W2, M->W3, M->W2T,
                       M->W3T);
}

int bind(hvx_ctx* ctx) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess))
        return mfail(ctx, HVX_ERR_HIP, "cannot bind device %d", ctx->device);
    return HVX_OK;
}

}  // namespace

extern "C" {

int hvx_mlp_create(hvx_ctx* ctx, float lod_max, uint64_t seed, hvx_mlp** out) {
    if (!ctx || !out) return HVX_ERR_NULL;
    *out = nullptr;
    if (!(lod_max > 0.0f) || lod_max > 64.0f) return mfail(ctx, HVX_ERR_INVALID_ARG, "lod_max must be in (0, 64]");
    int st;
    if ((st = bind(ctx))) return st;
    hvx_mlp* M = new (std::nothrow) hvx_mlp();
    if (!M) return HVX_ERR_HIP;
    M->ctx = ctx;
    M->lod_max = lod_max;
    hipError_t e;
    auto& L = M->allocs;
    const size_t np = hvx::kMlpParams;
    if ((e = dalloc(L, &M->P, np)) || (e = dalloc(L, &M->G, np)) || (e = dalloc(L, &M->m, np)) ||
        (e = dalloc(L, &M->v, np)) || (e = dalloc(L, &M->W1, (size_t)kH * kIn)) ||
        (e = dalloc(L, &M->W2, (size_t)kH * kH)) || (e = dalloc(L, &M->W3, (size_t)kH * kH)) ||
        (e = dalloc(L, &M->W2T, (size_t)kH * kH)) || (e = dalloc(L, &M->W3T, (size_t)kH * kH))) {
        hvx_mlp_destroy(M);
        return mfail(ctx, HVX_ERR_HIP, "hvx_mlp_create: %s", hipGetErrorString(e));
    }
    // uniform He initialisation (bound sqrt(6 / fan_in)), zero biases, from a
    // SplitMix64 stream of the seed
    std::vector<float> h(np, 0.0f);
    uint64_t state = seed;
    auto next = [&]() {
        uint64_t x = (state += 0x9E3779B97F4A7C15ull);
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int ofs, int count, int fan_in) {
        const double bound = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; ++i) h[ofs + i] = (float)((2.0 * next() - 1.0) * bound);
    };
    fill(hvx::kMlpW1, kH * kIn, kIn);
    fill(hvx::kMlpW2, kH * kH, kH);
    fill(hvx::kMlpW3, kH * kH, kH);
    fill(hvx::kMlpW4, 3 * kH, kH);
    if ((e = hipMemcpy(M->P, h.data(), np * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(M->m, 0, np * 4)) != hipSuccess || (e = hipMemset(M->v, 0, np * 4)) != hipSuccess) {
        hvx_mlp_destroy(M);
        return mfail(ctx, HVX_ERR_HIP, "hvx_mlp_create: %s", hipGetErrorString(e));
    }
    cast(M, nullptr);
    if ((e = hipDeviceSynchronize()) != hipSuccess) {
        hvx_mlp_destroy(M);
        return mfail(ctx, HVX_ERR_HIP, "hvx_mlp_create: %s", hipGetErrorString(e));
    }
    *out = M;
    return HVX_OK;
}

int hvx_mlp_destroy(hvx_mlp* M) {
    if (!M) return HVX_ERR_NULL;
    (void)hipDeviceSynchronize();
    for (void* p : M->batch_allocs) (void)hipFree(p);
    for (void* p : M->allocs) (void)hipFree(p);
    if (M->enc) (void)hipFree(M->enc);
    if (M->encT) (void)hipFree(M->encT);
    if (M->cell_lo) (void)hipFree(M->cell_lo);
    delete M;
    return HVX_OK;
}

int hvx_mlp_get_params(hvx_mlp* M, float* host_params) {
    if (!M || !host_params) return HVX_ERR_NULL;
    const hipError_t e = hipMemcpy(host_params, M->P, hvx::kMlpParams * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? HVX_OK : mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_get_params: %s", hipGetErrorString(e));
}

int hvx_mlp_set_params(hvx_mlp* M, const float* host_params) {
    if (!M || !host_params) return HVX_ERR_NULL;
    hipError_t e = hipMemcpy(M->P, host_params, hvx::kMlpParams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        cast(M, nullptr);
        e = hipDeviceSynchronize();
    }
    return e == hipSuccess ? HVX_OK : mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_set_params: %s", hipGetErrorString(e));
}

int hvx_mlp_sample(hvx_mlp* M, const float* d_chain, uint32_t face, uint32_t mips, uint64_t seed, uint32_t n,
                   float* d_inputs, float* d_targets, void* stream) {
    if (!M || !d_chain || !d_inputs || !d_targets) return HVX_ERR_NULL;
    if (face == 0 || face > 16384 || mips == 0 || mips > hvx_mip_count(face) || n == 0)
        return mfail(M->ctx, HVX_ERR_INVALID_ARG, "hvx_mlp_sample: bad chain sizes or n");
    int st;
    if ((st = bind(M->ctx))) return st;
    hipLaunchKernelGGL(mlp_data_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const float4*>(d_chain), face, mips, M->lod_max, seed, n,
                       reinterpret_cast<float4*>(d_inputs), reinterpret_cast<float4*>(d_targets));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? HVX_OK : mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_sample: %s", hipGetErrorString(e));
}

int hvx_mlp_gradient(hvx_mlp* M, const float* d_inputs, const float* d_targets, uint32_t batch, float* d_grads,
                     float* d_loss, void* stream) {
    if (!M || !d_inputs || !d_targets || !d_grads) return HVX_ERR_NULL;
    if (batch == 0 || batch % 256 || batch > kMaxBatch)
        return mfail(M->ctx, HVX_ERR_INVALID_ARG, "batch must be a positive multiple of 256 (<= 2^20)");
    int st;
    if ((st = bind(M->ctx))) return st;
    hipError_t e = ensure_batch(M, batch);
    if (e != hipSuccess) return mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_gradient: %s", hipGetErrorString(e));
    gradient(M, reinterpret_cast<const float4*>(d_inputs), reinterpret_cast<const float4*>(d_targets), batch, d_grads,
             d_loss, (hipStream_t)stream);
    e = hipGetLastError();
    return e == hipSuccess ? HVX_OK : mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_gradient: %s", hipGetErrorString(e));
}

int hvx_mlp_train(hvx_mlp* M, const float* d_inputs, const float* d_targets, uint32_t n, uint32_t batch,
                  uint32_t steps, float lr, float* d_loss_curve, void* stream) {
    if (!M || !d_inputs || !d_targets) return HVX_ERR_NULL;
    if (batch == 0 || batch % 256 || batch > kMaxBatch || n < batch || !(lr > 0.0f))
        return mfail(M->ctx, HVX_ERR_INVALID_ARG, "hvx_mlp_train: batch a multiple of 256 <= n, lr > 0");
    int st;
    if ((st = bind(M->ctx))) return st;
    hipError_t e = ensure_batch(M, batch);
    if (e != hipSuccess) return mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_train: %s", hipGetErrorString(e));
    hipStream_t s = (hipStream_t)stream;
    const uint32_t per_epoch = n / batch;
    const float4* in = reinterpret_cast<const float4*>(d_inputs);
    const float4* tg = reinterpret_cast<const float4*>(d_targets);
    // Encode the call's samples once (2 x 256 B per sample) when they are
    // revisited, instead of once per step; without the memory, each step
    // encodes its batch.
    const size_t npad = ((size_t)per_epoch * batch + 255) / 256 * 256;
    if (steps > per_epoch && M->enc_cap < npad) {
        if (M->enc) (void)hipFree(M->enc);
        if (M->encT) (void)hipFree(M->encT);
        M->enc = M->encT = nullptr;
        M->enc_cap = 0;
        if (hipMalloc(reinterpret_cast<void**>(&M->enc), npad * kIn * sizeof(bf16)) == hipSuccess &&
            hipMalloc(reinterpret_cast<void**>(&M->encT), npad * kIn * sizeof(bf16)) == hipSuccess) {
            M->enc_cap = npad;
        } else {
            if (M->enc) (void)hipFree(M->enc);
            M->enc = nullptr;
            (void)hipGetLastError();  // a failed allocation leaves the slower path, not an error
        }
    }
    const bool pre = steps > per_epoch && M->enc_cap >= npad;
    if (pre)
        hipLaunchKernelGGL(mlp_pe_kernel, dim3((uint32_t)(npad / 64)), dim3(64), 0, s, in, per_epoch * batch,
                           (uint32_t)npad, M->lod_max, M->enc, M->encT);
    for (uint32_t k = 0; k < steps; ++k) {
        const size_t off = (size_t)(k % per_epoch) * batch;
        if (pre)
            gradient(M, in + off, tg + off, batch, M->G, d_loss_curve ? d_loss_curve + k : nullptr, s,
                     M->enc + off * kIn, M->encT + off, (uint32_t)npad);
        else
            gradient(M, in + off, tg + off, batch, M->G, d_loss_curve ? d_loss_curve + k : nullptr, s);
        // cosine decay from lr to lr / 100 over the call
        const double ph = steps > 1 ? (double)k / (double)(steps - 1) : 0.0;
        const float lr_k = (float)(lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(hvx::kMlpPi * ph))));
        ++M->steps_done;
        const float bc1 = (float)(1.0 - std::pow(0.9, (double)M->steps_done));
        const float bc2 = (float)(1.0 - std::pow(0.999, (double)M->steps_done));
        hipLaunchKernelGGL(adam_kernel, dim3((hvx::kMlpParams + 255) / 256), dim3(256), 0, s, M->P, M->G, M->m, M->v,
                           (uint32_t)hvx::kMlpParams, lr_k, bc1, bc2, M->W1, M->W2, M->W3, M->W2T, M->W3T);
    }
    e = hipGetLastError();
    return e == hipSuccess ? HVX_OK : mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_train: %s", hipGetErrorString(e));
}

// Workgroups of mlp_infer_kernel: each stages the weights (34 KB) once and
// loops over 64-sample steps, so a few per CU suffice.
static uint32_t infer_blocks(const hvx_mlp* M, uint32_t n) {
    const uint32_t cap = 4u * (uint32_t)std::max(M->ctx->num_cus, 1);
    return std::max(1u, std::min(cap, (n + 63) / 64));
}

int hvx_mlp_eval(hvx_mlp* M, const float* d_inputs, uint32_t n, float* d_rgb, void* stream) {
    if (!M || !d_inputs || !d_rgb) return HVX_ERR_NULL;
    if (n == 0) return HVX_OK;
    int st;
    if ((st = bind(M->ctx))) return st;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(infer_blocks(M, n)), dim3(256), 0, (hipStream_t)stream, hvx::mlp_view(M),
                       reinterpret_cast<const float4*>(d_inputs), n, d_rgb);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? HVX_OK : mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_eval: %s", hipGetErrorString(e));
}

int hvx_mlp_cell_errors(hvx_mlp* M, const float* d_chain, uint32_t face, uint32_t mips, uint32_t cell_face,
                        uint32_t lod_bins, float* d_err, void* stream) {
    if (!M || !d_chain || !d_err) return HVX_ERR_NULL;
    const uint32_t mask_face = cell_face;
    if (face == 0 || face > 16384 || mips == 0 || mips > hvx_mip_count(face) || mask_face == 0 || mask_face > 1024 ||
        lod_bins == 0 || lod_bins > 256)
        return mfail(M->ctx, HVX_ERR_INVALID_ARG, "hvx_mlp_cell_errors: sizes out of range");
    int st;
    if ((st = bind(M->ctx))) return st;
    const uint32_t cells = lod_bins * 6 * mask_face * mask_face;
    const size_t npts = (size_t)cells * kMaskPoints;
    if (npts > (1u << 28)) return mfail(M->ctx, HVX_ERR_INVALID_ARG, "hvx_mlp_cell_errors: too many cells");
    // set-up call: the point, prediction and raw-error buffers are allocated here
    float4* pts = nullptr;
    float* rgb = nullptr;
    float* raw = nullptr;
    hipError_t e;
    if ((e = hipMalloc(&pts, npts * sizeof(float4))) != hipSuccess ||
        (e = hipMalloc(&rgb, npts * 3 * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&raw, (size_t)cells * sizeof(float))) != hipSuccess) {
        if (pts) (void)hipFree(pts);
        if (rgb) (void)hipFree(rgb);
        return mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_cell_errors: %s", hipGetErrorString(e));
    }
    // the lower estimates travel with the network for hvx_cf_shade_learned
    if (M->lo_face * M->lo_face * 6 * M->lo_bins != cells || !M->cell_lo) {
        (void)hipDeviceSynchronize();  // a shading call may still read the old ones
        if (M->cell_lo) (void)hipFree(M->cell_lo);
        M->cell_lo = nullptr;
        M->lo_face = M->lo_bins = 0;
        if ((e = hipMalloc(&M->cell_lo, (size_t)cells * sizeof(float4))) != hipSuccess) {
            M->cell_lo = nullptr;
            (void)hipFree(pts);
            (void)hipFree(rgb);
            (void)hipFree(raw);
            return mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_cell_errors: %s", hipGetErrorString(e));
        }
    }
    M->lo_face = mask_face;
    M->lo_bins = lod_bins;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(mask_points_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, mask_face, lod_bins,
                       M->lod_max, pts);
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(infer_blocks(M, npts)), dim3(256), 0, s,
                       hvx::mlp_view(M), pts, (uint32_t)npts, rgb);
    hipLaunchKernelGGL(cell_error_kernel, dim3((cells + 255) / 256), dim3(256), 0, s, pts, rgb,
                       reinterpret_cast<const float4*>(d_chain), face, mips, cells, raw, M->cell_lo);
    hipLaunchKernelGGL(cell_margin_kernel, dim3((cells + 255) / 256), dim3(256), 0, s, raw, cells, d_err);
    e = hipStreamSynchronize(s);
    (void)hipFree(pts);
    (void)hipFree(rgb);
    (void)hipFree(raw);
    return e == hipSuccess ? HVX_OK : mfail(M->ctx, HVX_ERR_HIP, "hvx_mlp_cell_errors: %s", hipGetErrorString(e));
}

}  // extern "C"
