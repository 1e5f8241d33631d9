// cf_kernels.hip — clustered forward on gfx950: cluster AABBs, light culling
// and clustered Cook-Torrance + IBL shading.
//
// Reference kernels (HelloVulkan/Shaders/ClusteredForward):
//   AABBGenerator.comp:39-89   -> aabb_kernel (one lane per cluster)
//   LightCulling.comp:59-87    -> light_prep_kernel + cull_kernel + scan/scatter
// (shading lives in shade_kernels.hip)
//
// Culling. The reference runs one invocation per cluster that walks every
// light and keeps the first `cap` hits. Here a workgroup owns a 4x4x1 block
// of clusters. Lights stream through in steps of 1024 (four per lane); each
// lane tests its lights against the union AABB of the block (a conservative
// filter: the per-axis distance terms of the union box are never larger than
// a member box's, term by term and under rounding, so a light that fails the
// union test fails every member test). Survivors are compacted into an LDS
// candidate list in ascending light order with wavefront ballots and a
// workgroup prefix sum (one barrier per step, the counts double-buffered).
// Once enough candidates are staged, every wave tests its clusters against
// them, 64 candidates per step, and appends hits with a second ballot + mbcnt
// prefix, keeping the first `cap` in ascending order — exactly the
// reference's list. The per-test arithmetic is the reference's, in the same
// order, with contraction off, so lists match it bit for bit.
#include <hip/hip_runtime.h>

#include "hvx_device.h"
#include "hvx_internal.h"

namespace hvx {
namespace {

// ------------------------------------------------------------------ AABBs
__global__ __launch_bounds__(256) void aabb_kernel(hvx_cf_ubo u, hvx_aabb* __restrict__ out) {
    const uint32_t X = u.slice_count_x, Y = u.slice_count_y, Z = u.slice_count_z;
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= X * Y * Z) return;
    const uint32_t x = id % X, y = (id / X) % Y, z = id / (X * Y);
    const float npx = 1.0f / (float)X, npy = 1.0f / (float)Y;
    float vmax[4], vmin[4];
    matvec4(u.camera_inverse_projection, (((float)x + 1.0f) * npx) * 2.0f - 1.0f,
            (((float)y + 1.0f) * npy) * 2.0f - 1.0f, 0.0f, 1.0f, vmax);
    matvec4(u.camera_inverse_projection, ((float)x * npx) * 2.0f - 1.0f, ((float)y * npy) * 2.0f - 1.0f, 0.0f,
            1.0f, vmin);
    const f3 maxVS = {vmax[0] / vmax[3], vmax[1] / vmax[3], vmax[2] / vmax[3]};
    const f3 minVS = {vmin[0] / vmin[3], vmin[1] / vmin[3], vmin[2] / vmin[3]};
    const float nearv = u.camera_near, farv = u.camera_far;
    const float ratio = farv / nearv;
    const float tileNear = (-nearv) * powf_(ratio, (float)z / (float)Z);
    const float tileFar = (-nearv) * powf_(ratio, (float)(z + 1) / (float)Z);
    const float* mn = &minVS.x;
    const float* mx = &maxVS.x;
    hvx_aabb box;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float p0 = (mn[c] * tileNear) / minVS.z;
        const float p1 = (mn[c] * tileFar) / minVS.z;
        const float p2 = (mx[c] * tileNear) / maxVS.z;
        const float p3 = (mx[c] * tileFar) / maxVS.z;
        box.min_point[c] = gmin(gmin(p0, p1), gmin(p2, p3));
        box.max_point[c] = gmax(gmax(p0, p1), gmax(p2, p3));
    }
    box.min_point[3] = 0.0f;
    box.max_point[3] = 0.0f;
    out[id] = box;
}

// -------------------------------------------------- view-space light prep
// posView = cameraView * posWorld (LightCulling.comp:53), r^2 as the test
// computes it (:56). Evaluated once per light instead of once per test.
__global__ __launch_bounds__(256) void light_prep_kernel(hvx_cf_ubo u, const hvx_light* __restrict__ lights,
                                                         uint32_t n, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const hvx_light& l = lights[i];
    float o[4];
    matvec4(u.camera_view, l.position[0], l.position[1], l.position[2], l.position[3], o);
    out[i] = make_float4(o[0], o[1], o[2], l.radius * l.radius);
}

struct Box {
    float mn[3], mx[3];
};

// SquaredDistancePointAABB + TestSphereAABB (LightCulling.comp:26-57): terms
// are added in axis order; an absent term adds nothing.
__device__ __forceinline__ bool sphere_box(float4 p, const Box& b) {
    const float v[3] = {p.x, p.y, p.z};
    float t[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float d = 0.0f;
        if (v[a] < b.mn[a]) d = b.mn[a] - v[a];
        else if (v[a] > b.mx[a]) d = b.mx[a] - v[a];
        t[a] = d * d;
    }
    const float sq = (t[0] + t[1]) + t[2];
    return sq < p.w;
}

constexpr int kWaves = 4;                 // 256 lanes
#ifndef HVX_CULL_LPL
#define HVX_CULL_LPL 4
#endif
constexpr int kLPL = HVX_CULL_LPL;        // lights per lane per level-1 step
constexpr int kTile = 256 * kLPL;         // lights per level-1 step
constexpr int kFlush = 256;               // staged candidates that trigger level 2
constexpr int kCandMax = kFlush + kTile;

__device__ __forceinline__ uint32_t lane_prefix(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <int GX, int GY, int GZ> struct CullShape {
    static constexpr int kG = GX * GY * GZ;
    static constexpr int kPerWave = kG / kWaves;
    static_assert(kG % kWaves == 0, "clusters per workgroup must split evenly over 4 waves");
};

template <int GX, int GY, int GZ> __host__ __device__ inline uint32_t cull_groups(uint32_t X, uint32_t Y, uint32_t Z) {
    return ((X + GX - 1) / GX) * ((Y + GY - 1) / GY) * ((Z + GZ - 1) / GZ);
}

struct CullShared {
    float4 pos[kCandMax];
    uint32_t idx[kCandMax];
    uint32_t wcnt[2][kLPL][kWaves];  // double-buffered by step parity
    uint32_t wdone[kWaves];
};

// Culls the GX x GY x GZ block of clusters `g` of job J (see the file header).
template <int GX, int GY, int GZ>
__device__ __forceinline__ void cull_group(const CullJob& J, uint32_t g, CullShared& sh) {
    using Shape = CullShape<GX, GY, GZ>;
    constexpr int kG = Shape::kG, kPerWave = Shape::kPerWave;
    const uint32_t X = J.X, Y = J.Y, Z = J.Z, n = J.n, cap = J.cap;
    const uint32_t ngx = (X + GX - 1) / GX, ngy = (Y + GY - 1) / GY;
    const uint32_t gx = g % ngx, gy = (g / ngx) % ngy, gz = g / (ngx * ngy);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __syncthreads();  // LDS reuse across consecutive groups of a persistent workgroup

    // union box of the valid clusters of this block (uniform)
    Box U;
    for (int a = 0; a < 3; ++a) { U.mn[a] = hvx::kInfF; U.mx[a] = -hvx::kInfF; }
    for (int k = 0; k < kG; ++k) {
        const uint32_t cx = gx * GX + (k % GX), cy = gy * GY + (k / GX) % GY, cz = gz * GZ + k / (GX * GY);
        if (cx >= X || cy >= Y || cz >= Z) continue;
        const hvx_aabb& b = J.clusters[cx + cy * X + cz * X * Y];
        for (int a = 0; a < 3; ++a) {
            U.mn[a] = gmin(U.mn[a], b.min_point[a]);
            U.mx[a] = gmax(U.mx[a], b.max_point[a]);
        }
    }

    // this wave's clusters
    Box cb[kPerWave];
    uint32_t cid[kPerWave], cnt[kPerWave], ovf[kPerWave];
    bool done[kPerWave];
#pragma unroll
    for (int j = 0; j < kPerWave; ++j) {
        const int k = (int)wave * kPerWave + j;
        const uint32_t cx = gx * GX + (k % GX), cy = gy * GY + (k / GX) % GY, cz = gz * GZ + k / (GX * GY);
        const bool valid = cx < X && cy < Y && cz < Z;
        cid[j] = valid ? cx + cy * X + cz * X * Y : 0u;
        cnt[j] = 0;
        ovf[j] = 0;
        done[j] = !valid;
        if (valid) {
            const hvx_aabb& b = J.clusters[cid[j]];
            for (int a = 0; a < 3; ++a) { cb[j].mn[a] = b.min_point[a]; cb[j].mx[a] = b.max_point[a]; }
        } else {
            for (int a = 0; a < 3; ++a) { cb[j].mn[a] = 0.0f; cb[j].mx[a] = 0.0f; }
        }
    }

    // union box of this wave's valid clusters (a wave with none never tests)
    Box UW;
    for (int a = 0; a < 3; ++a) { UW.mn[a] = hvx::kInfF; UW.mx[a] = -hvx::kInfF; }
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
        if (!done[j])
            for (int a = 0; a < 3; ++a) {
                UW.mn[a] = gmin(UW.mn[a], cb[j].mn[a]);
                UW.mx[a] = gmax(UW.mx[a], cb[j].mx[a]);
            }

    uint32_t base = 0, ncand = 0, ph = 0;
    // The next step's lights are loaded while this step is filtered. Loads
    // past the last light re-read it (the light table always holds at least
    // one entry), which keeps them unconditional; the lanes' hits are masked.
    const uint32_t last = n > 0 ? n - 1 : 0u;
    float4 pn[kLPL];
#pragma unroll
    for (int j = 0; j < kLPL; ++j) pn[j] = J.lvs[min(256u * j + tid, last)];
    for (;;) {
        // ---- level 1: kLPL lights per lane against the union box; sub-step
        // j covers lights base + 256 j .. base + 256 j + 255, so candidates
        // are compacted in (sub-step, wave, lane) order = ascending light order
        float4 p[kLPL];
        bool hit[kLPL];
        uint64_t m[kLPL];
#pragma unroll
        for (int j = 0; j < kLPL; ++j) {
            p[j] = pn[j];
            pn[j] = J.lvs[min(base + kTile + 256u * j + tid, last)];
        }
        __asm__ volatile("" ::: "memory");  // keeps the prefetch in this step
#pragma unroll
        for (int j = 0; j < kLPL; ++j) {
            const uint32_t li = base + 256u * j + tid;
            hit[j] = li < n && sphere_box(p[j], U);
        }
#pragma unroll
        for (int j = 0; j < kLPL; ++j) {
            m[j] = __ballot(hit[j]);
            if (lane == 0) sh.wcnt[ph][j][wave] = (uint32_t)__popcll(m[j]);
        }
        __syncthreads();
        uint32_t off[kLPL], tot = 0;
#pragma unroll
        for (int j = 0; j < kLPL; ++j) {
            off[j] = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = sh.wcnt[ph][j][w];
                if (w == (int)wave) off[j] = tot;
                tot += c;
            }
        }
#pragma unroll
        for (int j = 0; j < kLPL; ++j) {
            if (hit[j]) {
                const uint32_t pos = ncand + off[j] + lane_prefix(m[j]);
                sh.pos[pos] = p[j];
                sh.idx[pos] = base + 256u * j + tid;
            }
        }
        // the next step's counts go to the other buffer, so only level 2
        // (which reads every wave's candidates) needs a barrier here
        ph ^= 1u;
        ncand += tot;
        base += kTile;
        if (ncand < (uint32_t)kFlush && base < n) continue;
        __syncthreads();

        // ---- level 2: each wave tests its clusters against the staged
        // candidates, 64 per step: a lane reads its candidate once, skips
        // the step if no candidate meets the wave's union box, and otherwise
        // tests it against each cluster the wave still fills
        bool wave_done = true;
#pragma unroll
        for (int j = 0; j < kPerWave; ++j) wave_done = wave_done && done[j];
        for (uint32_t k0 = 0; k0 < ncand && !wave_done; k0 += 64) {
            const uint32_t k = k0 + lane;
            const bool valid = k < ncand;
            const float4 pk = valid ? sh.pos[k] : make_float4(0.f, 0.f, 0.f, 0.f);
            if (__ballot(valid && sphere_box(pk, UW)) == 0) continue;
            const uint32_t ik = valid ? sh.idx[k] : 0u;
            wave_done = true;
#pragma unroll
            for (int j = 0; j < kPerWave; ++j) {
                if (!done[j]) {
                    const bool h = valid && sphere_box(pk, cb[j]);
                    const uint64_t mm = __ballot(h);
                    if (cnt[j] < cap) {
                        const uint32_t pos = cnt[j] + lane_prefix(mm);
                        if (h && pos < cap) J.lists[(size_t)cid[j] * cap + pos] = ik;
                        const uint32_t nh = (uint32_t)__popcll(mm);
                        if (cnt[j] + nh > cap) ovf[j] = 1;
                        cnt[j] = min(cnt[j] + nh, cap);
                    } else if (mm != 0) {
                        ovf[j] = 1;
                    }
                    if (cnt[j] >= cap && (ovf[j] || !J.want_overflow)) done[j] = true;
                }
                wave_done = wave_done && done[j];
            }
        }
        if (lane == 0) sh.wdone[wave] = wave_done ? 1u : 0u;
        __syncthreads();
        ncand = 0;
        bool all_done = true;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) all_done = all_done && (sh.wdone[w] != 0);
        if (all_done || base >= n) break;
    }

    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < kPerWave; ++j) {
            const int k = (int)wave * kPerWave + j;
            const uint32_t cx = gx * GX + (k % GX), cy = gy * GY + (k / GX) % GY, cz = gz * GZ + k / (GX * GY);
            if (cx < X && cy < Y && cz < Z) {
                J.counts[cid[j]] = cnt[j];
                J.over[cid[j]] = ovf[j];
            }
        }
    }
}

template <int GX, int GY, int GZ> __global__ __launch_bounds__(256) void cull_kernel(CullJob J) {
    __shared__ CullShared sh;
    cull_group<GX, GY, GZ>(J, blockIdx.x, sh);
}

// Exclusive scan of the counts in cluster order -> LightCell offsets (the
// deterministic replacement of atomicAdd(globalIndex, count)), in three
// phases over blocks of 1024 clusters: block sums, a scan of the block sums,
// then each block's local scan plus its block offset.
constexpr uint32_t kScanBlock = 1024;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(v, d, 64);
        if ((int)lane >= d) v += u;
    }
    return v;
}

// 256-lane block exclusive scan; returns the block total in *total
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* s_w, uint32_t* total) {
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        woff += (k < (int)w) ? s_w[k] : 0u;
        tot += s_w[k];
    }
    __syncthreads();
    *total = tot;
    return woff + inc - v;
}

__device__ __forceinline__ void scan_partials_block(const uint32_t* __restrict__ counts,
                                                    const uint32_t* __restrict__ over, uint32_t C,
                                                    uint2* __restrict__ part, uint32_t blk, uint32_t* s_w,
                                                    uint32_t* s_v) {
    const uint32_t b0 = blk * kScanBlock;
    uint32_t sum = 0, ov = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t c = b0 + r * 256 + threadIdx.x;
        if (c < C) {
            sum += counts[c];
            ov += over[c];
        }
    }
    uint32_t tsum, tov;
    block_excl_scan256(sum, s_w, &tsum);
    block_excl_scan256(ov, s_v, &tov);
    if (threadIdx.x == 0) part[blk] = make_uint2(tsum, tov);
}

__global__ __launch_bounds__(256) void scan_partials_kernel(const uint32_t* __restrict__ counts,
                                                            const uint32_t* __restrict__ over, uint32_t C,
                                                            uint2* __restrict__ part) {
    __shared__ uint32_t s_w[4], s_v[4];
    scan_partials_block(counts, over, C, part, blockIdx.x, s_w, s_v);
}

// exclusive block offsets over nb partials with a workgroup of 1024 lanes
__device__ __forceinline__ void scan_top_block(uint2* __restrict__ part, uint32_t nb, uint32_t* __restrict__ total,
                                               uint32_t* __restrict__ novf, uint32_t* s_w, uint32_t* s_o) {
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6, nw = blockDim.x >> 6;
    uint32_t carry = 0, ocarry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += blockDim.x) {
        const uint32_t b = b0 + t;
        const uint2 p = b < nb ? part[b] : make_uint2(0u, 0u);
        const uint32_t inc = wave_incl_scan(p.x, lane);
        const uint32_t oinc = wave_incl_scan(p.y, lane);
        if (lane == 63) {
            s_w[w] = inc;
            s_o[w] = oinc;
        }
        __syncthreads();
        uint32_t woff = 0, tot = 0, otot = 0;
        for (uint32_t k = 0; k < nw; ++k) {
            woff += k < w ? s_w[k] : 0u;
            tot += s_w[k];
            otot += s_o[k];
        }
        __syncthreads();
        if (b < nb) part[b].x = carry + woff + inc - p.x;  // exclusive offset of block b
        carry += tot;
        ocarry += otot;
    }
    if (t == 0) {
        if (total) *total = carry;
        if (novf) *novf = ocarry;
    }
}

__global__ __launch_bounds__(1024) void scan_top_kernel(uint2* __restrict__ part, uint32_t nb,
                                                        uint32_t* __restrict__ total, uint32_t* __restrict__ novf) {
    __shared__ uint32_t s_w[16], s_o[16];
    scan_top_block(part, nb, total, novf, s_w, s_o);
}

__device__ __forceinline__ void scan_apply_block(const uint32_t* __restrict__ counts, uint32_t C,
                                                 const uint2* __restrict__ part, hvx_light_cell* __restrict__ cells,
                                                 uint32_t blk, uint32_t* s_w) {
    const uint32_t b0 = blk * kScanBlock;
    // thread t owns clusters b0 + 4t .. b0 + 4t + 3 (contiguous, in order)
    uint32_t k[4], local = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t c = b0 + threadIdx.x * 4 + r;
        k[r] = c < C ? counts[c] : 0u;
        local += k[r];
    }
    uint32_t tot;
    uint32_t run = part[blk].x + block_excl_scan256(local, s_w, &tot);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t c = b0 + threadIdx.x * 4 + r;
        if (c < C) {
            hvx_light_cell cell;
            cell.offset = run;
            cell.count = k[r];
            cells[c] = cell;
        }
        run += k[r];
    }
}

__global__ __launch_bounds__(256) void scan_apply_kernel(const uint32_t* __restrict__ counts, uint32_t C,
                                                         const uint2* __restrict__ part,
                                                         hvx_light_cell* __restrict__ cells) {
    __shared__ uint32_t s_w[4];
    scan_apply_block(counts, C, part, cells, blockIdx.x, s_w);
}

__device__ __forceinline__ void scatter_group(const uint32_t* __restrict__ lists, uint32_t cap, uint32_t C,
                                              const hvx_light_cell* __restrict__ cells, uint32_t* __restrict__ indices,
                                              uint32_t grp) {
    const uint32_t c = grp * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= C) return;
    const hvx_light_cell cell = cells[c];
    for (uint32_t j = lane; j < cell.count; j += 64) indices[cell.offset + j] = lists[(size_t)c * cap + j];
}

__global__ __launch_bounds__(256) void scatter_kernel(const uint32_t* __restrict__ lists, uint32_t cap, uint32_t C,
                                                      const hvx_light_cell* __restrict__ cells,
                                                      uint32_t* __restrict__ indices) {
    scatter_group(lists, cap, C, cells, indices, blockIdx.x);
}

// ------------------------------------- LightCullingBatch.comp semantics
// LightCullingBatch.comp:61-106: one lane per cluster (cluster = local index
// + kCullLdsBatch * workgroup), the lights staged in LDS in batches of
// kCullLdsBatch, every lane testing its cluster against the whole batch in
// slot order. The LDS entries hold the view-space centre and r^2 of each
// light (light_prep_kernel), the values the shader recomputes per test with
// the same arithmetic. Pass 0 counts, pass 1 writes at the scanned offsets;
// both make the same tests in the same order.
struct CullLdsJob {
    const hvx_aabb* clusters;
    const float4* lvs;
    uint32_t C, n, cap, flags;
    uint32_t* counts;
    uint32_t* over;
    const hvx_light_cell* cells;
    uint32_t* indices;
    size_t capacity;
};

template <bool WRITE>
__global__ __launch_bounds__(kCullLdsBatch) void cull_lds_batch_kernel(CullLdsJob J) {
    constexpr uint32_t T = kCullLdsBatch;
    __shared__ float4 sl[T];
    const uint32_t t = threadIdx.x, c = blockIdx.x * T + t;
    const bool valid = c < J.C;
    const bool ref_cap = (J.flags & HVX_CULL_BATCH_REFERENCE_CAP) != 0u;
    const bool ref_tail = (J.flags & HVX_CULL_BATCH_REFERENCE_TAIL) != 0u;
    Box b;
    if (valid) {
        const hvx_aabb& a = J.clusters[c];
        for (int k = 0; k < 3; ++k) { b.mn[k] = a.min_point[k]; b.mx[k] = a.max_point[k]; }
    } else {
        for (int k = 0; k < 3; ++k) { b.mn[k] = 0.0f; b.mx[k] = 0.0f; }
    }
    // slots the shader never writes are undefined there; zeros here (radius 0
    // never passes the strict test)
    sl[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t base = WRITE && valid ? J.cells[c].offset : 0u;
    uint32_t cnt = 0, ovf = 0;
    bool done = !valid;
    const uint32_t nb = (J.n + T - 1) / T;
    for (uint32_t batch = 0; batch < nb; ++batch) {
        // every lane has finished the previous batch before its slots are
        // overwritten (the shader has no barrier here, :86-93); a workgroup
        // whose clusters are all full stops
        if (__syncthreads_and(done ? 1 : 0)) break;
        const uint32_t li = batch * T + t;
        if (li < J.n) sl[t] = J.lvs[li];  // :79-82: the tail keeps the previous batch
        __syncthreads();
        const uint32_t m = (ref_tail || batch + 1 < nb) ? T : J.n - batch * T;  // slots tested
        if (done) continue;
        for (uint32_t i = 0; i < m; ++i) {
            if (!sphere_box(sl[i], b)) continue;
            if (!ref_cap && cnt >= J.cap) {  // a hit past the cap: the list is full
                ovf = 1;
                done = true;
                break;
            }
            if (WRITE) {
                const size_t pos = (size_t)base + cnt;
                if (pos < J.capacity) J.indices[pos] = batch * T + i;
            }
            ++cnt;
        }
    }
    if (!WRITE && valid) {
        J.counts[c] = cnt;
        J.over[c] = ovf;
    }
}

// ------------------------------------------------------ batched (multi-view)
// Every kernel below covers all views of a batch in one launch; a work item
// finds its view by the views' running item offsets (at most kMaxViews).
__device__ __forceinline__ uint32_t view_of(const CullBatchView* __restrict__ V, uint32_t nv, uint32_t item,
                                            uint32_t CullBatchView::*begin) {
    uint32_t v = 0;
    while (v + 1 < nv && item >= V[v + 1].*begin) ++v;
    return v;
}

__global__ __launch_bounds__(64) void batch_reset_kernel(uint32_t* __restrict__ q, uint32_t n) {
    if (threadIdx.x < n) q[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void batch_light_prep_kernel(const CullBatchView* __restrict__ V, uint32_t nv,
                                                               uint32_t total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const CullBatchView& v = V[view_of(V, nv, i, &CullBatchView::light_begin)];
    const uint32_t li = i - v.light_begin;
    const hvx_light& l = v.lights[li];
    float o[4];
    matvec4(v.view, l.position[0], l.position[1], l.position[2], l.position[3], o);
    const_cast<float4*>(v.job.lvs)[li] = make_float4(o[0], o[1], o[2], l.radius * l.radius);
}

// persistent: workgroups pull (view, cluster block) items from a queue
template <int GX, int GY, int GZ>
__global__ __launch_bounds__(256) void cull_batch_kernel(const CullBatchView* __restrict__ V, uint32_t nv,
                                                         uint32_t total_groups, uint32_t* __restrict__ queue) {
    __shared__ CullShared sh;
    __shared__ uint32_t s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= total_groups) break;
        const CullBatchView& v = V[view_of(V, nv, item, &CullBatchView::group_begin)];
        cull_group<GX, GY, GZ>(v.job, item - v.group_begin, sh);
    }
}

__global__ __launch_bounds__(256) void batch_scan_partials_kernel(const CullBatchView* __restrict__ V, uint32_t nv) {
    __shared__ uint32_t s_w[4], s_v[4];
    const CullBatchView& v = V[view_of(V, nv, blockIdx.x, &CullBatchView::block_begin)];
    scan_partials_block(v.job.counts, v.job.over, v.C, v.part, blockIdx.x - v.block_begin, s_w, s_v);
}

__global__ __launch_bounds__(1024) void batch_scan_top_kernel(const CullBatchView* __restrict__ V) {
    __shared__ uint32_t s_w[16], s_o[16];
    const CullBatchView& v = V[blockIdx.x];
    scan_top_block(v.part, v.nb, v.total, v.novf, s_w, s_o);
}

__global__ __launch_bounds__(256) void batch_scan_apply_kernel(const CullBatchView* __restrict__ V, uint32_t nv) {
    __shared__ uint32_t s_w[4];
    const CullBatchView& v = V[view_of(V, nv, blockIdx.x, &CullBatchView::block_begin)];
    scan_apply_block(v.job.counts, v.C, v.part, v.cells, blockIdx.x - v.block_begin, s_w);
}

__global__ __launch_bounds__(256) void batch_scatter_kernel(const CullBatchView* __restrict__ V, uint32_t nv) {
    const CullBatchView& v = V[view_of(V, nv, blockIdx.x, &CullBatchView::scatter_begin)];
    scatter_group(v.job.lists, v.job.cap, v.C, v.cells, v.indices, blockIdx.x - v.scatter_begin);
}

}  // namespace

// ================================================================ launchers
// cluster block of the single-view culling kernel (3840x2160, 32x18x48
// clusters, 65,536 lights: 4x4x1 0.186 ms,
// 8x4x1 0.191, 4x4x2 0.245, 2x2x4 0.245, 8x8x1 0.307)
#ifndef HVX_CULL_GX
#define HVX_CULL_GX 4
#define HVX_CULL_GY 4
#define HVX_CULL_GZ 1
#endif
constexpr int kCullGX = HVX_CULL_GX, kCullGY = HVX_CULL_GY, kCullGZ = HVX_CULL_GZ;
hipError_t launch_build_aabbs(const hvx_cf_ubo& ubo, hvx_aabb* d_clusters, hipStream_t s) {
    const uint32_t C = ubo.slice_count_x * ubo.slice_count_y * ubo.slice_count_z;
    hipLaunchKernelGGL(aabb_kernel, dim3((C + 255) / 256), dim3(256), 0, s, ubo, d_clusters);
    return hipGetLastError();
}

hipError_t launch_cull(const hvx_cf_ubo& ubo, const hvx_aabb* d_clusters, const hvx_light* d_lights,
                       uint32_t n_lights, uint32_t cap, hvx_light_cell* d_cells, uint32_t* d_indices,
                       uint32_t* d_total, uint32_t* d_overflow, float4* light_vs, uint32_t* lists, uint32_t* counts,
                       uint32_t* over, hipStream_t s) {
    const uint32_t X = ubo.slice_count_x, Y = ubo.slice_count_y, Z = ubo.slice_count_z;
    const uint32_t C = X * Y * Z;
    if (n_lights > 0) {
        hipLaunchKernelGGL(light_prep_kernel, dim3((n_lights + 255) / 256), dim3(256), 0, s, ubo, d_lights, n_lights,
                           light_vs);
    }
    CullJob J;
    J.X = X;
    J.Y = Y;
    J.Z = Z;
    J.n = n_lights;
    J.cap = cap;
    J.want_overflow = d_overflow != nullptr ? 1u : 0u;
    J.clusters = d_clusters;
    J.lvs = light_vs;
    J.lists = lists;
    J.counts = counts;
    J.over = over;
    hipLaunchKernelGGL((cull_kernel<kCullGX, kCullGY, kCullGZ>), dim3(cull_groups<kCullGX, kCullGY, kCullGZ>(X, Y, Z)),
                       dim3(256), 0, s, J);
    const uint32_t nb = (C + kScanBlock - 1) / kScanBlock;
    uint2* part = reinterpret_cast<uint2*>(over + C);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(nb), dim3(256), 0, s, counts, over, C, part);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(1024), 0, s, part, nb, d_total, d_overflow);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(256), 0, s, counts, C, part, d_cells);
    hipLaunchKernelGGL(scatter_kernel, dim3((C + 3) / 4), dim3(256), 0, s, lists, cap, C, d_cells, d_indices);
    return hipGetLastError();
}

hipError_t launch_cull_lds_batch(const hvx_cf_ubo& ubo, const hvx_aabb* d_clusters, const hvx_light* d_lights,
                                 uint32_t n_lights, uint32_t cap, uint32_t flags, hvx_light_cell* d_cells,
                                 uint32_t* d_indices, size_t indices_capacity, uint32_t* d_total, float4* light_vs,
                                 uint32_t* counts, uint32_t* over, hipStream_t s) {
    const uint32_t C = ubo.slice_count_x * ubo.slice_count_y * ubo.slice_count_z;
    if (n_lights > 0)
        hipLaunchKernelGGL(light_prep_kernel, dim3((n_lights + 255) / 256), dim3(256), 0, s, ubo, d_lights, n_lights,
                           light_vs);
    CullLdsJob J{d_clusters, light_vs, C, n_lights, cap, flags, counts, over, d_cells, d_indices, indices_capacity};
    const dim3 grid((C + kCullLdsBatch - 1) / kCullLdsBatch), block(kCullLdsBatch);
    hipLaunchKernelGGL(cull_lds_batch_kernel<false>, grid, block, 0, s, J);
    const uint32_t nb = (C + kScanBlock - 1) / kScanBlock;
    uint2* part = reinterpret_cast<uint2*>(over + C);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(nb), dim3(256), 0, s, counts, over, C, part);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(1024), 0, s, part, nb, d_total, nullptr);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(256), 0, s, counts, C, part, d_cells);
    hipLaunchKernelGGL(cull_lds_batch_kernel<true>, grid, block, 0, s, J);
    return hipGetLastError();
}

// ------------------------------------------------------ batched launchers
uint32_t cull_batch_groups(int shape, uint32_t X, uint32_t Y, uint32_t Z) {
    switch (shape) {
        case 0: return cull_groups<4, 4, 1>(X, Y, Z);
        case 1: return cull_groups<8, 4, 1>(X, Y, Z);
        case 2: return cull_groups<4, 4, 2>(X, Y, Z);
        default: return cull_groups<2, 2, 4>(X, Y, Z);
    }
}

hipError_t launch_cull_batch(const CullBatchView* d_views, uint32_t nv, const CullBatchTotals& t, int shape,
                             uint32_t persistent_wgs, uint32_t* d_queue, hipStream_t s) {
    if (t.lights > 0)
        hipLaunchKernelGGL(batch_light_prep_kernel, dim3((t.lights + 255) / 256), dim3(256), 0, s, d_views, nv,
                           t.lights);
    const dim3 grid(persistent_wgs), block(256);
    switch (shape) {
        case 0: hipLaunchKernelGGL((cull_batch_kernel<4, 4, 1>), grid, block, 0, s, d_views, nv, t.groups, d_queue); break;
        case 1: hipLaunchKernelGGL((cull_batch_kernel<8, 4, 1>), grid, block, 0, s, d_views, nv, t.groups, d_queue); break;
        case 2: hipLaunchKernelGGL((cull_batch_kernel<4, 4, 2>), grid, block, 0, s, d_views, nv, t.groups, d_queue); break;
        default: hipLaunchKernelGGL((cull_batch_kernel<2, 2, 4>), grid, block, 0, s, d_views, nv, t.groups, d_queue); break;
    }
    hipLaunchKernelGGL(batch_scan_partials_kernel, dim3(t.blocks), dim3(256), 0, s, d_views, nv);
    hipLaunchKernelGGL(batch_scan_top_kernel, dim3(nv), dim3(1024), 0, s, d_views);
    hipLaunchKernelGGL(batch_scan_apply_kernel, dim3(t.blocks), dim3(256), 0, s, d_views, nv);
    hipLaunchKernelGGL(batch_scatter_kernel, dim3(t.scatter), dim3(256), 0, s, d_views, nv);
    return hipGetLastError();
}

hipError_t launch_queue_reset(uint32_t* d_queues, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, d_queues, n);
    return hipGetLastError();
}

}  // namespace hvx
