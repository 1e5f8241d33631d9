// hvx_api.hip — the C ABI of include/hvx.h: argument validation, context and
// scratch management, and dispatch to the gfx950 kernels.
//
// Error behaviour: every entry point returns an hvx_status and records a
// message on the context (hvx_ctx_last_error). Nothing is launched when a
// check fails. The reference logs Vulkan errors and carries on
// (HelloVulkan/Source/Vulkan/VulkanCheck.cpp:3-8); a drop-in that must not
// corrupt memory instead refuses the call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "hvx_device.h"
#include "hvx_internal.h"

namespace {

int fail(hvx_ctx* ctx, int status, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(hvx_ctx* ctx, int status, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return status;
}

int hip_fail(hvx_ctx* ctx, hipError_t e, const char* where) {
    return fail(ctx, HVX_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

bool misaligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) != 0; }

int bind_device(hvx_ctx* ctx) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipGetDevice");
    if (cur != ctx->device) {
        e = hipSetDevice(ctx->device);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    }
    return HVX_OK;
}

bool is_pow2(uint32_t v) { return v != 0 && (v & (v - 1)) == 0; }

int check_ubo(hvx_ctx* ctx, const hvx_cf_ubo* u) {
    if (!u) return fail(ctx, HVX_ERR_NULL, "ubo is NULL");
    const uint64_t C = (uint64_t)u->slice_count_x * u->slice_count_y * u->slice_count_z;
    if (u->slice_count_x == 0 || u->slice_count_y == 0 || u->slice_count_z == 0 || C > (1u << 24))
        return fail(ctx, HVX_ERR_INVALID_ARG, "cluster grid %ux%ux%u out of range", u->slice_count_x,
                    u->slice_count_y, u->slice_count_z);
    if (!(u->camera_near > 0.0f) || !(u->camera_far > u->camera_near))
        return fail(ctx, HVX_ERR_INVALID_ARG, "need 0 < near < far");
    return HVX_OK;
}

}  // namespace

namespace hvx {
bool stream_capturing(hipStream_t stream) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &st) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return st != hipStreamCaptureStatusNone;
}

hipError_t ensure_scratch(hvx_ctx* ctx, void** buf, size_t* cur, size_t bytes, hipStream_t stream) {
    if (*cur >= bytes && *buf) return hipSuccess;
    if (stream_capturing(stream)) return hipErrorStreamCaptureUnsupported;
    ctx->scratch_gen += 1;
    if (*buf) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) return e;
        e = hipFree(*buf);
        if (e != hipSuccess) return e;
        *buf = nullptr;
        *cur = 0;
    }
    const size_t sz = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(buf, sz);
    if (e != hipSuccess) {
        *buf = nullptr;
        return e;
    }
    *cur = sz;
    return hipSuccess;
}
}  // namespace hvx

// Fixed culling scratch of hvx_cf_cull_lights_noalloc: fixed-slot lists,
// counts/overflow/partials and view-space lights of one call, carved in that
// order at 256-byte boundaries. Every GPU the library runs on holds its own
// copy for the life of the process. In host code `fixed_cull_scratch` names a
// host-side stand-in, never device memory: host code uses the address
// hipGetSymbolAddress returns for the device (hvx_ctx::fixed_scratch).
constexpr size_t kFixedCullScratchBytes = size_t(24) << 20;
static __device__ __align__(256) unsigned char fixed_cull_scratch[kFixedCullScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_fixed_mutex;
unsigned char* g_fixed_addr[kMaxDevices] = {};

// Device address of `device`'s fixed scratch, looked up on the first request
// for that device (with the device current) and kept for the process.
hipError_t fixed_scratch_address(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_fixed_mutex);
    if (!g_fixed_addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(fixed_cull_scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_fixed_addr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_fixed_addr[device];
    return hipSuccess;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Status of a failed scratch growth: a refusal under graph capture, or the
// runtime's error.
int scratch_fail(hvx_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipErrorStreamCaptureUnsupported)
        return fail(ctx, HVX_ERR_UNSUPPORTED,
                    "%s: the call needs more scratch while its stream is being captured; run it once eagerly at "
                    "this size before capturing", what);
    return hip_fail(ctx, e, what);
}
}  // namespace

extern "C" {

int hvx_abi_version(void) { return HVX_ABI_VERSION; }

uint64_t hvx_ctx_scratch_generation(const hvx_ctx* ctx) { return ctx ? ctx->scratch_gen : 0; }

const char* hvx_status_str(int s) {
    switch (s) {
        case HVX_OK: return "ok";
        case HVX_ERR_NULL: return "null pointer";
        case HVX_ERR_INVALID_ARG: return "invalid argument";
        case HVX_ERR_CAPACITY: return "output buffer too small";
        case HVX_ERR_ALIGN: return "misaligned pointer";
        case HVX_ERR_HIP: return "HIP runtime error";
        case HVX_ERR_UNSUPPORTED: return "unsupported option";
        case HVX_ERR_NO_DEVICE: return "no gfx950 device";
        case HVX_ERR_STALE_GRAPH: return "graph captured before a scratch reallocation";
        default: return "unknown status";
    }
}

const char* hvx_ctx_last_error(const hvx_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int hvx_ctx_set_option(hvx_ctx* ctx, int option, int64_t value) {
    if (!ctx) return HVX_ERR_NULL;
    switch (option) {
        case HVX_OPT_SHADE_EXACT:
            if (value != 0 && value != 1) return fail(ctx, HVX_ERR_INVALID_ARG, "HVX_OPT_SHADE_EXACT is 0 or 1");
            ctx->shade_exact = (int)value;
            return HVX_OK;
        case HVX_OPT_SHADE_KERNEL_EVENTS: {
            if (value < 0 || value > 256)
                return fail(ctx, HVX_ERR_INVALID_ARG, "HVX_OPT_SHADE_KERNEL_EVENTS is in [0, 256]");
            int st;
            if ((st = bind_device(ctx))) return st;
            (void)hipDeviceSynchronize();  // the old events may be pending
            for (hipEvent_t e : ctx->kev) (void)hipEventDestroy(e);
            ctx->kev.clear();
            ctx->kev_next = ctx->kev_count = 0;
            for (int64_t k = 0; k < 2 * value; ++k) {
                hipEvent_t e = nullptr;
                const hipError_t err = hipEventCreate(&e);
                if (err != hipSuccess) {
                    for (hipEvent_t x : ctx->kev) (void)hipEventDestroy(x);
                    ctx->kev.clear();
                    return hip_fail(ctx, err, "HVX_OPT_SHADE_KERNEL_EVENTS");
                }
                ctx->kev.push_back(e);
            }
            return HVX_OK;
        }
        case HVX_OPT_MLP_SEPARATE_LAUNCHES:
            if (value != 0 && value != 1)
                return fail(ctx, HVX_ERR_INVALID_ARG, "HVX_OPT_MLP_SEPARATE_LAUNCHES is 0 or 1");
            ctx->mlp_separate = (int)value;
            return HVX_OK;
        default: return fail(ctx, HVX_ERR_INVALID_ARG, "unknown option %d", option);
    }
}

int hvx_ctx_get_option(const hvx_ctx* ctx, int option, int64_t* value) {
    if (!ctx || !value) return HVX_ERR_NULL;
    switch (option) {
        case HVX_OPT_SHADE_EXACT: *value = ctx->shade_exact; return HVX_OK;
        case HVX_OPT_SHADE_KERNEL_EVENTS: *value = (int64_t)(ctx->kev.size() / 2); return HVX_OK;
        case HVX_OPT_MLP_SEPARATE_LAUNCHES: *value = ctx->mlp_separate; return HVX_OK;
        default: return HVX_ERR_INVALID_ARG;
    }
}

int hvx_ctx_kernel_times(hvx_ctx* ctx, float* ms_out, uint32_t max, uint32_t* n_out) {
    if (!ctx || !n_out || (max && !ms_out)) return HVX_ERR_NULL;
    *n_out = 0;
    const uint32_t pairs = (uint32_t)(ctx->kev.size() / 2);
    if (pairs == 0) return fail(ctx, HVX_ERR_INVALID_ARG, "HVX_OPT_SHADE_KERNEL_EVENTS is off");
    const uint32_t n = std::min(ctx->kev_count, pairs), first = (ctx->kev_next + pairs - n) % pairs;
    const uint32_t out = std::min(n, max);
    for (uint32_t k = 0; k < out; ++k) {
        const uint32_t slot = (first + (n - out) + k) % pairs;
        hipError_t e = hipEventSynchronize(ctx->kev[2 * slot + 1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms_out[k], ctx->kev[2 * slot], ctx->kev[2 * slot + 1]);
        if (e != hipSuccess) return hip_fail(ctx, e, "hvx_ctx_kernel_times");
    }
    ctx->kev_count = 0;
    *n_out = out;
    return HVX_OK;
}

int hvx_ctx_create(int device, hvx_ctx** out) {
    if (!out) return HVX_ERR_NULL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return HVX_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return HVX_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return HVX_ERR_NO_DEVICE;
    hvx_ctx* c = new (std::nothrow) hvx_ctx();
    if (!c) return HVX_ERR_HIP;
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    // the device's fixed culling scratch; without it only
    // hvx_cf_cull_lights_noalloc is unavailable (it reports the error)
    const hipError_t e = fixed_scratch_address(device, &c->fixed_scratch);
    if (e != hipSuccess) {
        c->fixed_scratch = nullptr;
        c->fixed_scratch_error = hipGetErrorString(e);
    }
    *out = c;
    return HVX_OK;
}

int hvx_ctx_destroy(hvx_ctx* ctx) {
    if (!ctx) return HVX_ERR_NULL;
    if (bind_device(ctx) == HVX_OK) {
        (void)hipDeviceSynchronize();
        void* bufs[] = {ctx->cull_lists,   ctx->cull_meta,    ctx->light_vs,  ctx->ggx_table,
                        ctx->shade_lights, ctx->shade_bins,   ctx->diffuse_stage};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        for (hipEvent_t e : ctx->kev) (void)hipEventDestroy(e);
    }
    delete ctx;
    return HVX_OK;
}

uint32_t hvx_mip_count(uint32_t size) {
    // Utility::MipMapCount (HelloVulkan/Header/Utility.h:42-50)
    uint32_t levels = 1;
    while (levels < 32 && (size >> levels)) levels += 1;
    return levels;
}

size_t hvx_cube_level_offset_floats(uint32_t face_size, uint32_t level) {
    return hvx::cube_level_offset_floats(face_size, level);
}

size_t hvx_cube_chain_floats(uint32_t face_size, uint32_t mips) {
    return hvx::cube_level_offset_floats(face_size, mips);
}

// ------------------------------------------------------------- clustered
int hvx_cf_build_aabbs(hvx_ctx* ctx, const hvx_cf_ubo* ubo, hvx_aabb* d_clusters, void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    int st = check_ubo(ctx, ubo);
    if (st) return st;
    if (!d_clusters) return fail(ctx, HVX_ERR_NULL, "d_clusters is NULL");
    if (misaligned(d_clusters, 16)) return fail(ctx, HVX_ERR_ALIGN, "d_clusters must be 16-byte aligned");
    if ((st = bind_device(ctx))) return st;
    hipError_t e = hvx::launch_build_aabbs(*ubo, d_clusters, (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "hvx_cf_build_aabbs launch");
}

}  // extern "C"

namespace {
// Argument checks shared by hvx_cf_cull_lights and its allocation-free
// variant; on success *C_out is the cluster count.
int check_cull_args(hvx_ctx* ctx, const hvx_cf_ubo* ubo, const hvx_aabb* d_clusters, const hvx_light* d_lights,
                    uint32_t n_lights, uint32_t cap, const hvx_light_cell* d_cells, const uint32_t* d_indices,
                    size_t indices_capacity, uint64_t* C_out) {
    int st = check_ubo(ctx, ubo);
    if (st) return st;
    if (!d_clusters || !d_cells || !d_indices) return fail(ctx, HVX_ERR_NULL, "clusters/cells/indices NULL");
    if (n_lights > 0 && !d_lights) return fail(ctx, HVX_ERR_NULL, "d_lights is NULL");
    if (cap < 1 || cap > 1024) return fail(ctx, HVX_ERR_INVALID_ARG, "cap %u not in [1, 1024]", cap);
    if (n_lights > (1u << 28)) return fail(ctx, HVX_ERR_INVALID_ARG, "too many lights (%u)", n_lights);
    if (misaligned(d_lights, 16) || misaligned(d_clusters, 16) || misaligned(d_cells, 8) || misaligned(d_indices, 4))
        return fail(ctx, HVX_ERR_ALIGN, "lights/clusters need 16-byte, cells 8-byte alignment");
    const uint64_t C = (uint64_t)ubo->slice_count_x * ubo->slice_count_y * ubo->slice_count_z;
    if (C * cap > 0xffffffffull)
        return fail(ctx, HVX_ERR_INVALID_ARG, "clusters*cap = %llu exceeds 32-bit list offsets",
                    (unsigned long long)(C * cap));
    if ((uint64_t)indices_capacity < C * cap)
        return fail(ctx, HVX_ERR_CAPACITY, "indices_capacity %zu < clusters*cap = %llu", indices_capacity,
                    (unsigned long long)(C * cap));
    *C_out = C;
    return HVX_OK;
}
}  // namespace

extern "C" {

int hvx_cf_cull_lights(hvx_ctx* ctx, const hvx_cf_ubo* ubo, const hvx_aabb* d_clusters, const hvx_light* d_lights,
                       uint32_t n_lights, uint32_t cap, hvx_light_cell* d_cells, uint32_t* d_indices,
                       size_t indices_capacity, uint32_t* d_total, uint32_t* d_overflow_count, void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    uint64_t C = 0;
    int st = check_cull_args(ctx, ubo, d_clusters, d_lights, n_lights, cap, d_cells, d_indices, indices_capacity, &C);
    if (st) return st;
    if ((st = bind_device(ctx))) return st;
    const hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if ((e = hvx::ensure_scratch(ctx, &ctx->cull_lists, &ctx->cull_lists_bytes, C * cap * 4, s)) != hipSuccess ||
        (e = hvx::ensure_scratch(ctx, &ctx->cull_meta, &ctx->cull_meta_bytes, C * 8 + ((C + 1023) / 1024) * 8 + 64,
                                 s)) != hipSuccess ||
        (e = hvx::ensure_scratch(ctx, &ctx->light_vs, &ctx->light_vs_bytes, (size_t)n_lights * 16, s)) != hipSuccess)
        return scratch_fail(ctx, e, "culling scratch allocation");
    uint32_t* counts = static_cast<uint32_t*>(ctx->cull_meta);
    e = hvx::launch_cull(*ubo, d_clusters, d_lights, n_lights, cap, d_cells, d_indices, d_total, d_overflow_count,
                         static_cast<float4*>(ctx->light_vs), static_cast<uint32_t*>(ctx->cull_lists), counts,
                         counts + C, s);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "hvx_cf_cull_lights launch");
}

size_t hvx_cf_cull_noalloc_scratch_bytes(uint32_t clusters, uint32_t cap, uint32_t n_lights) {
    const size_t C = clusters;
    return align256(C * cap * 4) + align256(C * 8 + ((C + 1023) / 1024) * 8 + 64) + align256((size_t)n_lights * 16);
}

size_t hvx_cf_cull_noalloc_scratch_limit(void) { return kFixedCullScratchBytes; }

int hvx_cf_cull_lights_noalloc(hvx_ctx* ctx, const hvx_cf_ubo* ubo, const hvx_aabb* d_clusters,
                               const hvx_light* d_lights, uint32_t n_lights, uint32_t cap, hvx_light_cell* d_cells,
                               uint32_t* d_indices, size_t indices_capacity, uint32_t* d_total,
                               uint32_t* d_overflow_count, void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    uint64_t C = 0;
    int st = check_cull_args(ctx, ubo, d_clusters, d_lights, n_lights, cap, d_cells, d_indices, indices_capacity, &C);
    if (st) return st;
    if (!ctx->fixed_scratch)
        return fail(ctx, HVX_ERR_HIP, "the fixed culling scratch of device %d is unavailable: %s", ctx->device,
                    ctx->fixed_scratch_error.c_str());
    const size_t lists_b = align256(C * cap * 4), meta_b = align256(C * 8 + ((C + 1023) / 1024) * 8 + 64);
    const size_t need = hvx_cf_cull_noalloc_scratch_bytes((uint32_t)C, cap, n_lights);
    if (need > kFixedCullScratchBytes)
        return fail(ctx, HVX_ERR_CAPACITY,
                    "%llu clusters x cap %u with %u lights need %zu bytes of culling scratch; the fixed scratch of "
                    "hvx_cf_cull_lights_noalloc holds %zu (use hvx_cf_cull_lights)",
                    (unsigned long long)C, cap, n_lights, need, kFixedCullScratchBytes);
    if ((st = bind_device(ctx))) return st;
    // lists, counts + overflow + partials, view-space lights: every word the
    // kernels read is written earlier in the same call, so nothing is cleared
    unsigned char* base = ctx->fixed_scratch;
    uint32_t* lists = reinterpret_cast<uint32_t*>(base);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + lists_b);
    float4* lvs = reinterpret_cast<float4*>(base + lists_b + meta_b);
    const hipError_t e = hvx::launch_cull(*ubo, d_clusters, d_lights, n_lights, cap, d_cells, d_indices, d_total,
                                          d_overflow_count, lvs, lists, counts, counts + C, (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "hvx_cf_cull_lights_noalloc launch");
}

int hvx_cf_cull_lights_batch(hvx_ctx* ctx, const hvx_cf_ubo* ubo, const hvx_aabb* d_clusters,
                             const hvx_light* d_lights, uint32_t n_lights, uint32_t cap, uint32_t flags,
                             hvx_light_cell* d_cells, uint32_t* d_indices, size_t indices_capacity, uint32_t* d_total,
                             void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    int st = check_ubo(ctx, ubo);
    if (st) return st;
    if (!d_clusters || !d_cells || !d_indices) return fail(ctx, HVX_ERR_NULL, "clusters/cells/indices NULL");
    if (n_lights > 0 && !d_lights) return fail(ctx, HVX_ERR_NULL, "d_lights is NULL");
    if (cap < 1 || cap > 1024) return fail(ctx, HVX_ERR_INVALID_ARG, "cap %u not in [1, 1024]", cap);
    if (flags & ~(HVX_CULL_BATCH_REFERENCE_CAP | HVX_CULL_BATCH_REFERENCE_TAIL))
        return fail(ctx, HVX_ERR_INVALID_ARG, "unknown flags 0x%x", flags);
    if (n_lights > (1u << 28)) return fail(ctx, HVX_ERR_INVALID_ARG, "too many lights (%u)", n_lights);
    if (misaligned(d_lights, 16) || misaligned(d_clusters, 16) || misaligned(d_cells, 8) || misaligned(d_indices, 4))
        return fail(ctx, HVX_ERR_ALIGN, "lights/clusters need 16-byte, cells 8-byte alignment");
    const uint64_t C = (uint64_t)ubo->slice_count_x * ubo->slice_count_y * ubo->slice_count_z;
    if (C * cap > 0xffffffffull)
        return fail(ctx, HVX_ERR_INVALID_ARG, "clusters*cap = %llu exceeds 32-bit list offsets",
                    (unsigned long long)(C * cap));
    if ((flags & HVX_CULL_BATCH_REFERENCE_CAP) && C * (uint64_t)(n_lights + hvx::kCullLdsBatch) > 0xffffffffull)
        return fail(ctx, HVX_ERR_INVALID_ARG, "uncapped lists of %llu clusters x %u lights exceed 32-bit offsets",
                    (unsigned long long)C, n_lights);
    if ((uint64_t)indices_capacity < C * cap)
        return fail(ctx, HVX_ERR_CAPACITY, "indices_capacity %zu < clusters*cap = %llu", indices_capacity,
                    (unsigned long long)(C * cap));
    if ((st = bind_device(ctx))) return st;
    const hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if ((e = hvx::ensure_scratch(ctx, &ctx->cull_meta, &ctx->cull_meta_bytes, C * 8 + ((C + 1023) / 1024) * 8 + 64,
                                 s)) != hipSuccess ||
        (e = hvx::ensure_scratch(ctx, &ctx->light_vs, &ctx->light_vs_bytes, (size_t)n_lights * 16, s)) != hipSuccess)
        return scratch_fail(ctx, e, "culling scratch allocation");
    uint32_t* counts = static_cast<uint32_t*>(ctx->cull_meta);
    e = hvx::launch_cull_lds_batch(*ubo, d_clusters, d_lights, n_lights, cap, flags, d_cells, d_indices,
                                   indices_capacity, d_total, static_cast<float4*>(ctx->light_vs), counts, counts + C,
                                   s);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "hvx_cf_cull_lights_batch launch");
}

}  // extern "C"

namespace {
int shade_impl(hvx_ctx* ctx, const hvx_frag_soa* fr, const hvx_camera_ubo* cam, const hvx_cf_ubo* ubo,
               const hvx_pbr_pc* pc, const hvx_light* d_lights, uint32_t n_lights, const hvx_light_cell* d_cells,
               const uint32_t* d_indices, const hvx_ibl_maps* ibl, float* d_rgb, const hvx::LearnedShade* learned,
               void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    int st = check_ubo(ctx, ubo);
    if (st) return st;
    if (!fr || !cam || !pc || !ibl || !d_rgb || !d_cells || !d_indices)
        return fail(ctx, HVX_ERR_NULL, "NULL argument");
    if (!fr->depth || !fr->roughness || !fr->metallic) return fail(ctx, HVX_ERR_NULL, "fragment plane NULL");
    for (int c = 0; c < 3; ++c)
        if (!fr->normal[c] || !fr->albedo[c]) return fail(ctx, HVX_ERR_NULL, "fragment plane NULL");
    if (fr->width == 0 || fr->height == 0 || (uint64_t)fr->width * fr->height > (1ull << 30))
        return fail(ctx, HVX_ERR_INVALID_ARG, "fragment grid %ux%u out of range", fr->width, fr->height);
    if (ubo->screen_size[0] != (float)fr->width || ubo->screen_size[1] != (float)fr->height)
        return fail(ctx, HVX_ERR_INVALID_ARG, "ubo screen size %gx%g != fragment grid %ux%u", ubo->screen_size[0],
                    ubo->screen_size[1], fr->width, fr->height);
    if (n_lights > 0 && !d_lights) return fail(ctx, HVX_ERR_NULL, "d_lights is NULL");
    if (misaligned(d_lights, 16)) return fail(ctx, HVX_ERR_ALIGN, "d_lights must be 16-byte aligned");
    if (!ibl->specular || !ibl->diffuse || !ibl->brdf_lut) return fail(ctx, HVX_ERR_NULL, "IBL map NULL");
    if (ibl->specular_size == 0 || ibl->specular_mips == 0 || ibl->specular_mips > hvx_mip_count(ibl->specular_size) ||
        ibl->diffuse_size == 0 || ibl->lut_width == 0 || ibl->lut_height == 0)
        return fail(ctx, HVX_ERR_INVALID_ARG, "IBL map sizes out of range");
    if (misaligned(ibl->specular, 16) || misaligned(ibl->diffuse, 16) || misaligned(ibl->brdf_lut, 8))
        return fail(ctx, HVX_ERR_ALIGN, "IBL maps must be 16-byte (cubes) / 8-byte (LUT) aligned");
    if ((st = bind_device(ctx))) return st;
    hipError_t e = hipSuccess;
    if (!ctx->shade_exact &&
        (e = hvx::ensure_scratch(ctx, &ctx->shade_lights, &ctx->shade_lights_bytes, (size_t)n_lights * 32,
                                 (hipStream_t)stream)) != hipSuccess)
        return scratch_fail(ctx, e, "shading light table allocation");
    const bool binned = hvx::shade_binned(*ubo);
    if (binned && (e = hvx::ensure_scratch(ctx, &ctx->shade_bins, &ctx->shade_bins_bytes,
                                           hvx::shade_bin_scratch_bytes(*fr, *ubo), (hipStream_t)stream)) !=
                      hipSuccess)
        return scratch_fail(ctx, e, "shading bin allocation");
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (!ctx->kev.empty()) {
        const uint32_t pairs = (uint32_t)(ctx->kev.size() / 2), slot = ctx->kev_next;
        ev0 = ctx->kev[2 * slot];
        ev1 = ctx->kev[2 * slot + 1];
        ctx->kev_next = (slot + 1) % pairs;
        ctx->kev_count = std::min(ctx->kev_count + 1, pairs);
    }
    e = hvx::launch_shade(*fr, *cam, *ubo, *pc, d_lights, n_lights, d_cells, d_indices, *ibl, d_rgb,
                          ctx->shade_exact != 0, static_cast<float4*>(ctx->shade_lights),
                          binned ? static_cast<uint32_t*>(ctx->shade_bins) : nullptr, learned, (hipStream_t)stream,
                          ev0, ev1);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "shading launch");
}
}  // namespace

extern "C" {

int hvx_cf_shade(hvx_ctx* ctx, const hvx_frag_soa* fr, const hvx_camera_ubo* cam, const hvx_cf_ubo* ubo,
                 const hvx_pbr_pc* pc, const hvx_light* d_lights, uint32_t n_lights, const hvx_light_cell* d_cells,
                 const uint32_t* d_indices, const hvx_ibl_maps* ibl, float* d_rgb, void* stream) {
    return shade_impl(ctx, fr, cam, ubo, pc, d_lights, n_lights, d_cells, d_indices, ibl, d_rgb, nullptr, stream);
}

int hvx_cf_shade_learned(hvx_ctx* ctx, const hvx_frag_soa* fr, const hvx_camera_ubo* cam, const hvx_cf_ubo* ubo,
                         const hvx_pbr_pc* pc, const hvx_light* d_lights, uint32_t n_lights,
                         const hvx_light_cell* d_cells, const uint32_t* d_indices, const hvx_ibl_maps* ibl,
                         const hvx_mlp* mlp, const float* d_cell_err, uint32_t cell_face, uint32_t lod_bins,
                         float tol, float* d_rgb, uint32_t* d_served, void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    if (!mlp || !d_cell_err || !pc) return fail(ctx, HVX_ERR_NULL, "mlp, cell errors and push constants are required");
    if (cell_face == 0 || cell_face > 1024 || lod_bins == 0 || lod_bins > 256 || !(tol >= 0.0f))
        return fail(ctx, HVX_ERR_INVALID_ARG, "cell_face in [1, 1024], lod_bins in [1, 256], tol >= 0");
    hvx::LearnedShade L;
    L.mlp = hvx::mlp_view(mlp);
    if (!(pc->max_reflection_lod >= 0.0f) || pc->max_reflection_lod > L.mlp.lod_max)
        return fail(ctx, HVX_ERR_INVALID_ARG, "maxReflectionLod %g exceeds the network's lod range %g",
                    pc->max_reflection_lod, L.mlp.lod_max);
    L.cell_err = d_cell_err;
    L.cell_lo = hvx::mlp_cell_lower(mlp, cell_face, lod_bins);
    L.mask_face = cell_face;
    L.lod_bins = lod_bins;
    L.tol = tol;
    L.served = d_served;
    return shade_impl(ctx, fr, cam, ubo, pc, d_lights, n_lights, d_cells, d_indices, ibl, d_rgb, &L, stream);
}

// -------------------------------------------------------------- tonemap
int hvx_post_tonemap(hvx_ctx* ctx, const float* d_rgb, uint32_t width, uint32_t height, int op, int gamma,
                     float* d_out_rgb, uint32_t* d_out_bgra8, void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    if (!d_rgb || (!d_out_rgb && !d_out_bgra8)) return fail(ctx, HVX_ERR_NULL, "input or both outputs NULL");
    if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 30))
        return fail(ctx, HVX_ERR_INVALID_ARG, "image %ux%u out of range", width, height);
    if (op < HVX_TONEMAP_REINHARD || op > HVX_TONEMAP_UNCHARTED2)
        return fail(ctx, HVX_ERR_INVALID_ARG, "unknown tonemap operator %d", op);
    if (misaligned(d_rgb, 4) || misaligned(d_out_rgb, 4) || misaligned(d_out_bgra8, 4))
        return fail(ctx, HVX_ERR_ALIGN, "buffers must be 4-byte aligned");
    int st;
    if ((st = bind_device(ctx))) return st;
    const hipError_t e = hvx::launch_tonemap(d_rgb, width * height, op, gamma ? 1 : 0, d_out_rgb, d_out_bgra8,
                                             (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "hvx_post_tonemap launch");
}

// ------------------------------------------------------------------- IBL
int hvx_ibl_brdf_lut(hvx_ctx* ctx, uint32_t width, uint32_t height, uint32_t sample_count, float* d_lut_rg,
                     void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    if (!d_lut_rg) return fail(ctx, HVX_ERR_NULL, "d_lut_rg is NULL");
    if (width == 0 || width > 1024 || height == 0 || height > 65535 || sample_count == 0)
        return fail(ctx, HVX_ERR_INVALID_ARG, "LUT %ux%u (width <= 1024) with %u samples", width, height,
                    sample_count);
    if (misaligned(d_lut_rg, 8)) return fail(ctx, HVX_ERR_ALIGN, "d_lut_rg must be 8-byte aligned");
    int st;
    if ((st = bind_device(ctx))) return st;
    hipError_t e = hvx::launch_brdf_lut(width, height, sample_count, d_lut_rg, (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "hvx_ibl_brdf_lut launch");
}

int hvx_ibl_equirect_to_cube(hvx_ctx* ctx, const float* d_eq, uint32_t w, uint32_t h, float* d_cube,
                             uint32_t face_size, void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    if (!d_eq || !d_cube) return fail(ctx, HVX_ERR_NULL, "NULL image");
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 28) || face_size == 0 || face_size > 16384)
        return fail(ctx, HVX_ERR_INVALID_ARG, "sizes out of range");
    if (misaligned(d_eq, 16) || misaligned(d_cube, 16)) return fail(ctx, HVX_ERR_ALIGN, "images must be 16-byte aligned");
    int st;
    if ((st = bind_device(ctx))) return st;
    hipError_t e = hvx::launch_equirect_to_cube(d_eq, w, h, d_cube, face_size, (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "hvx_ibl_equirect_to_cube launch");
}

int hvx_ibl_cube_mips(hvx_ctx* ctx, float* d_chain, uint32_t face_size, uint32_t levels, void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    if (!d_chain) return fail(ctx, HVX_ERR_NULL, "d_cube_chain is NULL");
    if (!is_pow2(face_size) || levels == 0 || levels > hvx_mip_count(face_size))
        return fail(ctx, HVX_ERR_INVALID_ARG, "face_size must be a power of two, 1 <= levels <= %u",
                    hvx_mip_count(face_size));
    if (misaligned(d_chain, 16)) return fail(ctx, HVX_ERR_ALIGN, "d_cube_chain must be 16-byte aligned");
    int st;
    if ((st = bind_device(ctx))) return st;
    hipError_t e = hvx::launch_cube_mips(d_chain, face_size, levels, (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "hvx_ibl_cube_mips launch");
}

int hvx_ibl_filter(hvx_ctx* ctx, int type, const float* d_cube, uint32_t face_size, float* d_out,
                   uint32_t out_face_size, uint32_t out_mips, uint32_t sample_count, const hvx_filter_opts* opts,
                   void* stream) {
    if (!ctx) return HVX_ERR_NULL;
    if (!d_cube || !d_out) return fail(ctx, HVX_ERR_NULL, "NULL cube");
    if (face_size == 0 || face_size > 16384 || out_face_size == 0 || out_face_size > 16384)
        return fail(ctx, HVX_ERR_INVALID_ARG, "face sizes out of range");
    if (misaligned(d_cube, 16) || misaligned(d_out, 16)) return fail(ctx, HVX_ERR_ALIGN, "cubes must be 16-byte aligned");
    const bool emulate = opts && opts->emulate_static_viewport;
    int st;
    if (type == HVX_FILTER_DIFFUSE) {
        if (out_mips != 1) return fail(ctx, HVX_ERR_INVALID_ARG, "diffuse output has exactly 1 mip");
        if ((st = bind_device(ctx))) return st;
        // staged footprints in context scratch; under stream capture with no
        // buffer large enough yet, the taps are gathered from the cube (the
        // same results, without the allocation capture forbids)
        const size_t need = hvx::diffuse_stage_size(face_size);
        float* stage = nullptr;
        if (need && (ctx->diffuse_stage_bytes >= need || !hvx::stream_capturing((hipStream_t)stream))) {
            hipError_t e = hvx::ensure_scratch(ctx, &ctx->diffuse_stage, &ctx->diffuse_stage_bytes, need,
                                               (hipStream_t)stream);
            if (e != hipSuccess) return scratch_fail(ctx, e, "diffuse staging allocation");
            stage = static_cast<float*>(ctx->diffuse_stage);
        }
        hipError_t e = hvx::launch_diffuse(d_cube, face_size, d_out, out_face_size, stage, (hipStream_t)stream);
        return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "diffuse launch");
    }
    if (type != HVX_FILTER_SPECULAR) return fail(ctx, HVX_ERR_INVALID_ARG, "unknown filter type %d", type);
    if (out_mips == 0 || out_mips > hvx_mip_count(out_face_size) || out_mips > 16)
        return fail(ctx, HVX_ERR_INVALID_ARG, "out_mips %u not in [1, %u]", out_mips,
                    std::min(16u, hvx_mip_count(out_face_size)));
    if (sample_count == 0 || sample_count > (1u << 20))
        return fail(ctx, HVX_ERR_INVALID_ARG, "sample_count %u not in [1, 2^20]", sample_count);
    if ((st = bind_device(ctx))) return st;
    hipError_t e = hvx::ensure_scratch(ctx, &ctx->ggx_table, &ctx->ggx_table_bytes, (size_t)out_mips * sample_count * 16,
                                       (hipStream_t)stream);
    if (e != hipSuccess) return scratch_fail(ctx, e, "GGX table allocation");
    e = hvx::launch_specular(d_cube, face_size, d_out, out_face_size, out_mips, sample_count, emulate,
                             static_cast<float4*>(ctx->ggx_table), (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : hip_fail(ctx, e, "specular launch");
}

}  // extern "C"
