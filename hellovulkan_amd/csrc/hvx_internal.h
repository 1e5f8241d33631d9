// hvx_internal.h — context and kernel-launcher declarations shared by the
// translation units of libhvx (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/hvx.h"

struct hvx_ctx {
    int device = 0;
    int num_cus = 0;
    std::string last_error;
    // growable device scratch, owned by the context
    void* cull_lists = nullptr;   // uint32 [clusters * cap] fixed-slot lists
    size_t cull_lists_bytes = 0;
    void* cull_meta = nullptr;    // uint32 counts[C], overflow[C], uint2 block partials
    size_t cull_meta_bytes = 0;
    void* light_vs = nullptr;     // float4 [lights]: view-space centre, r^2
    size_t light_vs_bytes = 0;
    void* ggx_table = nullptr;    // float4 [mips * samples]: tangent-space GGX H
    size_t ggx_table_bytes = 0;
    void* shade_lights = nullptr; // float4 [2 * lights]: compact shading light table
    size_t shade_lights_bytes = 0;
    // options (hvx_ctx_set_option)
    int shade_exact = 0;
};

namespace hvx {

// Grows a scratch buffer to at least `bytes` (synchronises the device when it
// has to replace an existing allocation). Returns a hipError_t.
hipError_t ensure_scratch(hvx_ctx* ctx, void** buf, size_t* cur, size_t bytes);

// --- clustered forward (cf_kernels.hip)
hipError_t launch_build_aabbs(const hvx_cf_ubo& ubo, hvx_aabb* d_clusters, hipStream_t s);
hipError_t launch_cull(const hvx_cf_ubo& ubo, const hvx_aabb* d_clusters, const hvx_light* d_lights,
                       uint32_t n_lights, uint32_t cap, hvx_light_cell* d_cells, uint32_t* d_indices,
                       uint32_t* d_total, uint32_t* d_overflow, float4* light_vs, uint32_t* lists,
                       uint32_t* counts, uint32_t* over, hipStream_t s);
// --- shading (shade_kernels.hip); light_scratch holds 2 * n_lights float4
hipError_t launch_shade(const hvx_frag_soa& fr, const hvx_camera_ubo& cam, const hvx_cf_ubo& ubo,
                        const hvx_pbr_pc& pc, const hvx_light* d_lights, uint32_t n_lights,
                        const hvx_light_cell* d_cells, const uint32_t* d_indices, const hvx_ibl_maps& ibl,
                        float* d_rgb, bool exact, float4* light_scratch, hipStream_t s);

// --- IBL (ibl_kernels.hip)
hipError_t launch_brdf_lut(uint32_t w, uint32_t h, uint32_t n, float* d_lut, hipStream_t s);
hipError_t launch_equirect_to_cube(const float* d_eq, uint32_t w, uint32_t h, float* d_cube, uint32_t S,
                                   hipStream_t s);
hipError_t launch_cube_mips(float* d_chain, uint32_t S, uint32_t levels, hipStream_t s);
hipError_t launch_diffuse(const float* d_cube, uint32_t S, float* d_out, uint32_t Sd, hipStream_t s);
hipError_t launch_specular(const float* d_cube, uint32_t S, float* d_out, uint32_t So, uint32_t mips,
                           uint32_t n, bool emulate_static_viewport, float4* ggx_table, hipStream_t s);

}  // namespace hvx
