// hvx_internal.h — context and kernel-launcher declarations shared by the
// translation units of libhvx (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hvx.h"
#include "mlp_device.h"

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
    void* diffuse_stage = nullptr; // the diffuse filter's staged footprints (64 B each)
    size_t diffuse_stage_bytes = 0;
    void* shade_lights = nullptr; // float4 [2 * lights]: compact shading light table
    size_t shade_lights_bytes = 0;
    void* shade_bins = nullptr;   // uint32 [fragments + X + Y + 2]: cluster-binned fragment order
    size_t shade_bins_bytes = 0;
    // bumped whenever a scratch buffer is (re)allocated: a graph captured
    // before the bump may hold pointers to freed scratch
    uint64_t scratch_gen = 0;
    // device address of the library's fixed culling scratch on `device`
    // (hvx_cf_cull_lights_noalloc), looked up once per device at creation
    unsigned char* fixed_scratch = nullptr;
    std::string fixed_scratch_error;  // why the lookup failed, if it did
    // options (hvx_ctx_set_option)
    int shade_exact = 0;
    int mlp_separate = 0;  // HVX_OPT_MLP_SEPARATE_LAUNCHES
    // HVX_OPT_SHADE_KERNEL_EVENTS: event pairs around shading-kernel launches
    // (a ring of kev.size() / 2 pairs; kev_count recorded since the last read)
    std::vector<hipEvent_t> kev;
    uint32_t kev_next = 0, kev_count = 0;
};

namespace hvx {

// One view's culling work (cf_kernels.hip).
struct CullJob {
    uint32_t X, Y, Z;
    uint32_t n, cap, want_overflow;
    const hvx_aabb* clusters;
    const float4* lvs;
    uint32_t* lists;
    uint32_t* counts;
    uint32_t* over;
};

// A view of a culling batch, device-resident: its CullJob, its inputs and
// outputs, and the running item offsets that map batched work items to views.
struct CullBatchView {
    CullJob job;
    const hvx_light* lights;
    float view[16];
    uint2* part;
    hvx_light_cell* cells;
    uint32_t* indices;
    uint32_t* total;
    uint32_t* novf;
    uint32_t C, nb;
    uint32_t light_begin, group_begin, block_begin, scatter_begin;
};
struct CullBatchTotals {
    uint32_t lights, groups, blocks, scatter;
};

// One probe of a prefilter batch (ibl_kernels.hip): level-0 input cube,
// output chain, and the offset of its mip count's GGX table.
struct ProbeJob {
    const float4* cube;
    float4* out;
    uint32_t S, So, mips, n;
    uint32_t table;
};
// A work item of one mip of one probe: texels [first, first + count) of mip
// 0, or padded 4x4-texel tiles [first, first + count) of a mip >= 1.
struct ProbeItem {
    uint32_t probe, mip, first, count;
};

// Grows a scratch buffer to at least `bytes` (synchronises the device when it
// has to replace an existing allocation). A growth while `stream` is being
// captured into a graph is refused with hipErrorStreamCaptureUnsupported
// (allocation and synchronisation are illegal there). Returns a hipError_t.
hipError_t ensure_scratch(hvx_ctx* ctx, void** buf, size_t* cur, size_t bytes, hipStream_t stream);
// Whether `stream` is being captured into a graph.
bool stream_capturing(hipStream_t stream);

// --- clustered forward (cf_kernels.hip)
hipError_t launch_build_aabbs(const hvx_cf_ubo& ubo, hvx_aabb* d_clusters, hipStream_t s);
hipError_t launch_cull(const hvx_cf_ubo& ubo, const hvx_aabb* d_clusters, const hvx_light* d_lights,
                       uint32_t n_lights, uint32_t cap, hvx_light_cell* d_cells, uint32_t* d_indices,
                       uint32_t* d_total, uint32_t* d_overflow, float4* light_vs, uint32_t* lists,
                       uint32_t* counts, uint32_t* over, hipStream_t s);
// LightCullingBatch.comp semantics (one lane per cluster, lights staged in
// LDS batches of kCullLdsBatch): pass 0 writes per-cluster counts and
// overflow flags, pass 1 (after the scan) writes the packed lists. flags are
// HVX_CULL_BATCH_REFERENCE_* of hvx.h.
constexpr uint32_t kCullLdsBatch = 576;
hipError_t launch_cull_lds_batch(const hvx_cf_ubo& ubo, const hvx_aabb* d_clusters, const hvx_light* d_lights,
                                 uint32_t n_lights, uint32_t cap, uint32_t flags, hvx_light_cell* d_cells,
                                 uint32_t* d_indices, size_t indices_capacity, uint32_t* d_total, float4* light_vs,
                                 uint32_t* counts, uint32_t* over, hipStream_t s);
// cull group shapes: 0 = 4x4x1, 1 = 8x4x1, 2 = 4x4x2, 3 = 2x2x4
constexpr int kCullShapes = 4;
uint32_t cull_batch_groups(int shape, uint32_t X, uint32_t Y, uint32_t Z);
hipError_t launch_cull_batch(const CullBatchView* d_views, uint32_t nv, const CullBatchTotals& t, int shape,
                             uint32_t persistent_wgs, uint32_t* d_queue, hipStream_t s);
hipError_t launch_queue_reset(uint32_t* d_queues, uint32_t n, hipStream_t s);

// --- shading (shade_kernels.hip); light_scratch holds 2 * n_lights float4 (a 32-byte record per light)
// Learned prefiltered radiance for shading: the network and its fallback mask.
struct LearnedShade {
    MlpView mlp;
    const float* cell_err;
    const float4* cell_lo;  // nullable: per-cell estimate from below of the network's output (RGB, pad)
    uint32_t mask_face, lod_bins;
    float tol;
    uint32_t* served;
};
MlpView mlp_view(const hvx_mlp* mlp);
// the cell lower estimates of the last hvx_mlp_cell_errors call on that grid, else nullptr
const float4* mlp_cell_lower(const hvx_mlp* mlp, uint32_t cell_face, uint32_t lod_bins);

hipError_t launch_tonemap(const float* d_rgb, uint32_t n, int op, int gamma, float* d_out, uint32_t* d_bgra,
                          hipStream_t s);

hipError_t launch_shade(const hvx_frag_soa& fr, const hvx_camera_ubo& cam, const hvx_cf_ubo& ubo,
                        const hvx_pbr_pc& pc, const hvx_light* d_lights, uint32_t n_lights,
                        const hvx_light_cell* d_cells, const uint32_t* d_indices, const hvx_ibl_maps& ibl,
                        float* d_rgb, bool exact, float4* light_scratch, uint32_t* bin_scratch,
                        const LearnedShade* learned, hipStream_t s, hipEvent_t ev_begin = nullptr,
                        hipEvent_t ev_end = nullptr);
// Whether shading bins fragments by cluster for this grid, and the scratch
// it needs for that (the fragment order and the tile bounds).
bool shade_binned(const hvx_cf_ubo& ubo);
size_t shade_bin_scratch_bytes(const hvx_frag_soa& fr, const hvx_cf_ubo& ubo);

// --- IBL (ibl_kernels.hip)
hipError_t launch_brdf_lut(uint32_t w, uint32_t h, uint32_t n, float* d_lut, hipStream_t s);
hipError_t launch_equirect_to_cube(const float* d_eq, uint32_t w, uint32_t h, float* d_cube, uint32_t S,
                                   hipStream_t s);
// n chains chain_stride_floats apart (one chain by default)
hipError_t launch_cube_mips(float* d_chain, uint32_t S, uint32_t levels, hipStream_t s, uint32_t n = 1,
                            size_t chain_stride_floats = 0);
// d_dst + p * dst_stride_floats <- d_src[p] (a device array of n pointers), `floats` floats each
hipError_t launch_gather_copy(const float* const* d_src, size_t floats, float* d_dst, size_t dst_stride_floats,
                              uint32_t n, hipStream_t s);
// Bytes of the diffuse filter's staged footprints for a level-0 face of S
// (0: not staged: S < 2, or more than kDiffuseStageMaxBytes).
constexpr size_t kDiffuseStageMaxBytes = size_t(1) << 30;
size_t diffuse_stage_size(uint32_t S);
// d_stage: diffuse_stage_size(S) bytes of scratch, or nullptr (taps
// gathered from the cube; the same results)
hipError_t launch_diffuse(const float* d_cube, uint32_t S, float* d_out, uint32_t Sd, float* d_stage, hipStream_t s);
hipError_t launch_ggx_tables(uint32_t mips, uint32_t n, float4* table, hipStream_t s);
// d_queues: the eight per-XCD probe queues (zeroed before each run)
hipError_t launch_probe_batch(const ProbeJob* d_jobs, const ProbeItem* d_items, const uint32_t* d_xcd_begin,
                              const float4* d_tables, uint32_t persistent_wgs, uint32_t* d_queues, hipStream_t s);
// Work items of a probe batch (probe, mip, tile order; items of about
// `texels` output texels) and the item runs of the eight XCD queues.
void make_probe_items(const ProbeJob* jobs, uint32_t n_probes, uint32_t texels, std::vector<ProbeItem>& items,
                      uint32_t xcd_begin[9]);
constexpr uint32_t kProbeQueues = 8;  // per-XCD probe queues; the culling queue follows them
hipError_t launch_specular(const float* d_cube, uint32_t S, float* d_out, uint32_t So, uint32_t mips,
                           uint32_t n, bool emulate_static_viewport, float4* ggx_table, hipStream_t s);

}  // namespace hvx
