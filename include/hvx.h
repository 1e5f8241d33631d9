/*
 * hvx.h — C ABI of the MI355X-native clustered-forward + image-based-lighting
 * compute path.
 *
 * Each entry point replaces one pipeline stage of HelloVulkan's
 * clustered-forward / IBL path (the stage "plugins" derived from PipelineBase,
 * HelloVulkan/Header/Pipelines/PipelineBase.h:24-95). The caller owns all
 * device memory; every call is asynchronous on the HIP stream passed as
 * `void* stream` (NULL = the default stream), returns an hvx_status code and
 * never throws. A context must not be used by two host threads at once, and
 * the stages of one context run in order on one stream: they share the
 * context's scratch buffers, so calls on different streams race (use one
 * context per stream for concurrent work).
 *
 * Graph capture: a call that would grow the context's scratch while its
 * stream is being captured is refused (HVX_ERR_UNSUPPORTED), since
 * allocation and synchronisation are illegal there; run each stage once
 * eagerly at its largest size first. A graph holds the scratch pointers it
 * was captured with: when a later eager call grows the scratch (the value
 * of hvx_ctx_scratch_generation changes) those pointers are freed, and the
 * graph must be captured again before it is replayed.
 *
 * Struct layouts mirror the reference's GPU-visible layouts byte for byte
 * (std430 / std140 as declared in the GLSL and the C++ UBO structs) so that a
 * reference-side caller can hand its buffers over unchanged.
 */
#ifndef HVX_H_
#define HVX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HVX_ABI_VERSION 1

/* ------------------------------------------------------------------ status */
typedef enum hvx_status {
    HVX_OK = 0,
    HVX_ERR_NULL = 1,          /* a required pointer is NULL                   */
    HVX_ERR_INVALID_ARG = 2,   /* a size / count / enum is out of range        */
    HVX_ERR_CAPACITY = 3,      /* an output buffer is too small                */
    HVX_ERR_ALIGN = 4,         /* a device pointer is misaligned               */
    HVX_ERR_HIP = 5,           /* the HIP runtime reported an error            */
    HVX_ERR_UNSUPPORTED = 6,   /* option combination not implemented, or a call
                                  that would allocate during graph capture      */
    HVX_ERR_NO_DEVICE = 7,     /* no usable gfx950 device                      */
    HVX_ERR_STALE_GRAPH = 8    /* a captured graph refers to context scratch
                                  that has since been reallocated              */
} hvx_status;

/* ------------------------------------------------------------------ types  */

/* LightData (HelloVulkan/Header/Resources/ResourcesLight.h:13-21,
 * Shaders/LightData.glsl:1-6): std430 stride 48 B. */
typedef struct hvx_light {
    float position[4]; /* world space, w = 1 */
    float color[4];    /* linear RGB, w = 1  */
    float radius;
    float _pad[3];
} hvx_light;

/* AABB (ResourcesLight.h:24-30, Shaders/AABB/AABB.glsl:1-5): view space. */
typedef struct hvx_aabb {
    float min_point[4]; /* w = 0 */
    float max_point[4]; /* w = 0 */
} hvx_aabb;

/* LightCell (ResourcesLight.h:33-39, ClusteredForward/Header.glsl:1-5). */
typedef struct hvx_light_cell {
    uint32_t offset;
    uint32_t count;
} hvx_light_cell;

/* ClusterForwardUBO (HelloVulkan/Header/UBOs.h:67-89,
 * ClusteredForward/Header.glsl:7-19). Matrices are column-major (glm). 176 B. */
typedef struct hvx_cf_ubo {
    float camera_inverse_projection[16];
    float camera_view[16];
    float screen_size[2];
    float slice_scaling;
    float slice_bias;
    float camera_near;
    float camera_far;
    uint32_t slice_count_x;
    uint32_t slice_count_y;
    uint32_t slice_count_z;
    uint32_t _pad[3];
} hvx_cf_ubo;

/* CameraUBO (UBOs.h:7-19, Shaders/CameraUBO.glsl:1-8). 160 B. */
typedef struct hvx_camera_ubo {
    float projection[16];
    float view[16];
    float position[4];
    float camera_near;
    float camera_far;
    float _pad[2];
} hvx_camera_ubo;

/* PushConstPBR (HelloVulkan/Header/PushConstants.h:22-29). 20 B. */
typedef struct hvx_pbr_pc {
    float light_intensity;
    float base_reflectivity;
    float max_reflection_lod;
    float light_falloff;
    float albedo_multiplier;
} hvx_pbr_pc;

/* Struct-of-arrays fragments: the rasterised G-buffer inputs of
 * ClusteredForward/Scene.frag:60-149, one fragment per pixel, row-major
 * [height][width] (pixel (x, y) has gl_FragCoord.xy = (x + .5, y + .5)).
 * World position is rebuilt from (pixel centre, depth) with inv_view_proj. */
typedef struct hvx_frag_soa {
    const float* depth;       /* gl_FragCoord.z, NDC depth in [0,1)          */
    const float* normal[3];   /* world-space unit normal, planes x, y, z     */
    const float* albedo[3];   /* linear albedo, planes r, g, b               */
    const float* roughness;   /* [0,1]                                       */
    const uint8_t* metallic;  /* UNORM8 (value / 255)                        */
    uint32_t width;
    uint32_t height;
    float inv_view_proj[16];  /* column-major inverse(P * V)                 */
} hvx_frag_soa;

/* The three IBL maps Scene.frag samples (bindings 7, 8, 9). Cube maps are
 * RGBA32F, mip-major, then face (+X,-X,+Y,-Y,+Z,-Z), then row, then column;
 * level m has side max(size >> m, 1). The LUT is RG32F [height][width]. */
typedef struct hvx_ibl_maps {
    const float* specular;
    uint32_t specular_size;
    uint32_t specular_mips;
    const float* diffuse;
    uint32_t diffuse_size;
    uint32_t _pad0;
    const float* brdf_lut;
    uint32_t lut_width;
    uint32_t lut_height;
} hvx_ibl_maps;

/* Cube-filter type (PipelineCubeFilter.h CubeFilterType). */
typedef enum hvx_filter_type {
    HVX_FILTER_DIFFUSE = 0,
    HVX_FILTER_SPECULAR = 1
} hvx_filter_type;

/* Cube-filter options. All-zero = the reference's behaviour as it renders
 * its intended mapping (every output mip covers its whole face). */
typedef struct hvx_filter_opts {
    /* 1: reproduce the reference's static base-size viewport for specular
     * mips >= 1 (PipelineCubeFilter.cpp:189-192, :303): mip i then only
     * covers uv in [-1, -1 + 2^(1-i)] of each face. */
    uint32_t emulate_static_viewport;
    uint32_t _reserved[3];
} hvx_filter_opts;

typedef struct hvx_ctx hvx_ctx;

/* ------------------------------------------------------------------ context */

/* Creates a context bound to HIP device `device` (must be gfx950). Replaces
 * the VulkanContext the reference's stages are constructed with
 * (PipelineBase.h:27, VulkanContext.cpp:11-44). */
int hvx_ctx_create(int device, hvx_ctx** out);
int hvx_ctx_destroy(hvx_ctx* ctx);
const char* hvx_status_str(int status);
/* Text of the last error raised on this context ("" if none). */
const char* hvx_ctx_last_error(const hvx_ctx* ctx);
int hvx_abi_version(void);

/* Context options (no counterpart in the reference, whose arithmetic is fixed
 * by the driver's shader compiler). */
typedef enum hvx_option {
    /* 0 (default): shading keeps the light/half vectors and their cosines
     * exact and evaluates the remaining per-light terms with fused
     * multiply-add and hardware reciprocal/exp2/log2 (within 1e-5 relative of
     * the exact evaluation; waves whose fragments are all rough take hardware
     * reciprocal square roots, and which fragments share a wave can change
     * between calls, so results may differ in the last bits from call to
     * call); 1: every operation exactly as the FP32 definition of the path
     * (bit-reproducible on any IEEE machine). */
    HVX_OPT_SHADE_EXACT = 1,
    /* 0 (default): off; k in [1, 256]: record a pair of HIP events around
     * each launch of the shading kernel itself (not the binning and light
     * table launches of the stage), keeping the last k pairs, for
     * hvx_ctx_kernel_times. For measurement. */
    HVX_OPT_SHADE_KERNEL_EVENTS = 2,
    /* 0 (default): a network training step (hvx_mlp_gradient, each step of
     * hvx_mlp_train) runs its per-sample work (three forward layers, output
     * layer and loss, two backward layer products) in one launch; 1: as
     * separate layer launches. Both give the same bits; 1 is the reference
     * the fused step is tested against. */
    HVX_OPT_MLP_SEPARATE_LAUNCHES = 3
} hvx_option;
int hvx_ctx_set_option(hvx_ctx* ctx, int option, int64_t value);
int hvx_ctx_get_option(const hvx_ctx* ctx, int option, int64_t* value);
/* Durations (ms, oldest first) of the shading-kernel launches recorded
 * since the last call (at most the HVX_OPT_SHADE_KERNEL_EVENTS depth):
 * waits for the last recorded launch, writes min(recorded, max) values and
 * their count, and clears the record. No counterpart in the reference. */
int hvx_ctx_kernel_times(hvx_ctx* ctx, float* ms_out, uint32_t max, uint32_t* n_out);
/* Counter of the context's scratch (re)allocations. A graph captured at one
 * value may be replayed only while the value is unchanged. No counterpart in
 * the reference (its descriptor sets bind fixed buffers). */
uint64_t hvx_ctx_scratch_generation(const hvx_ctx* ctx);

/* ------------------------------------------------------- clustered forward */

/* Cluster AABBs for every froxel of the X*Y*Z grid of `ubo`.
 * Replaces PipelineAABBGenerator::FillCommandBuffer
 * (HelloVulkan/Source/Pipelines/PipelineAABBGenerator.cpp:36-83) running
 * Shaders/ClusteredForward/AABBGenerator.comp:39-89.
 * d_clusters: X*Y*Z entries, index x + y*X + z*X*Y. */
int hvx_cf_build_aabbs(hvx_ctx* ctx, const hvx_cf_ubo* ubo,
                       hvx_aabb* d_clusters, void* stream);

/* Per-cluster light lists: for each cluster, the first `cap` lights in
 * ascending index whose sphere intersects the cluster AABB (strict
 * d^2 < r^2). Replaces PipelineLightCulling::ResetGlobalIndex +
 * FillCommandBuffer (PipelineLightCulling.cpp:38-119) running
 * Shaders/ClusteredForward/LightCulling.comp:59-87 (cap there is 150).
 * The reference reserves list ranges with a global atomicAdd; here offsets
 * are the exclusive prefix sum of counts in cluster order (deterministic).
 *   d_indices:           indices_capacity entries; must be >= X*Y*Z*cap
 *                        (the size ResourcesClusterForward.cpp:10-13
 *                        allocates), else HVX_ERR_CAPACITY before launch
 *   d_total:             optional (NULL ok): receives sum(counts), the
 *                        value the reference leaves in globalIndex
 *   d_overflow_count:    optional (NULL ok): receives the number of
 *                        clusters whose list was truncated at `cap`.
 * 1 <= cap <= 1024. */
int hvx_cf_cull_lights(hvx_ctx* ctx, const hvx_cf_ubo* ubo,
                       const hvx_aabb* d_clusters,
                       const hvx_light* d_lights, uint32_t n_lights,
                       uint32_t cap,
                       hvx_light_cell* d_cells,
                       uint32_t* d_indices, size_t indices_capacity,
                       uint32_t* d_total, uint32_t* d_overflow_count,
                       void* stream);

/* hvx_cf_cull_lights without any allocation: no counterpart in the
 * reference. It is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight per-frame serving loops, hosts that forbid
 * hipMalloc). Same arguments, checks, kernels and results as
 * hvx_cf_cull_lights; its scratch (fixed-slot lists, counts, view-space
 * lights) is a fixed device array of the library, one per GPU, held until
 * the process ends, instead of the context's growable scratch. Calls whose
 * scratch (hvx_cf_cull_noalloc_scratch_bytes) exceeds
 * hvx_cf_cull_noalloc_scratch_limit() (24 MiB: e.g. 32x18x48 clusters, cap
 * 128 and 65,536 lights need 15.4 MB) are refused with HVX_ERR_CAPACITY
 * before anything is launched. Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (other streams, other contexts,
 * other host threads) race: serialise them. It never calls hipMalloc,
 * hipFree or a synchronising runtime function. */
int hvx_cf_cull_lights_noalloc(hvx_ctx* ctx, const hvx_cf_ubo* ubo,
                               const hvx_aabb* d_clusters,
                               const hvx_light* d_lights, uint32_t n_lights,
                               uint32_t cap,
                               hvx_light_cell* d_cells,
                               uint32_t* d_indices, size_t indices_capacity,
                               uint32_t* d_total, uint32_t* d_overflow_count,
                               void* stream);
/* Bytes of fixed scratch a hvx_cf_cull_lights_noalloc call needs, and the
 * size of the fixed scratch (the limit). */
size_t hvx_cf_cull_noalloc_scratch_bytes(uint32_t clusters, uint32_t cap, uint32_t n_lights);
size_t hvx_cf_cull_noalloc_scratch_limit(void);

/* The batched variant of the culling shader, LightCullingBatch.comp:61-106
 * (the reference builds it but dispatches LightCulling.comp instead,
 * PipelineLightCulling.cpp:25-26, :80-83): one invocation per cluster,
 * cluster index = local index + 576 * workgroup (with the reference's
 * 16x9x4 workgroups over a 16x9xZ grid this is x + y*X + z*X*Y, the index
 * of every other stage), and the lights staged through shared memory in
 * batches of 576, each invocation testing its cluster against the whole
 * batch (strict d^2 < r^2, as LightCulling.comp). Offsets are the exclusive
 * prefix sum of the counts (the reference: atomicAdd).
 * flags == 0 fixes the shader's two defects: the list stops at `cap`
 * and entries of the last batch past n_lights are not tested. The lists are
 * then those of hvx_cf_cull_lights, bit for bit. The flags restore the
 * reference's behaviour for comparison (such lists must not be shaded):
 *   HVX_CULL_BATCH_REFERENCE_CAP:  no cap (:86-92 writes past the 150-entry
 *       array): counts are not limited; entries whose packed position is
 *       >= indices_capacity are not written, and *d_total receives the
 *       unlimited sum, so a caller sees when its buffer was too small;
 *   HVX_CULL_BATCH_REFERENCE_TAIL: the last batch's slots past n_lights keep
 *       the previous batch's lights (never written, :79-82) and are tested,
 *       emitting indices >= n_lights; slots never written hold zeros (the
 *       shader leaves them undefined), whose radius 0 never passes the test.
 * d_overflow_count is not produced; other arguments as hvx_cf_cull_lights. */
#define HVX_CULL_BATCH_REFERENCE_CAP 1u
#define HVX_CULL_BATCH_REFERENCE_TAIL 2u
int hvx_cf_cull_lights_batch(hvx_ctx* ctx, const hvx_cf_ubo* ubo,
                             const hvx_aabb* d_clusters,
                             const hvx_light* d_lights, uint32_t n_lights,
                             uint32_t cap, uint32_t flags,
                             hvx_light_cell* d_cells,
                             uint32_t* d_indices, size_t indices_capacity,
                             uint32_t* d_total, void* stream);

/* Clustered Cook-Torrance + split-sum IBL shading of SoA fragments.
 * Replaces PipelinePBRClusterForward::FillCommandBuffer
 * (PipelinePBRClusterForward.cpp:77-113) running
 * Shaders/ClusteredForward/Scene.frag:60-149 with Radiance.glsl:2-56 and
 * PBR/Ambient.glsl:1-29. Output is FP32 RGB, planar [3][height*width]
 * (the reference writes vec4 fragColor to an 8-bit attachment).
 * Fragments whose depth maps to slice Z (depth == far) are clamped to Z-1. */
int hvx_cf_shade(hvx_ctx* ctx, const hvx_frag_soa* frags,
                 const hvx_camera_ubo* camera, const hvx_cf_ubo* ubo,
                 const hvx_pbr_pc* pc,
                 const hvx_light* d_lights, uint32_t n_lights,
                 const hvx_light_cell* d_cells, const uint32_t* d_indices,
                 const hvx_ibl_maps* ibl, float* d_rgb, void* stream);

/* --------------------------------------------------------------------- IBL */

/* Split-sum BRDF LUT, RG32F [height][width]. Replaces
 * PipelineBRDFLUT::CreateLUT (PipelineBRDFLUT.cpp:32-108) running
 * Shaders/IBL/BRDFLUT.comp:26-71. */
int hvx_ibl_brdf_lut(hvx_ctx* ctx, uint32_t width, uint32_t height,
                     uint32_t sample_count, float* d_lut_rg, void* stream);

/* Equirectangular RGBA32F [h][w] (row 0 = texture t = 0) to cube level 0
 * (6 faces of face_size^2, RGBA32F). Replaces
 * PipelineEquirect2Cube::OffscreenRender (PipelineEquirect2Cube.cpp:232-285)
 * running Shaders/IBL/Equirect2Cube.frag:96-107. */
int hvx_ibl_equirect_to_cube(hvx_ctx* ctx, const float* d_equirect_rgba,
                             uint32_t w, uint32_t h, float* d_cube_rgba,
                             uint32_t face_size, void* stream);

/* Fills levels 1..levels-1 of a cube chain from level 0 with 2x2 box means
 * (the 2:1 LINEAR blit chain of VulkanImage::GenerateMipmap,
 * VulkanImage.cpp:667-786). face_size must be a power of two. */
int hvx_ibl_cube_mips(hvx_ctx* ctx, float* d_cube_chain, uint32_t face_size,
                      uint32_t levels, void* stream);

/* Diffuse irradiance (type 0: out_mips must be 1, sample_count ignored —
 * the reference walks a fixed 0.025 rad grid) or GGX specular prefilter
 * (type 1: mip i uses roughness i/(out_mips-1), sample_count Hammersley
 * samples). Reads level 0 of d_cube only, bilinear with seamless edges, as
 * the reference's maxLod = 0 sampler does. Replaces
 * PipelineCubeFilter::OffscreenRender (PipelineCubeFilter.cpp:246-358)
 * running Shaders/IBL/CubeFilterDiffuse.frag:35-136 and
 * CubeFilterSpecular.frag:34-146. opts may be NULL.
 * Context scratch: the specular filter's sample table (16 B per mip and
 * sample); the diffuse filter's staged footprints, 64 B for each bilinear
 * footprint (6 (face_size+1)^2 entries, 403 MB at 1024^2; none when that
 * exceeds 1 GiB). Under stream capture, when the context holds no staging
 * buffer that large yet, the diffuse filter reads the taps from d_cube
 * instead and allocates nothing (the same results, bit for bit). */
int hvx_ibl_filter(hvx_ctx* ctx, int type, const float* d_cube,
                   uint32_t face_size, float* d_out_chain,
                   uint32_t out_face_size, uint32_t out_mips,
                   uint32_t sample_count, const hvx_filter_opts* opts,
                   void* stream);

/* ------------------------------------------------- contraction (MFMA) filter */

/* Arithmetic of the contraction filter's MFMA products. */
typedef enum hvx_precision {
    HVX_PRECISION_FP16_SCALED = 0, /* FP16 weights and radiance, power-of-two scaled per row / per face */
    HVX_PRECISION_BF16_SPLIT = 1   /* as FP16_SCALED, except the low-roughness (kind 2) mips, near and far
                                      fields: split-BF16 compensated products (hi*hi + hi*lo + lo*hi), twice
                                      the weights of those mips */
} hvx_precision;

typedef struct hvx_contract_plan hvx_contract_plan;

typedef struct hvx_contract_mip_info {
    int kind;                 /* 0 delta (roughness 0), 1 full-sphere sum, 2 near window + far field */
    uint32_t out_face;
    uint32_t in_face;         /* input level summed (kind 1) or near level (kind 2) */
    uint32_t far_face;        /* kind 2: far level and far grid */
    float window_t0, window_t1; /* kind 2: angular window radii (radians) */
    uint64_t weight_bytes;    /* precomputed weights of this mip */
    uint32_t min_input_level; /* deepest level of the input chain the plan reads */
    uint32_t _pad;
} hvx_contract_mip_info;

/* The cube filters recast as dense contractions of precomputed GGX (type 1)
 * or cosine (type 0) weight matrices with radiance texels, on FP16 MFMA with
 * FP32 accumulation; the weights' row normalisation sums are FP64, on the
 * GPU. This is a deterministic quadrature of the integral the reference's
 * filters estimate (CubeFilterSpecular.frag:34-75 by 1024 GGX samples,
 * CubeFilterDiffuse.frag:35-65 by a 252x63 grid), not a bit-compatible
 * variant of them: hvx_ibl_filter remains the drop-in. The plan builds the
 * weights for an input cube of face `face` and an output chain of face
 * `out_face` with `out_mips` levels (powers of two; diffuse: one level). A
 * plan serves any number of environments; calls sharing one plan must not
 * overlap in time (its staging buffers are reused). */
int hvx_ibl_contract_plan_create(hvx_ctx* ctx, int type, uint32_t face, uint32_t out_face, uint32_t out_mips,
                                 int precision, hvx_contract_plan** out);
int hvx_ibl_contract_plan_destroy(hvx_contract_plan* plan);
int hvx_ibl_contract_plan_query(const hvx_contract_plan* plan, uint32_t mip, hvx_contract_mip_info* info);
/* d_chain: the input cube with its box mip chain (hvx_ibl_cube_mips), at
 * least min_input_level + 1 levels; d_out: hvx_cube_chain_floats(out_face,
 * out_mips) floats. Allocates nothing; may be captured into a graph. */
int hvx_ibl_filter_contract(hvx_contract_plan* plan, const float* d_chain, uint32_t chain_levels, float* d_out,
                            void* stream);

/* Bright-texel split for single-probe calls (max_hot in [1, 4096]; 0 turns
 * it off). Level 0 is split per channel into min(L, T) + (L - T)+ at the
 * threshold T that at most max_hot texels reach (largest channel binned in
 * steps of 2^(1/8)); the contraction filters the clamped part, and the
 * excess of those texels is summed exactly over their level-0 texel centres,
 * w(n.l) dw (L - T)+, over the closed-form integral of w, and added to every
 * non-delta output texel. A sun-like source then no longer passes through
 * the box-filtered levels. The first call that enables it allocates the
 * clamped chain and the lists; filter calls still allocate nothing.
 * hvx_ibl_filter_contract_multi refuses plans with the split on. No
 * reference counterpart. */
int hvx_ibl_contract_plan_set_hot(hvx_contract_plan* plan, uint32_t max_hot);
/* Threshold and number of hot texels of the plan's last split call (waits
 * for the device). */
int hvx_ibl_contract_hot_info(hvx_contract_plan* plan, float* threshold, uint32_t* n_hot);

/* Several probes of the plan's sizes in one call (the same launches, with a
 * probe dimension): d_chains and d_outs are DEVICE arrays of n pointers
 * (inputs as for hvx_ibl_filter_contract). Each probe's result equals the
 * single-probe call's bit for bit. The plan must first reserve per-probe
 * scratch for at least n probes (hvx_ibl_contract_plan_reserve, which
 * allocates; n <= 4096); the call itself allocates nothing. No reference
 * counterpart (batches many ResourcesIBL specular prefilters). */
int hvx_ibl_contract_plan_reserve(hvx_contract_plan* plan, uint32_t max_probes);
int hvx_ibl_filter_contract_multi(hvx_contract_plan* plan, uint32_t n, const float* const* d_chains,
                                  uint32_t chain_levels, float* const* d_outs, void* stream);

/* --------------------------------------------- learned radiance (MLP) */

/* A per-environment network standing in for the prefiltered specular chain:
 * positional encoding (16 frequencies) of (direction, lod) -> 3 x 64 ReLU ->
 * RGB log-radiance (the network is defined in hellovulkan_amd/csrc/hvx_mlp.h;
 * 16771 FP32 parameters). It has no counterpart in the reference: it learns
 * the value the shading reads with textureLod(specularMap, R, lod)
 * (Ambient.glsl:1-29) for lod in [0, lod_max]. Trained with BF16 MFMA
 * forward/backward kernels and FP32 master weights (Adam). */
typedef struct hvx_mlp hvx_mlp;

int hvx_mlp_create(hvx_ctx* ctx, float lod_max, uint64_t seed, hvx_mlp** out);
int hvx_mlp_destroy(hvx_mlp* mlp);
int hvx_mlp_get_params(hvx_mlp* mlp, float* host_params /* 16771 */);
int hvx_mlp_set_params(hvx_mlp* mlp, const float* host_params);
/* n training samples from a prefiltered chain (hvx_ibl_filter type 1):
 * inputs float4 (dir.xyz, lod), targets float4 (log(rgb + 1e-3), 0). */
int hvx_mlp_sample(hvx_mlp* mlp, const float* d_chain, uint32_t face, uint32_t mips, uint64_t seed, uint32_t n,
                   float* d_inputs, float* d_targets, void* stream);
/* Gradient of the mean squared log-radiance error of one batch (a multiple
 * of 256 samples) with respect to the parameters, in their layout; *d_loss
 * (nullable) receives the loss. */
int hvx_mlp_gradient(hvx_mlp* mlp, const float* d_inputs, const float* d_targets, uint32_t batch, float* d_grads,
                     float* d_loss, void* stream);
/* `steps` Adam steps over consecutive batches of the n samples (cosine
 * learning-rate decay from lr to lr/100); d_loss_curve (nullable) receives
 * each step's loss. A call that revisits its samples (steps > n / batch)
 * encodes them once into network-owned memory (512 B per sample, kept for
 * later calls); without that memory each step encodes its batch. */
int hvx_mlp_train(hvx_mlp* mlp, const float* d_inputs, const float* d_targets, uint32_t n, uint32_t batch,
                  uint32_t steps, float lr, float* d_loss_curve, void* stream);
/* Radiance (RGB, n x 3 floats) predicted at n inputs (dir.xyz, lod), by the
 * inference path the shading kernel uses. */
int hvx_mlp_eval(hvx_mlp* mlp, const float* d_inputs, uint32_t n, float* d_rgb, void* stream);
/* Error of the network for hvx_cf_shade_learned: for each cell (bin b,
 * face f, texel y, x) of a cell_face cube and lod_bins bins over
 * [0, lod_max], d_err[(b * 6 + f) * cell_face^2 + y * cell_face + x]
 * receives twice the largest relative error (every channel) of the
 * network against the chain's textureLod value at the cell's 75 test points
 * (a 5 x 5 grid over the texel, corners included, at both bin ends and its
 * middle): a margin over the measured error, since the error between test
 * points is not measured. It also keeps, with the network, half the
 * smallest output the network gave in each cell, which hvx_cf_shade_learned
 * uses on the same cell grid to skip the network where even that output
 * could not be used. A set-up call: it allocates and synchronises the
 * stream. */
int hvx_mlp_cell_errors(hvx_mlp* mlp, const float* d_chain, uint32_t face, uint32_t mips, uint32_t cell_face,
                        uint32_t lod_bins, float* d_err, void* stream);
/* hvx_cf_shade with the prefiltered specular radiance taken from the network
 * (inferred inside the shading kernel with BF16 MFMA) wherever the error it
 * can cause in the shaded radiance stays within `tol` relative: with e the
 * cell's measured error, e/(1-e) * network value * (F brdf.x + brdf.y) must
 * not exceed tol times the fragment's radiance without the specular ambient
 * term, in every channel. Elsewhere the specular map is sampled as
 * hvx_cf_shade does (the fallback). d_served (nullable, zeroed by the
 * caller) counts the fragments that used the network. pc->max_reflection_lod
 * must not exceed the network's lod_max. The cell error is measured at test
 * points, so it estimates, not bounds, the error between them. */
int hvx_cf_shade_learned(hvx_ctx* ctx, const hvx_frag_soa* frags, const hvx_camera_ubo* camera, const hvx_cf_ubo* ubo,
                         const hvx_pbr_pc* pc, const hvx_light* d_lights, uint32_t n_lights,
                         const hvx_light_cell* d_cells, const uint32_t* d_indices, const hvx_ibl_maps* ibl,
                         const hvx_mlp* mlp, const float* d_cell_err, uint32_t cell_face, uint32_t lod_bins,
                         float tol, float* d_rgb, uint32_t* d_served, void* stream);

/* ---------------------------------------------------------------- batches */

/* One light probe of a prefilter batch: its level-0 cube and where its GGX
 * prefiltered chain goes (as hvx_ibl_filter type 1 writes it). */
typedef struct hvx_probe {
    const float* cube;         /* RGBA32F, 6 x face_size^2 (level 0 only is read) */
    uint32_t face_size;
    uint32_t out_face_size;
    uint32_t out_mips;
    uint32_t _pad;
    float* out_chain;          /* hvx_cube_chain_floats(out_face_size, out_mips) */
} hvx_probe;

/* One clustered-forward view of a culling batch (hvx_cf_cull_lights
 * arguments). Cluster AABBs must already be built (hvx_cf_build_aabbs). */
typedef struct hvx_view {
    hvx_cf_ubo ubo;
    const hvx_aabb* clusters;
    const hvx_light* lights;
    uint32_t n_lights;
    uint32_t cap;
    hvx_light_cell* cells;
    uint32_t* indices;
    size_t indices_capacity;   /* >= X*Y*Z*cap */
    uint32_t* total;           /* nullable */
} hvx_view;

/* Tile shapes of the batched kernels. hvx_batch_autotune fills the timings. */
typedef struct hvx_batch_tuning {
    uint32_t probe_tile;        /* output texels per probe work item, 32..8192 */
    uint32_t probe_wgs_per_cu;  /* persistent probe workgroups per CU, 1..16 */
    uint32_t cull_shape;        /* cluster block: 0 = 4x4x1, 1 = 8x4x1, 2 = 4x4x2, 3 = 2x2x4 */
    uint32_t cull_wgs_per_cu;   /* persistent culling workgroups per CU, 1..16 */
    float probe_ms;             /* measured by hvx_batch_autotune for the chosen shape */
    float cull_ms;
} hvx_batch_tuning;

typedef struct hvx_batch hvx_batch;

/* A batch of probes (GGX prefilter with sample_count samples, as
 * hvx_ibl_filter type 1) and views (light culling, as hvx_cf_cull_lights),
 * run by two persistent kernels that pull work items from device queues.
 * There is no single reference entry point: it batches what
 * ResourcesIBL::Create (ResourcesIBL.cpp:18-40) does for one probe and what
 * PipelineLightCulling::FillCommandBuffer (PipelineLightCulling.cpp:51-119)
 * does for one view. Every output equals the per-item entry point's output
 * bit for bit. The arrays are copied; all scratch is allocated here, so
 * hvx_batch_run allocates nothing and may be captured into a graph. */
int hvx_batch_create(hvx_ctx* ctx, const hvx_probe* probes, uint32_t n_probes, uint32_t sample_count,
                     const hvx_view* views, uint32_t n_views, hvx_batch** out);
int hvx_batch_destroy(hvx_batch* batch);
int hvx_batch_set_tuning(hvx_batch* batch, const hvx_batch_tuning* tuning);
int hvx_batch_get_tuning(const hvx_batch* batch, hvx_batch_tuning* tuning);
/* Times every candidate tile shape on this batch (synchronously, on
 * `stream`; outputs are overwritten) and keeps the fastest. */
int hvx_batch_autotune(hvx_batch* batch, void* stream);
/* Contraction mode (enable != 0): the probes are filtered by the MFMA
 * contraction (hvx_ibl_filter_contract, `precision` an hvx_precision) instead
 * of the sampled GGX prefilter: one plan, scratch input chain and stream per
 * distinct (face_size, out_face_size, out_mips); per probe, level 0 is copied
 * into the chain, its box mips are built and the plan filters it. Outputs then
 * equal the per-probe contraction filter's bit for bit. Allocates the plans
 * and scratch here; enable = 0 returns to the sampled prefilter. The mode runs
 * through hvx_batch_run only: hvx_batch_capture returns HVX_ERR_UNSUPPORTED
 * while it is enabled. No reference counterpart (it composes two additions of
 * this library). */
int hvx_batch_set_contract(hvx_batch* batch, int enable, int precision);
/* Runs the batch: the probe and culling kernels on forked streams, joined
 * back into `stream` before returning control to it. */
int hvx_batch_run(hvx_batch* batch, void* stream);
/* Captures hvx_batch_run into a graph (re-captured after a tuning change);
 * hvx_batch_launch replays it on `stream`. */
int hvx_batch_capture(hvx_batch* batch, void* stream);
int hvx_batch_launch(hvx_batch* batch, void* stream);

/* ------------------------------------------------------------- tonemap */

typedef enum hvx_tonemap_op {
    HVX_TONEMAP_REINHARD = 0,   /* Tonemap.frag main(): Reinhard, pure white 1.0 */
    HVX_TONEMAP_ACES = 1,
    HVX_TONEMAP_FILMIC = 2,
    HVX_TONEMAP_UNCHARTED2 = 3
} hvx_tonemap_op;

/* PipelineTonemap::FillCommandBuffer (PipelineTonemap.cpp:34-53) running
 * Tonemap.frag: maps the planar FP32 RGB shading output (width*height pixels
 * per plane) with curve `op`, then gamma 2.2 when `gamma` is set (main()
 * uses HVX_TONEMAP_REINHARD with gamma). Writes planar FP32 RGB to d_out_rgb
 * and/or packed B8G8R8A8_UNORM (round to nearest even, NaN -> 0, alpha 255)
 * to d_out_bgra8; either may be NULL, not both. */
int hvx_post_tonemap(hvx_ctx* ctx, const float* d_rgb, uint32_t width, uint32_t height, int op, int gamma,
                     float* d_out_rgb, uint32_t* d_out_bgra8, void* stream);

/* ------------------------------------------------------------ HDR input */

/* Decodes a Radiance RGBE (.hdr) file held in memory into RGBA32F (alpha 1),
 * as VulkanImage::CreateFromHDR loads the environment
 * (VulkanImage.cpp:120-146: stbi_loadf with flip-on-load). With rgba NULL it
 * only reads width and height; otherwise rgba receives width * height * 4
 * floats, row 0 = the file's last scanline when flip_vertically is set.
 * HVX_ERR_UNSUPPORTED for other pixel formats or orientations,
 * HVX_ERR_INVALID_ARG for malformed or truncated data. Host memory only. */
int hvx_hdr_decode(const uint8_t* data, size_t size, int flip_vertically, uint32_t* width, uint32_t* height,
                   float* rgba);

/* ------------------------------------------------------------------ camera */

/* Camera description: position/target, vertical fov (radians), clip planes,
 * viewport and cluster grid. Mirrors Camera::SetPositionAndTarget +
 * UpdateInternal (HelloVulkan/Source/Camera.cpp:78-139) with worldUp =
 * (0,1,0), GLM_FORCE_DEPTH_ZERO_TO_ONE and the Vulkan Y flip (P[1][1] *= -1),
 * plus GetCameraUBO / GetClusterForwardUBO (:160-197). */
typedef struct hvx_camera_desc {
    float eye[3];
    float target[3];
    float fovy;
    float znear;
    float zfar;
    uint32_t width;
    uint32_t height;
    uint32_t slice_count_x;
    uint32_t slice_count_y;
    uint32_t slice_count_z;
} hvx_camera_desc;

/* Fills the CameraUBO, the ClusterForwardUBO and inverse(P * V). Matrices are
 * formed in double precision and rounded once to float (the reference forms
 * them in float with glm; they are inputs of the path, shared by every
 * stage). Any output pointer may be NULL. */
int hvx_camera_build(const hvx_camera_desc* desc, hvx_camera_ubo* camera,
                     hvx_cf_ubo* cf, float inv_view_proj[16]);

/* Size helpers (float counts) for cube chains. */
size_t hvx_cube_chain_floats(uint32_t face_size, uint32_t mips);
size_t hvx_cube_level_offset_floats(uint32_t face_size, uint32_t level);
/* Utility::MipMapCount (HelloVulkan/Header/Utility.h:42-50). */
uint32_t hvx_mip_count(uint32_t size);

#ifdef __cplusplus
}
#endif

#endif /* HVX_H_ */
