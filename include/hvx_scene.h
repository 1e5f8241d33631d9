/*
 * hvx_scene.h — seeded synthetic scenes for the clustered-forward + IBL path.
 *
 * The reference feeds its stages from an application layer (lights from
 * AppPBRClusterForward::InitLights, HelloVulkan/Source/Apps/
 * AppPBRClusterForward.cpp:81-118; rasterised Sponza fragments; an HDR file
 * decoded by stb_image). None of those assets exist here, so these host-side
 * generators produce the inputs instead, deterministically from a 64-bit seed
 * (SplitMix64 counter hashing: element i of stream k draws
 * splitmix64(seed + (k << 40) + i), so any sub-range can be regenerated
 * independently and in parallel). All buffers are host memory, caller-owned.
 */
#ifndef HVX_SCENE_H_
#define HVX_SCENE_H_

#include <stdint.h>

#include "hvx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* n point lights: position U([-15,15] x [0,12] x [-7,7]) with w = 1, radius
 * U[0.5, 3], colour U[0,1]^3 with w = 1. */
int hvx_scene_lights(uint64_t seed, uint32_t n, hvx_light* out);

/* One fragment per pixel of a width x height view through `camera`:
 * linear depth log-uniform in [near, far) stored as NDC depth, unit normal
 * uniform on the hemisphere facing the camera, albedo U[0,1]^3, roughness
 * U[0.05, 1], metallic Bernoulli(0.3) as UNORM8 {0, 255}. Planar outputs of
 * width*height elements each; inv_view_proj as from hvx_camera_build. */
int hvx_scene_fragments(uint64_t seed, const hvx_camera_desc* camera, const float inv_view_proj[16],
                        float* depth, float* nx, float* ny, float* nz, float* ar, float* ag, float* ab,
                        float* roughness, uint8_t* metallic, int threads);

/* Procedural HDR equirect, RGBA32F [h][w] in texture order (row 0 at t = 0,
 * direction y = cos(pi * (1 - t))): sky gradient over elevation, a sun disk of
 * radiance 1e4, and 32 hot spots with log-normal peak radiance and Gaussian
 * angular falloff. */
int hvx_scene_equirect(uint64_t seed, uint32_t w, uint32_t h, float* rgba, int threads);

#ifdef __cplusplus
}
#endif

#endif /* HVX_SCENE_H_ */
