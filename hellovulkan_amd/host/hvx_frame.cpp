// hvx_frame — C++20 frame driver of the clustered-forward + IBL path.
//
// The role of AppPBRClusterForward (HelloVulkan/Source/Apps/
// AppPBRClusterForward.cpp:19-136) for this port: it builds a seeded scene
// (lights, SoA fragments, procedural environment), runs the IBL stages once
// (ResourcesIBL), then per frame updates the UBOs and runs light culling and
// clustered shading — first with eager launches, then replayed from one
// captured hipGraph — and reports per-frame times as one JSON line.
//
//   hvx_frame [--width W --height H --grid X Y Z --lights N --cap C]
//             [--cube S --diffuse S --specular S --mips M --lut S]
//             [--frames K --warmup W] [--dump PREFIX]
//             [--contract PRECISION] [--tonemap OP [--gamma 0|1]] [--orbit K]
//
// --orbit K then moves the camera for K more frames (AppPBRClusterForward::
// UpdateUBOs pushes new camera and cluster UBOs every frame): the eye circles
// the target, each frame runs eagerly and from the same executable graph,
// updated in place by FrameGraph::Capture, and the two frames must be equal.
//
// --contract builds the specular chain with the MFMA contraction filter
// (hvx_precision) instead of the sampled prefilter; --tonemap adds the post
// step (hvx_tonemap_op) after shading, eagerly and in the graph.
// --dump writes PREFIX.cells (hvx_light_cell[C]), PREFIX.indices (u32, the
// packed total), PREFIX.rgb (float[3][W*H], eager frame), PREFIX.rgb_graph
// (graph replay), PREFIX.spec (the specular chain) and with --tonemap
// PREFIX.bgra8 / PREFIX.bgra8_graph, for comparison against the C ABI
// driven from Python.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hvx.h"
#include "hvx/pipelines.hpp"
#include "hvx_scene.h"

namespace {

struct Args {
    uint32_t width = 3840, height = 2160, gx = 32, gy = 18, gz = 48, lights = 65536, cap = 128;
    uint32_t cube = 1024, diffuse = 32, specular = 512, mips = 10, lut = 512;
    uint32_t frames = 20, warmup = 3, orbit = 0;
    uint64_t light_seed = 1001, frag_seed = 1002, env_seed = 77;
    int contract = -1, tonemap = -1, gamma = 1;
    std::string dump;
};

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        auto next = [&]() -> uint32_t {
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + k);
            return (uint32_t)std::strtoul(argv[++i], nullptr, 10);
        };
        if (k == "--width") a.width = next();
        else if (k == "--height") a.height = next();
        else if (k == "--grid") { a.gx = next(); a.gy = next(); a.gz = next(); }
        else if (k == "--lights") a.lights = next();
        else if (k == "--cap") a.cap = next();
        else if (k == "--cube") a.cube = next();
        else if (k == "--diffuse") a.diffuse = next();
        else if (k == "--specular") a.specular = next();
        else if (k == "--mips") a.mips = next();
        else if (k == "--lut") a.lut = next();
        else if (k == "--frames") a.frames = next();
        else if (k == "--warmup") a.warmup = next();
        else if (k == "--orbit") a.orbit = next();
        else if (k == "--light-seed") a.light_seed = next();
        else if (k == "--frag-seed") a.frag_seed = next();
        else if (k == "--dump") { if (i + 1 >= argc) throw std::runtime_error("--dump PREFIX"); a.dump = argv[++i]; }
        else if (k == "--contract") a.contract = (int)next();
        else if (k == "--tonemap") a.tonemap = (int)next();
        else if (k == "--gamma") a.gamma = (int)next();
        else throw std::runtime_error("unknown argument " + k);
    }
    return a;
}

template <class T> void write_file(const std::string& path, const std::vector<T>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    hvx::hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    return ms;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Args a = parse(argc, argv);
        hvx::Context ctx(0);
        hipStream_t stream;
        hvx::hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");

        // camera and UBOs (Camera::GetCameraUBO / GetClusterForwardUBO)
        hvx_camera_desc desc{};
        const float eye[3] = {0.0f, 1.0f, 6.0f}, target[3] = {0.0f, 2.5f, 0.0f};
        std::memcpy(desc.eye, eye, sizeof(eye));
        std::memcpy(desc.target, target, sizeof(target));
        desc.fovy = 0.785398f;
        desc.znear = 0.1f;
        desc.zfar = 100.0f;
        desc.width = a.width;
        desc.height = a.height;
        desc.slice_count_x = a.gx;
        desc.slice_count_y = a.gy;
        desc.slice_count_z = a.gz;
        hvx_camera_ubo cam{};
        hvx_cf_ubo ubo{};
        float ivp[16];
        ctx.check(hvx_camera_build(&desc, &cam, &ubo, ivp), "hvx_camera_build");

        // lights (InitLights) and fragments
        std::vector<hvx_light> lights(a.lights);
        ctx.check(hvx_scene_lights(a.light_seed, a.lights, lights.data()), "hvx_scene_lights");
        hvx::ResourcesLight resLight;
        resLight.AddLights(lights);
        const size_t F = (size_t)a.width * a.height;
        std::vector<float> planes[8];
        for (auto& p : planes) p.resize(F);
        std::vector<uint8_t> metal(F);
        ctx.check(hvx_scene_fragments(a.frag_seed, &desc, ivp, planes[0].data(), planes[1].data(), planes[2].data(),
                                      planes[3].data(), planes[4].data(), planes[5].data(), planes[6].data(),
                                      planes[7].data(), metal.data(), 16),
                  "hvx_scene_fragments");
        hvx::FragmentsSoA frags;
        frags.Create(a.width, a.height, ivp);
        frags.depth_.upload(planes[0].data(), F);
        for (int c = 0; c < 3; ++c) {
            frags.normal_[c].upload(planes[1 + c].data(), F);
            frags.albedo_[c].upload(planes[4 + c].data(), F);
        }
        frags.roughness_.upload(planes[7].data(), F);
        frags.metallic_.upload(metal.data(), F);

        // image-based lighting resources (ResourcesIBL::Create)
        const uint32_t ew = 2 * a.cube, eh = a.cube;
        std::vector<float> eq((size_t)ew * eh * 4);
        ctx.check(hvx_scene_equirect(a.env_seed, ew, eh, eq.data(), 16), "hvx_scene_equirect");
        hvx::DeviceBuffer<float> d_eq(eq.size());
        d_eq.upload(eq.data(), eq.size());
        hvx::ResourcesIBL ibl;
        ibl.cubeSize_ = a.cube;
        ibl.diffuseSize_ = a.diffuse;
        ibl.specularSize_ = a.specular;
        ibl.specularMips_ = a.mips;
        ibl.lutWidth_ = ibl.lutHeight_ = a.lut;
        const auto t_ibl0 = std::chrono::steady_clock::now();
        hvx::CreateIBL(ctx, ibl, d_eq.data(), ew, eh, stream);
        if (a.contract >= 0) {  // the specular chain through the contraction filter instead
            hvx::PipelineCubeFilterContract filter(ctx, &ibl.environmentCubemap_, ibl.cubeSize_, stream);
            filter.OffscreenRender(ibl.specularCubemap_, hvx::CubeFilterType::Specular, ibl.specularSize_,
                                   ibl.specularMips_, (hvx_precision)a.contract, stream);
            hvx::hip_check(hipStreamSynchronize(stream), "contraction sync");
        }
        hvx::hip_check(hipStreamSynchronize(stream), "ibl sync");
        const double ibl_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_ibl0).count();

        // stages (AppPBRClusterForward::Init) and UBOs (UpdateUBOs)
        hvx::ResourcesClusterForward resCF;
        resCF.CreateBuffers(a.gx, a.gy, a.gz, a.cap);
        hvx::DeviceBuffer<float> rgb(3 * F), rgb_graph(3 * F);
        hvx::DeviceBuffer<uint32_t> bgra(a.tonemap >= 0 ? F : 0), bgra_graph(a.tonemap >= 0 ? F : 0);
        hvx::PipelineAABBGenerator aabb(ctx, &resCF);
        hvx::PipelineLightCulling cull(ctx, &resLight, &resCF);
        hvx::PipelinePBRClusterForward pbr(ctx, &frags, &resLight, &resCF, &ibl, &rgb);
        hvx::PipelineTonemap tone(ctx, &rgb, a.width, a.height, &bgra), tone_g(ctx, &rgb_graph, a.width, a.height,
                                                                               &bgra_graph);
        if (a.tonemap >= 0) {
            tone.SetOperator((hvx_tonemap_op)a.tonemap, a.gamma != 0);
            tone_g.SetOperator((hvx_tonemap_op)a.tonemap, a.gamma != 0);
        }
        aabb.SetClusterForwardUBO(ubo);
        cull.ResetGlobalIndex();
        cull.SetClusterForwardUBO(ubo);
        pbr.SetClusterForwardUBO(ubo);
        pbr.SetCameraUBO(cam);
        aabb.FillCommandBuffer(stream);  // dirty at creation

        hipEvent_t e0, e1;
        hvx::hip_check(hipEventCreate(&e0), "hipEventCreate");
        hvx::hip_check(hipEventCreate(&e1), "hipEventCreate");
        // eager frames
        std::vector<hvx::PipelineBase*> eager = {&cull, &pbr};
        if (a.tonemap >= 0) eager.push_back(&tone);
        for (uint32_t i = 0; i < a.warmup; ++i)
            for (hvx::PipelineBase* p : eager) p->FillCommandBuffer(stream);
        hvx::hip_check(hipEventRecord(e0, stream), "hipEventRecord");
        for (uint32_t i = 0; i < a.frames; ++i)
            for (hvx::PipelineBase* p : eager) p->FillCommandBuffer(stream);
        hvx::hip_check(hipEventRecord(e1, stream), "hipEventRecord");
        hvx::hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        const float eager_ms = elapsed_ms(e0, e1) / (float)a.frames;

        // the same frame replayed from one captured graph, into a second output
        hvx::PipelinePBRClusterForward pbr_g(ctx, &frags, &resLight, &resCF, &ibl, &rgb_graph);
        pbr_g.SetClusterForwardUBO(ubo);
        pbr_g.SetCameraUBO(cam);
        std::vector<hvx::PipelineBase*> staged = {&cull, &pbr_g};
        if (a.tonemap >= 0) staged.push_back(&tone_g);
        hvx::FrameGraph graph(staged);
        graph.Capture();
        for (uint32_t i = 0; i < a.warmup; ++i) graph.Launch();
        hvx::hip_check(hipEventRecord(e0, graph.stream()), "hipEventRecord");
        for (uint32_t i = 0; i < a.frames; ++i) graph.Launch();
        hvx::hip_check(hipEventRecord(e1, graph.stream()), "hipEventRecord");
        hvx::hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        const float graph_ms = elapsed_ms(e0, e1) / (float)a.frames;
        hvx::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");

        // moving camera: new UBOs per frame through the same executable graph
        std::vector<float> o_eager(3 * F), o_graph(3 * F);
        bool orbit_match = true;
        std::vector<double> orbit_sums;
        for (uint32_t k = 0; k < a.orbit; ++k) {
            const float ang = 0.35f * (float)(k + 1), radius = 5.0f;
            hvx_camera_desc d = desc;
            d.eye[0] = target[0] + radius * std::sin(ang);
            d.eye[1] = eye[1] + 0.25f * (float)k;
            d.eye[2] = target[2] + radius * std::cos(ang);
            hvx_camera_ubo c2{};
            hvx_cf_ubo u2{};
            float ivp2[16];
            ctx.check(hvx_camera_build(&d, &c2, &u2, ivp2), "hvx_camera_build");
            frags.SetInvViewProj(ivp2);
            cull.SetClusterForwardUBO(u2);
            for (hvx::PipelinePBRClusterForward* p : {&pbr, &pbr_g}) {
                p->SetClusterForwardUBO(u2);
                p->SetCameraUBO(c2);
            }
            for (hvx::PipelineBase* p : eager) p->FillCommandBuffer(stream);
            hvx::hip_check(hipStreamSynchronize(stream), "orbit eager sync");
            graph.Capture();
            graph.Launch();
            hvx::hip_check(hipStreamSynchronize(graph.stream()), "orbit graph sync");
            rgb.download(o_eager.data(), o_eager.size());
            rgb_graph.download(o_graph.data(), o_graph.size());
            orbit_match = orbit_match && std::memcmp(o_eager.data(), o_graph.data(), o_eager.size() * 4) == 0;
            double sum = 0.0;
            for (float v : o_eager) sum += v;
            orbit_sums.push_back(sum);
        }
        std::string sums = "[";
        for (size_t k = 0; k < orbit_sums.size(); ++k) {
            char b[32];
            std::snprintf(b, sizeof(b), "%s%.6e", k ? ", " : "", orbit_sums[k]);
            sums += b;
        }
        sums += "]";
        if (a.orbit) {  // back to the first camera for the outputs below
            frags.SetInvViewProj(ivp);
            cull.SetClusterForwardUBO(ubo);
            for (hvx::PipelinePBRClusterForward* p : {&pbr, &pbr_g}) {
                p->SetClusterForwardUBO(ubo);
                p->SetCameraUBO(cam);
            }
            for (hvx::PipelineBase* p : eager) p->FillCommandBuffer(stream);
            hvx::hip_check(hipStreamSynchronize(stream), "eager sync");
            graph.Capture();
            graph.Launch();
            hvx::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        }

        uint32_t total = 0;
        resCF.globalIndex_.download(&total, 1);
        std::vector<float> h_rgb(3 * F), h_rgb_g(3 * F);
        rgb.download(h_rgb.data(), h_rgb.size());
        rgb_graph.download(h_rgb_g.data(), h_rgb_g.size());
        bool graph_matches = std::memcmp(h_rgb.data(), h_rgb_g.data(), h_rgb.size() * sizeof(float)) == 0;
        std::vector<uint32_t> h_bgra(bgra.size()), h_bgra_g(bgra_graph.size());
        if (a.tonemap >= 0) {
            bgra.download(h_bgra.data(), F);
            bgra_graph.download(h_bgra_g.data(), F);
            graph_matches = graph_matches && h_bgra == h_bgra_g;
        }
        double rgb_sum = 0.0;
        for (float v : h_rgb) rgb_sum += v;
        if (!a.dump.empty()) {
            std::vector<hvx_light_cell> cells(resCF.clusters());
            resCF.lightCellsBuffer_.download(cells.data(), cells.size());
            std::vector<uint32_t> idx(total);
            if (total) resCF.lightIndicesBuffer_.download(idx.data(), total);
            write_file(a.dump + ".cells", cells);
            write_file(a.dump + ".indices", idx);
            write_file(a.dump + ".rgb", h_rgb);
            write_file(a.dump + ".rgb_graph", h_rgb_g);
            std::vector<float> spec(ibl.specularCubemap_.size());
            ibl.specularCubemap_.download(spec.data(), spec.size());
            write_file(a.dump + ".spec", spec);
            if (a.tonemap >= 0) {
                write_file(a.dump + ".bgra8", h_bgra);
                write_file(a.dump + ".bgra8_graph", h_bgra_g);
            }
        }
        std::printf(
            "{\"eager_ms_per_frame\": %.4f, \"graph_ms_per_frame\": %.4f, \"graph_nodes\": %zu, "
            "\"graph_matches_eager\": %s, \"mfrag_per_s_graph\": %.2f, \"ibl_build_ms\": %.2f, \"list_entries\": %u, "
            "\"rgb_sum\": %.6e, \"fragments\": %zu, \"contract\": %d, \"tonemap\": %d, \"orbit_frames\": %u, "
            "\"orbit_frames_match\": %s, \"orbit_rgb_sums\": %s, \"graph_updates_in_place\": %zu}\n",
            eager_ms, graph_ms, graph.node_count(), graph_matches ? "true" : "false",
            (double)F / (graph_ms * 1e-3) / 1e6, ibl_ms, total, rgb_sum, F, a.contract, a.tonemap, a.orbit,
            orbit_match ? "true" : "false", sums.c_str(), graph.update_count());
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(stream);
        return graph_matches && orbit_match ? 0 : 3;
    } catch (const hvx::Error& e) {
        std::fprintf(stderr, "hvx_frame: %s\n", e.what());
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "hvx_frame: %s\n", e.what());
        return 1;
    }
}
