// hvx/pipelines.hpp — C++20 host-side stage classes over the C ABI of hvx.h.
//
// They mirror the reference's pipeline-stage plugins and resource holders
// (HelloVulkan/Header/Pipelines/PipelineBase.h:24-95 and the stage headers,
// HelloVulkan/Header/Resources/Resources{Light,ClusterForward,IBL}.h) with the
// same names, constructor arguments and per-frame methods, so the
// clustered-forward app's frame (AppPBRClusterForward::UpdateUBOs +
// AppBase::FillCommandBuffer) reads the same on this port. Differences: a
// HIP stream replaces the VkCommandBuffer, device buffers are hipMalloc'd,
// and errors raise hvx::Error (the C ABI itself never throws).
// FrameGraph captures a sequence of stages into one hipGraph and replays it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hvx.h"

namespace hvx {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void hip_check(hipError_t e, const char* where) {
    if (e != hipSuccess) throw Error(HVX_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

// Owns an hvx_ctx (the VulkanContext the reference's stages are built with).
class Context {
public:
    explicit Context(int device = 0) {
        const int st = hvx_ctx_create(device, &ctx_);
        if (st) throw Error(st, std::string("hvx_ctx_create: ") + hvx_status_str(st));
    }
    ~Context() {
        if (ctx_) hvx_ctx_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    hvx_ctx* get() const { return ctx_; }
    void check(int st, const char* where) const {
        if (st) throw Error(st, std::string(where) + ": " + hvx_status_str(st) + ": " + hvx_ctx_last_error(ctx_));
    }

private:
    hvx_ctx* ctx_ = nullptr;
};

// RAII device buffer (VulkanBuffer's role).
template <class T> class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
    void resize(size_t n) {
        release();
        if (n) hip_check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void upload(const T* host, size_t n, size_t offset = 0) {
        hip_check(hipMemcpy(p_ + offset, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void download(T* host, size_t n, size_t offset = 0) const {
        hip_check(hipMemcpy(host, p_ + offset, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
    T* data() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// ---------------------------------------------------------------- resources

// ResourcesLight (HelloVulkan/Header/Resources/ResourcesLight.h:41-81).
class ResourcesLight {
public:
    void AddLights(const std::vector<hvx_light>& lights) {
        lights_ = lights;
        buffer_.resize(lights_.size() ? lights_.size() : 1);
        if (!lights_.empty()) buffer_.upload(lights_.data(), lights_.size());
    }
    // ResourcesLight::UpdateLightPosition: moves one light (world space)
    void UpdateLightPosition(uint32_t index, const float position[3]) {
        if (index >= lights_.size()) throw Error(HVX_ERR_INVALID_ARG, "UpdateLightPosition: index out of range");
        std::memcpy(lights_[index].position, position, 3 * sizeof(float));
        buffer_.upload(&lights_[index], 1, index);
    }
    uint32_t GetLightCount() const { return (uint32_t)lights_.size(); }
    const hvx_light* device() const { return buffer_.data(); }

private:
    std::vector<hvx_light> lights_;
    DeviceBuffer<hvx_light> buffer_;
};

// ResourcesClusterForward (ResourcesClusterForward.h / .cpp:5-46): AABBs,
// light cells, light indices sized cap * clusters, and the list total that
// plays globalIndex's role.
class ResourcesClusterForward {
public:
    void CreateBuffers(uint32_t sliceX, uint32_t sliceY, uint32_t sliceZ, uint32_t maxLightsPerCluster) {
        clusters_ = sliceX * sliceY * sliceZ;
        cap_ = maxLightsPerCluster;
        aabbBuffer_.resize(clusters_);
        lightCellsBuffer_.resize(clusters_);
        lightIndicesBuffer_.resize((size_t)clusters_ * cap_);
        globalIndex_.resize(1);
        aabbDirty_ = true;
    }
    uint32_t clusters() const { return clusters_; }
    uint32_t cap() const { return cap_; }

    bool aabbDirty_ = true;
    DeviceBuffer<hvx_aabb> aabbBuffer_;
    DeviceBuffer<hvx_light_cell> lightCellsBuffer_;
    DeviceBuffer<uint32_t> lightIndicesBuffer_;
    DeviceBuffer<uint32_t> globalIndex_;

private:
    uint32_t clusters_ = 0, cap_ = 0;
};

// Struct-of-arrays fragment planes (what the rasterised G-buffer provides).
class FragmentsSoA {
public:
    void Create(uint32_t width, uint32_t height, const float invViewProj[16]) {
        width_ = width;
        height_ = height;
        const size_t n = (size_t)width * height;
        depth_.resize(n);
        for (auto& p : normal_) p.resize(n);
        for (auto& p : albedo_) p.resize(n);
        roughness_.resize(n);
        metallic_.resize(n);
        std::memcpy(ivp_, invViewProj, sizeof(ivp_));
    }
    // a new camera over the same G-buffer planes (world positions are rebuilt
    // from pixel centre and depth with this matrix)
    void SetInvViewProj(const float invViewProj[16]) { std::memcpy(ivp_, invViewProj, sizeof(ivp_)); }
    hvx_frag_soa view() const {
        hvx_frag_soa f{};
        f.depth = depth_.data();
        for (int c = 0; c < 3; ++c) {
            f.normal[c] = normal_[c].data();
            f.albedo[c] = albedo_[c].data();
        }
        f.roughness = roughness_.data();
        f.metallic = metallic_.data();
        f.width = width_;
        f.height = height_;
        std::memcpy(f.inv_view_proj, ivp_, sizeof(ivp_));
        return f;
    }
    uint32_t width() const { return width_; }
    uint32_t height() const { return height_; }

    DeviceBuffer<float> depth_, normal_[3], albedo_[3], roughness_;
    DeviceBuffer<uint8_t> metallic_;

private:
    uint32_t width_ = 0, height_ = 0;
    float ivp_[16] = {};
};

// ------------------------------------------------------------------ stages

// PipelineBase (PipelineBase.h:24-95): FillCommandBuffer records the stage's
// work on a stream instead of a VkCommandBuffer.
class PipelineBase {
public:
    explicit PipelineBase(Context& ctx) : ctx_(ctx) {}
    virtual ~PipelineBase() = default;
    virtual void FillCommandBuffer(hipStream_t stream) = 0;
    virtual void OnWindowResized() {}
    virtual void SetCameraUBO(const hvx_camera_ubo& ubo) { cameraUBO_ = ubo; }
    Context& context() const { return ctx_; }

protected:
    Context& ctx_;
    hvx_camera_ubo cameraUBO_{};
};

// PipelineAABBGenerator (PipelineAABBGenerator.h:17-28, .cpp:36-83): runs only
// when the cluster resources are dirty (created or resized).
class PipelineAABBGenerator final : public PipelineBase {
public:
    PipelineAABBGenerator(Context& ctx, ResourcesClusterForward* cf) : PipelineBase(ctx), cf_(cf) {}
    void SetClusterForwardUBO(const hvx_cf_ubo& ubo) { ubo_ = ubo; }
    void OnWindowResized() override { cf_->aabbDirty_ = true; }
    void FillCommandBuffer(hipStream_t stream) override {
        if (!cf_->aabbDirty_) return;
        ctx_.check(hvx_cf_build_aabbs(ctx_.get(), &ubo_, cf_->aabbBuffer_.data(), stream), "hvx_cf_build_aabbs");
        cf_->aabbDirty_ = false;
    }

private:
    ResourcesClusterForward* cf_;
    hvx_cf_ubo ubo_{};
};

// PipelineLightCulling (PipelineLightCulling.h:18-24, .cpp:38-119).
class PipelineLightCulling final : public PipelineBase {
public:
    PipelineLightCulling(Context& ctx, ResourcesLight* lights, ResourcesClusterForward* cf)
        : PipelineBase(ctx), lights_(lights), cf_(cf) {}
    void SetClusterForwardUBO(const hvx_cf_ubo& ubo) { ubo_ = ubo; }
    // The reference zeroes its atomic list counter each frame; offsets here
    // come from a prefix sum, so there is nothing to reset.
    void ResetGlobalIndex() {}
    void FillCommandBuffer(hipStream_t stream) override {
        ctx_.check(hvx_cf_cull_lights(ctx_.get(), &ubo_, cf_->aabbBuffer_.data(), lights_->device(),
                                      lights_->GetLightCount(), cf_->cap(), cf_->lightCellsBuffer_.data(),
                                      cf_->lightIndicesBuffer_.data(), cf_->lightIndicesBuffer_.size(),
                                      cf_->globalIndex_.data(), nullptr, stream),
                   "hvx_cf_cull_lights");
    }

private:
    ResourcesLight* lights_;
    ResourcesClusterForward* cf_;
    hvx_cf_ubo ubo_{};
};

// ResourcesIBL (ResourcesIBL.cpp:18-56): environment cube, irradiance map,
// prefiltered specular chain and BRDF LUT, built by the IBL stages below.
class ResourcesIBL {
public:
    uint32_t cubeSize_ = 1024, diffuseSize_ = 32, specularSize_ = 128, specularMips_ = 8;
    uint32_t lutWidth_ = 256, lutHeight_ = 256, lutSamples_ = 1024, specularSamples_ = 1024;
    DeviceBuffer<float> environmentCubemap_, diffuseCubemap_, specularCubemap_, brdfLut_;

    hvx_ibl_maps maps() const {
        hvx_ibl_maps m{};
        m.specular = specularCubemap_.data();
        m.specular_size = specularSize_;
        m.specular_mips = specularMips_;
        m.diffuse = diffuseCubemap_.data();
        m.diffuse_size = diffuseSize_;
        m.brdf_lut = brdfLut_.data();
        m.lut_width = lutWidth_;
        m.lut_height = lutHeight_;
        return m;
    }
};

// PipelineEquirect2Cube (PipelineEquirect2Cube.h:15-18, .cpp:232-285).
class PipelineEquirect2Cube final : public PipelineBase {
public:
    PipelineEquirect2Cube(Context& ctx, const float* d_equirect, uint32_t w, uint32_t h)
        : PipelineBase(ctx), eq_(d_equirect), w_(w), h_(h) {}
    void OffscreenRender(DeviceBuffer<float>& cube, uint32_t faceSize, hipStream_t stream = nullptr) {
        const uint32_t mips = hvx_mip_count(faceSize);
        cube.resize(hvx_cube_chain_floats(faceSize, mips));
        ctx_.check(hvx_ibl_equirect_to_cube(ctx_.get(), eq_, w_, h_, cube.data(), faceSize, stream),
                   "hvx_ibl_equirect_to_cube");
    }
    void FillCommandBuffer(hipStream_t) override {}

private:
    const float* eq_;
    uint32_t w_, h_;
};

enum class CubeFilterType : uint8_t { Diffuse = 0, Specular = 1 };

// PipelineCubeFilter (PipelineCubeFilter.h:10-29, .cpp:11-72, :246-358): the
// constructor builds the input cube's mip chain as the reference does.
class PipelineCubeFilter final : public PipelineBase {
public:
    PipelineCubeFilter(Context& ctx, DeviceBuffer<float>* inputCube, uint32_t faceSize, hipStream_t stream = nullptr)
        : PipelineBase(ctx), in_(inputCube), size_(faceSize) {
        ctx_.check(hvx_ibl_cube_mips(ctx_.get(), in_->data(), size_, hvx_mip_count(size_), stream),
                   "hvx_ibl_cube_mips");
    }
    void OffscreenRender(DeviceBuffer<float>& out, CubeFilterType type, uint32_t outSize, uint32_t mips,
                         uint32_t samples, const hvx_filter_opts* opts = nullptr, hipStream_t stream = nullptr) {
        out.resize(hvx_cube_chain_floats(outSize, mips));
        ctx_.check(hvx_ibl_filter(ctx_.get(), (int)type, in_->data(), size_, out.data(), outSize, mips, samples, opts,
                                  stream),
                   "hvx_ibl_filter");
    }
    void FillCommandBuffer(hipStream_t) override {}

private:
    DeviceBuffer<float>* in_;
    uint32_t size_;
};

// PipelineBRDFLUT (PipelineBRDFLUT.h:15-21, .cpp:32-108).
class PipelineBRDFLUT final : public PipelineBase {
public:
    explicit PipelineBRDFLUT(Context& ctx) : PipelineBase(ctx) {}
    void CreateLUT(DeviceBuffer<float>& lut, uint32_t w, uint32_t h, uint32_t samples, hipStream_t stream = nullptr) {
        lut.resize((size_t)w * h * 2);
        ctx_.check(hvx_ibl_brdf_lut(ctx_.get(), w, h, samples, lut.data(), stream), "hvx_ibl_brdf_lut");
    }
    void FillCommandBuffer(hipStream_t) override {}
};

// Owner of a contraction-filter plan (hvx_ibl_contract_plan_*): the weights
// of one (type, face, out_face, out_mips, precision), built once.
class ContractPlan {
public:
    ContractPlan(Context& ctx, CubeFilterType type, uint32_t face, uint32_t outFace, uint32_t outMips,
                 hvx_precision precision = HVX_PRECISION_FP16_SCALED)
        : ctx_(ctx), type_(type), face_(face), outFace_(outFace), outMips_(outMips), precision_(precision) {
        ctx_.check(hvx_ibl_contract_plan_create(ctx_.get(), (int)type, face, outFace, outMips, precision, &plan_),
                   "hvx_ibl_contract_plan_create");
    }
    ~ContractPlan() {
        if (plan_) (void)hvx_ibl_contract_plan_destroy(plan_);
    }
    ContractPlan(const ContractPlan&) = delete;
    ContractPlan& operator=(const ContractPlan&) = delete;
    bool matches(CubeFilterType t, uint32_t face, uint32_t outFace, uint32_t mips, hvx_precision pr) const {
        return t == type_ && face == face_ && outFace == outFace_ && mips == outMips_ && pr == precision_;
    }
    // chain: the input cube with its box mips (levels levels)
    void Filter(const float* chain, uint32_t levels, float* out, hipStream_t stream = nullptr) {
        ctx_.check(hvx_ibl_filter_contract(plan_, chain, levels, out, stream), "hvx_ibl_filter_contract");
    }
    hvx_contract_mip_info Info(uint32_t mip) const {
        hvx_contract_mip_info i{};
        ctx_.check(hvx_ibl_contract_plan_query(plan_, mip, &i), "hvx_ibl_contract_plan_query");
        return i;
    }

private:
    Context& ctx_;
    CubeFilterType type_;
    uint32_t face_, outFace_, outMips_;
    hvx_precision precision_;
    hvx_contract_plan* plan_ = nullptr;
};

// PipelineCubeFilter's interface over the MFMA contraction filter (no
// reference counterpart: a deterministic quadrature of the same integral in
// place of the sampled estimate). Plans are built on first use and kept.
class PipelineCubeFilterContract final : public PipelineBase {
public:
    PipelineCubeFilterContract(Context& ctx, DeviceBuffer<float>* inputCube, uint32_t faceSize,
                               hipStream_t stream = nullptr)
        : PipelineBase(ctx), in_(inputCube), size_(faceSize), levels_(hvx_mip_count(faceSize)) {
        ctx_.check(hvx_ibl_cube_mips(ctx_.get(), in_->data(), size_, levels_, stream), "hvx_ibl_cube_mips");
    }
    void OffscreenRender(DeviceBuffer<float>& out, CubeFilterType type, uint32_t outSize, uint32_t mips,
                         hvx_precision precision = HVX_PRECISION_FP16_SCALED, hipStream_t stream = nullptr) {
        ContractPlan* plan = nullptr;
        for (auto& p : plans_)
            if (p->matches(type, size_, outSize, mips, precision)) plan = p.get();
        if (!plan) {
            plans_.push_back(std::make_unique<ContractPlan>(ctx_, type, size_, outSize, mips, precision));
            plan = plans_.back().get();
        }
        out.resize(hvx_cube_chain_floats(outSize, mips));
        plan->Filter(in_->data(), levels_, out.data(), stream);
    }
    void FillCommandBuffer(hipStream_t) override {}

private:
    DeviceBuffer<float>* in_;
    uint32_t size_, levels_;
    std::vector<std::unique_ptr<ContractPlan>> plans_;
};

// Owner of a per-environment radiance network (hvx_mlp_*): trained on
// samples of a prefiltered specular chain, with its measured cell errors for
// learned shading (no reference counterpart).
class RadianceNetwork {
public:
    RadianceNetwork(Context& ctx, float lodMax, uint64_t seed) : ctx_(ctx) {
        ctx_.check(hvx_mlp_create(ctx_.get(), lodMax, seed, &mlp_), "hvx_mlp_create");
    }
    ~RadianceNetwork() {
        if (mlp_) (void)hvx_mlp_destroy(mlp_);
    }
    RadianceNetwork(const RadianceNetwork&) = delete;
    RadianceNetwork& operator=(const RadianceNetwork&) = delete;
    // samples of the chain, then `steps` Adam steps of `batch` samples
    void Train(const ResourcesIBL& ibl, uint32_t samples, uint32_t batch, uint32_t steps, float lr, uint64_t seed,
               hipStream_t stream = nullptr) {
        inputs_.resize((size_t)samples * 4);
        targets_.resize((size_t)samples * 4);
        ctx_.check(hvx_mlp_sample(mlp_, ibl.specularCubemap_.data(), ibl.specularSize_, ibl.specularMips_, seed,
                                  samples, inputs_.data(), targets_.data(), stream),
                   "hvx_mlp_sample");
        ctx_.check(hvx_mlp_train(mlp_, inputs_.data(), targets_.data(), samples, batch, steps, lr, nullptr, stream),
                   "hvx_mlp_train");
    }
    // the fallback mask: measured error per (lod bin, face, cell)
    void MeasureCells(const ResourcesIBL& ibl, uint32_t cellFace, uint32_t lodBins, hipStream_t stream = nullptr) {
        cellFace_ = cellFace;
        lodBins_ = lodBins;
        cellErr_.resize((size_t)lodBins * 6 * cellFace * cellFace);
        ctx_.check(hvx_mlp_cell_errors(mlp_, ibl.specularCubemap_.data(), ibl.specularSize_, ibl.specularMips_, cellFace,
                                       lodBins, cellErr_.data(), stream),
                   "hvx_mlp_cell_errors");
    }
    const hvx_mlp* get() const { return mlp_; }
    const float* cellErrors() const { return cellErr_.data(); }
    uint32_t cellFace() const { return cellFace_; }
    uint32_t lodBins() const { return lodBins_; }

private:
    Context& ctx_;
    hvx_mlp* mlp_ = nullptr;
    DeviceBuffer<float> inputs_, targets_, cellErr_;
    uint32_t cellFace_ = 0, lodBins_ = 0;
};

// Builds every map of ResourcesIBL from an equirect (ResourcesIBL::Create).
inline void CreateIBL(Context& ctx, ResourcesIBL& ibl, const float* d_equirect, uint32_t w, uint32_t h,
                      hipStream_t stream = nullptr) {
    PipelineEquirect2Cube e2c(ctx, d_equirect, w, h);
    e2c.OffscreenRender(ibl.environmentCubemap_, ibl.cubeSize_, stream);
    PipelineCubeFilter filter(ctx, &ibl.environmentCubemap_, ibl.cubeSize_, stream);
    filter.OffscreenRender(ibl.diffuseCubemap_, CubeFilterType::Diffuse, ibl.diffuseSize_, 1, 0, nullptr, stream);
    filter.OffscreenRender(ibl.specularCubemap_, CubeFilterType::Specular, ibl.specularSize_, ibl.specularMips_,
                           ibl.specularSamples_, nullptr, stream);
    PipelineBRDFLUT lut(ctx);
    lut.CreateLUT(ibl.brdfLut_, ibl.lutWidth_, ibl.lutHeight_, ibl.lutSamples_, stream);
}

// PipelinePBRClusterForward (PipelinePBRClusterForward.h:20-43, .cpp:77-113),
// shading SoA fragments into planar FP32 RGB.
class PipelinePBRClusterForward final : public PipelineBase {
public:
    PipelinePBRClusterForward(Context& ctx, const FragmentsSoA* frags, ResourcesLight* lights,
                              ResourcesClusterForward* cf, const ResourcesIBL* ibl, DeviceBuffer<float>* rgb)
        : PipelineBase(ctx), frags_(frags), lights_(lights), cf_(cf), ibl_(ibl), rgb_(rgb) {}
    void SetClusterForwardUBO(const hvx_cf_ubo& ubo) { ubo_ = ubo; }
    void SetPBRPushConstants(const hvx_pbr_pc& pc) { pc_ = pc; }
    // Learned specular IBL (hvx_cf_shade_learned) with fallback tolerance
    // `tol`; nullptr returns to the prefiltered map. served (nullable) counts
    // the fragments the network served.
    void SetLearned(const RadianceNetwork* net, float tol = 2e-3f, uint32_t* d_served = nullptr) {
        net_ = net;
        tol_ = tol;
        served_ = d_served;
    }
    void FillCommandBuffer(hipStream_t stream) override {
        const hvx_frag_soa f = frags_->view();
        const hvx_ibl_maps m = ibl_->maps();
        if (net_)
            ctx_.check(hvx_cf_shade_learned(ctx_.get(), &f, &cameraUBO_, &ubo_, &pc_, lights_->device(),
                                            lights_->GetLightCount(), cf_->lightCellsBuffer_.data(),
                                            cf_->lightIndicesBuffer_.data(), &m, net_->get(), net_->cellErrors(),
                                            net_->cellFace(), net_->lodBins(), tol_, rgb_->data(), served_, stream),
                       "hvx_cf_shade_learned");
        else
            ctx_.check(hvx_cf_shade(ctx_.get(), &f, &cameraUBO_, &ubo_, &pc_, lights_->device(),
                                    lights_->GetLightCount(), cf_->lightCellsBuffer_.data(),
                                    cf_->lightIndicesBuffer_.data(), &m, rgb_->data(), stream),
                       "hvx_cf_shade");
    }

private:
    const FragmentsSoA* frags_;
    ResourcesLight* lights_;
    ResourcesClusterForward* cf_;
    const ResourcesIBL* ibl_;
    DeviceBuffer<float>* rgb_;
    hvx_cf_ubo ubo_{};
    hvx_pbr_pc pc_{1.75f, 0.01f, 1.0f, 0.1f, 0.01f};  // UIData.h:22-29, AppPBRClusterForward.cpp:21
    const RadianceNetwork* net_ = nullptr;
    float tol_ = 2e-3f;
    uint32_t* served_ = nullptr;
};

// The post step after shading (Shaders/Common/Tonemap.frag:11-94): tone
// mapping of the planar FP32 frame into packed BGRA8 (and, optionally, the
// mapped FP32 planes).
class PipelineTonemap final : public PipelineBase {
public:
    PipelineTonemap(Context& ctx, const DeviceBuffer<float>* rgb, uint32_t width, uint32_t height,
                    DeviceBuffer<uint32_t>* bgra8, DeviceBuffer<float>* mapped = nullptr)
        : PipelineBase(ctx), rgb_(rgb), w_(width), h_(height), bgra8_(bgra8), mapped_(mapped) {}
    void SetOperator(hvx_tonemap_op op, bool gamma = true) {
        op_ = op;
        gamma_ = gamma;
    }
    void FillCommandBuffer(hipStream_t stream) override {
        ctx_.check(hvx_post_tonemap(ctx_.get(), rgb_->data(), w_, h_, (int)op_, gamma_ ? 1 : 0,
                                    mapped_ ? mapped_->data() : nullptr, bgra8_ ? bgra8_->data() : nullptr, stream),
                   "hvx_post_tonemap");
    }

private:
    const DeviceBuffer<float>* rgb_;
    uint32_t w_, h_;
    DeviceBuffer<uint32_t>* bgra8_;
    DeviceBuffer<float>* mapped_;
    hvx_tonemap_op op_ = HVX_TONEMAP_REINHARD;
    bool gamma_ = true;
};

// Captures the stages' FillCommandBuffer calls (in order) into one hipGraph
// on a capture stream and replays it (AppBase::FillCommandBuffer's role). The
// stages must have run once eagerly first so their scratch is allocated (the
// library refuses to allocate under capture). Call Capture() again after any
// stage parameter changes (camera/UBO, push constants): the executable graph
// is updated in place when the topology is unchanged. Launch() refuses a
// graph whose contexts have reallocated scratch since the capture (the
// graph's pointers to it are freed): capture again.
class FrameGraph {
public:
    explicit FrameGraph(std::vector<PipelineBase*> stages) : stages_(std::move(stages)) {
        hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    }
    ~FrameGraph() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    FrameGraph(const FrameGraph&) = delete;
    FrameGraph& operator=(const FrameGraph&) = delete;

    void Capture() {
        gens_.clear();
        for (PipelineBase* p : stages_) gens_.push_back(hvx_ctx_scratch_generation(p->context().get()));
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            for (PipelineBase* p : stages_) p->FillCommandBuffer(stream_);
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream_, &g), "hipStreamEndCapture");
        if (exec_) {
            hipGraphExecUpdateResult res;
            hipGraphNode_t err_node = nullptr;
            if (hipGraphExecUpdate(exec_, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
                (void)hipGraphDestroy(graph_);
                graph_ = g;
                updates_ += 1;
                return;
            }
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(exec_);
            exec_ = nullptr;
        }
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_ = g;
        hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    void Launch() {
        if (!exec_) throw Error(HVX_ERR_INVALID_ARG, "FrameGraph::Launch before Capture");
        for (size_t i = 0; i < stages_.size(); ++i)
            if (hvx_ctx_scratch_generation(stages_[i]->context().get()) != gens_[i])
                throw Error(HVX_ERR_STALE_GRAPH, "FrameGraph::Launch: a stage's context reallocated its scratch "
                                                 "after the capture; Capture() again");
        hip_check(hipGraphLaunch(exec_, stream_), "hipGraphLaunch");
    }
    // captures that updated the executable graph in place
    size_t update_count() const { return updates_; }
    hipStream_t stream() const { return stream_; }
    size_t node_count() const {
        size_t n = 0;
        if (graph_) (void)hipGraphGetNodes(graph_, nullptr, &n);
        return n;
    }

private:
    std::vector<PipelineBase*> stages_;
    std::vector<uint64_t> gens_;
    hipStream_t stream_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    size_t updates_ = 0;
};

}  // namespace hvx
