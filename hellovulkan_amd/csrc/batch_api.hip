// batch_api.hip — batched probe prefiltering and multi-view light culling
// (hvx_batch_* of include/hvx.h): work-item construction, device-resident
// job tables, the autotuner and graph capture. Kernels live in
// ibl_kernels.hip (probe_batch_kernel) and cf_kernels.hip (cull_batch_kernel
// and the batched scan/scatter).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "hvx_device.h"
#include "hvx_internal.h"

static_assert(sizeof(hvx_probe) == 32 && sizeof(hvx_view) == 232 && sizeof(hvx_batch_tuning) == 24,
              "batch descriptor layouts are part of the ABI");

namespace {

constexpr uint32_t kMaxBatchViews = 64;

template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        release();
        n = count;
        return hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct ViewScratch {
    DevBuf<uint32_t> lists, meta;  // meta: counts[C], over[C], uint2 partials[nb]
    DevBuf<float4> lvs;
};

}  // namespace

// Contraction mode: the probes of one (face, out_face, out_mips) class share
// a contraction plan (with per-probe scratch reserved) and a stream; their
// input chains sit `stride` floats apart in one block (level 0 gathered in,
// box mips built), and one multi-probe call filters them all. Classes run
// concurrently.
struct ContractClass {
    uint32_t S = 0, So = 0, mips = 0, levels = 0;
    size_t stride = 0;  // floats per input chain
    hvx_contract_plan* plan = nullptr;
    float* chains = nullptr;
    const float** d_src = nullptr;  // device arrays of n pointers: probe level 0,
    float** d_in = nullptr;         // input chains,
    float** d_out = nullptr;        // output chains
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<uint32_t> probes;
};

struct hvx_batch {
    hvx_ctx* ctx = nullptr;
    uint32_t n = 0;  // sample count
    std::vector<hvx_probe> probes;
    std::vector<hvx_view> views;
    std::vector<hvx::ProbeJob> jobs;
    DevBuf<hvx::ProbeJob> d_jobs;
    DevBuf<float4> d_tables;
    DevBuf<hvx::ProbeItem> d_items;
    DevBuf<uint32_t> d_xcd_begin;  // [9]: item runs of the per-XCD probe queues
    uint32_t n_items = 0;
    std::vector<ViewScratch> vscratch;
    DevBuf<hvx::CullBatchView> d_views;
    hvx::CullBatchTotals totals{};
    DevBuf<uint32_t> d_queues;  // [0, 8) per-XCD probe queues, [8] culling queue
    hvx_batch_tuning tuning{256, 4, 0, 8, 0.0f, 0.0f};
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, cfork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_stale = true;
    std::vector<ContractClass> classes;  // non-empty: probes use the contraction filter
};

namespace {

int bfail(hvx_batch* b, int st, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int bfail(hvx_batch* b, int st, const char* fmt, ...) {
    if (b && b->ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        b->ctx->last_error = buf;
    }
    return st;
}

int bhip(hvx_batch* b, hipError_t e, const char* where) {
    return bfail(b, HVX_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

bool misaligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) != 0; }

int upload_items(hvx_batch* b) {
    std::vector<hvx::ProbeItem> items;
    uint32_t xcd_begin[9];
    hvx::make_probe_items(b->jobs.data(), (uint32_t)b->jobs.size(), b->tuning.probe_tile, items, xcd_begin);
    hipError_t e;
    if (b->d_items.n < items.size() && (e = b->d_items.alloc(items.size())) != hipSuccess)
        return bhip(b, e, "batch item allocation");
    if (!b->d_xcd_begin.p && (e = b->d_xcd_begin.alloc(9)) != hipSuccess) return bhip(b, e, "batch item runs");
    if (!items.empty() &&
        (e = hipMemcpy(b->d_items.p, items.data(), items.size() * sizeof(hvx::ProbeItem), hipMemcpyHostToDevice)) !=
            hipSuccess)
        return bhip(b, e, "batch item upload");
    if ((e = hipMemcpy(b->d_xcd_begin.p, xcd_begin, sizeof(xcd_begin), hipMemcpyHostToDevice)) != hipSuccess)
        return bhip(b, e, "batch item runs upload");
    b->n_items = (uint32_t)items.size();
    return HVX_OK;
}

// Device view table for culling shape `shape` (group offsets depend on it).
int upload_views(hvx_batch* b) {
    std::vector<hvx::CullBatchView> v(b->views.size());
    hvx::CullBatchTotals t{};
    for (size_t i = 0; i < b->views.size(); ++i) {
        const hvx_view& w = b->views[i];
        hvx::CullBatchView& c = v[i];
        const uint32_t X = w.ubo.slice_count_x, Y = w.ubo.slice_count_y, Z = w.ubo.slice_count_z, C = X * Y * Z;
        const uint32_t nb = (C + 1023) / 1024;
        ViewScratch& s = b->vscratch[i];
        c.job.X = X;
        c.job.Y = Y;
        c.job.Z = Z;
        c.job.n = w.n_lights;
        c.job.cap = w.cap;
        c.job.want_overflow = 0;
        c.job.clusters = w.clusters;
        c.job.lvs = s.lvs.p;
        c.job.lists = s.lists.p;
        c.job.counts = s.meta.p;
        c.job.over = s.meta.p + C;
        c.lights = w.lights;
        std::memcpy(c.view, w.ubo.camera_view, sizeof(c.view));
        c.part = reinterpret_cast<uint2*>(s.meta.p + 2 * C);
        c.cells = w.cells;
        c.indices = w.indices;
        c.total = w.total;
        c.novf = nullptr;
        c.C = C;
        c.nb = nb;
        c.light_begin = t.lights;
        c.group_begin = t.groups;
        c.block_begin = t.blocks;
        c.scatter_begin = t.scatter;
        t.lights += w.n_lights;
        t.groups += hvx::cull_batch_groups((int)b->tuning.cull_shape, X, Y, Z);
        t.blocks += nb;
        t.scatter += (C + 3) / 4;
    }
    hipError_t e;
    if (b->d_views.n < v.size() && (e = b->d_views.alloc(v.size())) != hipSuccess)
        return bhip(b, e, "batch view table allocation");
    if (!v.empty() &&
        (e = hipMemcpy(b->d_views.p, v.data(), v.size() * sizeof(hvx::CullBatchView), hipMemcpyHostToDevice)) !=
            hipSuccess)
        return bhip(b, e, "batch view table upload");
    b->totals = t;
    return HVX_OK;
}

// Contraction mode: per class, on its own stream forked from `s`: gather the
// probes' level 0 into their chains, build the box mips of all chains, then
// one multi-probe contraction call.
hipError_t run_contract(hvx_batch* b, hipStream_t s) {
    hipError_t e;
    if ((e = hipEventRecord(b->cfork, s)) != hipSuccess) return e;
    for (ContractClass& c : b->classes) {
        const uint32_t n = (uint32_t)c.probes.size();
        if ((e = hipStreamWaitEvent(c.stream, b->cfork, 0)) != hipSuccess) return e;
        if ((e = hvx::launch_gather_copy(c.d_src, (size_t)6 * c.S * c.S * 4, c.chains, c.stride, n, c.stream)) !=
                hipSuccess ||
            (e = hvx::launch_cube_mips(c.chains, c.S, c.levels, c.stream, n, c.stride)) != hipSuccess)
            return e;
        if (hvx_ibl_filter_contract_multi(c.plan, n, c.d_in, c.levels, c.d_out, c.stream) != HVX_OK)
            return hipErrorLaunchFailure;
        if ((e = hipEventRecord(c.done, c.stream)) != hipSuccess || (e = hipStreamWaitEvent(s, c.done, 0)) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t run_probes(hvx_batch* b, hipStream_t s) {
    if (!b->classes.empty()) return run_contract(b, s);
    if (b->n_items == 0) return hipSuccess;
    return hvx::launch_probe_batch(b->d_jobs.p, b->d_items.p, b->d_xcd_begin.p, b->d_tables.p,
                                   (uint32_t)b->ctx->num_cus * b->tuning.probe_wgs_per_cu, b->d_queues.p, s);
}

hipError_t run_culls(hvx_batch* b, hipStream_t s) {
    if (b->views.empty()) return hipSuccess;
    return hvx::launch_cull_batch(b->d_views.p, (uint32_t)b->views.size(), b->totals, (int)b->tuning.cull_shape,
                                  (uint32_t)b->ctx->num_cus * b->tuning.cull_wgs_per_cu, b->d_queues.p + hvx::kProbeQueues, s);
}

hipError_t run_all(hvx_batch* b, hipStream_t s) {
    hipError_t e;
    if ((e = hvx::launch_queue_reset(b->d_queues.p, hvx::kProbeQueues + 1, s)) != hipSuccess) return e;
    if ((e = hipEventRecord(b->fork, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(b->side, b->fork, 0)) != hipSuccess) return e;
    if ((e = run_probes(b, b->side)) != hipSuccess) return e;
    if ((e = run_culls(b, s)) != hipSuccess) return e;
    if ((e = hipEventRecord(b->join, b->side)) != hipSuccess) return e;
    return hipStreamWaitEvent(s, b->join, 0);
}

void release_classes(hvx_batch* b) {
    for (ContractClass& c : b->classes) {
        if (c.plan) (void)hvx_ibl_contract_plan_destroy(c.plan);
        if (c.chains) (void)hipFree(c.chains);
        if (c.d_src) (void)hipFree(c.d_src);
        if (c.d_in) (void)hipFree(c.d_in);
        if (c.d_out) (void)hipFree(c.d_out);
        if (c.done) (void)hipEventDestroy(c.done);
        if (c.stream) (void)hipStreamDestroy(c.stream);
    }
    b->classes.clear();
}

int check_tuning(hvx_batch* b, const hvx_batch_tuning& t) {
    if (t.probe_tile < 32 || t.probe_tile > 8192 || t.probe_wgs_per_cu < 1 || t.probe_wgs_per_cu > 16 ||
        t.cull_shape >= (uint32_t)hvx::kCullShapes || t.cull_wgs_per_cu < 1 || t.cull_wgs_per_cu > 16)
        return bfail(b, HVX_ERR_INVALID_ARG, "tuning out of range");
    return HVX_OK;
}

float time_ms(hvx_batch* b, hipStream_t s, bool probes, int reps) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return std::numeric_limits<float>::infinity();
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return std::numeric_limits<float>::infinity();
    }
    float best = std::numeric_limits<float>::infinity();
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, s);
        hipError_t e = hvx::launch_queue_reset(b->d_queues.p, hvx::kProbeQueues + 1, s);
        if (e == hipSuccess) e = probes ? run_probes(b, s) : run_culls(b, s);
        (void)hipEventRecord(e1, s);
        if (e != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

extern "C" {

int hvx_batch_create(hvx_ctx* ctx, const hvx_probe* probes, uint32_t n_probes, uint32_t sample_count,
                     const hvx_view* views, uint32_t n_views, hvx_batch** out) {
    if (!ctx || !out) return HVX_ERR_NULL;
    *out = nullptr;
    if ((n_probes && !probes) || (n_views && !views)) return HVX_ERR_NULL;
    if (n_views > kMaxBatchViews) {
        ctx->last_error = "too many views in one batch (max 64)";
        return HVX_ERR_INVALID_ARG;
    }
    if (n_probes && (sample_count == 0 || sample_count > (1u << 20))) {
        ctx->last_error = "sample_count not in [1, 2^20]";
        return HVX_ERR_INVALID_ARG;
    }
    for (uint32_t i = 0; i < n_probes; ++i) {
        const hvx_probe& p = probes[i];
        if (!p.cube || !p.out_chain) return HVX_ERR_NULL;
        if (p.face_size == 0 || p.face_size > 16384 || p.out_face_size == 0 || p.out_face_size > 16384 ||
            p.out_mips == 0 || p.out_mips > hvx_mip_count(p.out_face_size) || p.out_mips > 16) {
            ctx->last_error = "probe " + std::to_string(i) + ": sizes out of range";
            return HVX_ERR_INVALID_ARG;
        }
        if (misaligned(p.cube, 16) || misaligned(p.out_chain, 16)) return HVX_ERR_ALIGN;
    }
    for (uint32_t i = 0; i < n_views; ++i) {
        const hvx_view& v = views[i];
        const uint64_t C = (uint64_t)v.ubo.slice_count_x * v.ubo.slice_count_y * v.ubo.slice_count_z;
        if (!v.clusters || !v.cells || !v.indices || (v.n_lights && !v.lights)) return HVX_ERR_NULL;
        if (C == 0 || C > (1u << 24) || v.cap < 1 || v.cap > 1024 || v.n_lights > (1u << 28) ||
            C * v.cap > 0xffffffffull || (uint64_t)v.indices_capacity < C * v.cap) {
            ctx->last_error = "view " + std::to_string(i) + ": grid/cap/capacity out of range";
            return C == 0 || v.cap < 1 || v.cap > 1024 ? HVX_ERR_INVALID_ARG : HVX_ERR_CAPACITY;
        }
        if (misaligned(v.clusters, 16) || misaligned(v.lights, 16) || misaligned(v.cells, 8)) return HVX_ERR_ALIGN;
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess))
        return HVX_ERR_HIP;
    hvx_batch* b = new (std::nothrow) hvx_batch();
    if (!b) return HVX_ERR_HIP;
    b->ctx = ctx;
    b->n = sample_count;
    b->probes.assign(probes, probes + n_probes);
    b->views.assign(views, views + n_views);
    auto fail_destroy = [&](int st) {
        hvx_batch_destroy(b);
        return st;
    };
    hipError_t e;
    // one GGX table per distinct mip count, computed once
    std::vector<uint32_t> mips_seen;
    std::vector<uint32_t> table_off;
    uint32_t table_len = 0;
    for (const hvx_probe& p : b->probes) {
        auto it = std::find(mips_seen.begin(), mips_seen.end(), p.out_mips);
        if (it == mips_seen.end()) {
            mips_seen.push_back(p.out_mips);
            table_off.push_back(table_len);
            table_len += p.out_mips * sample_count;
        }
    }
    if ((e = b->d_tables.alloc(table_len)) != hipSuccess) return fail_destroy(bhip(b, e, "GGX tables"));
    for (size_t k = 0; k < mips_seen.size(); ++k)
        if ((e = hvx::launch_ggx_tables(mips_seen[k], sample_count, b->d_tables.p + table_off[k], nullptr)) !=
            hipSuccess)
            return fail_destroy(bhip(b, e, "GGX tables"));
    b->jobs.resize(n_probes);
    for (uint32_t i = 0; i < n_probes; ++i) {
        const hvx_probe& p = b->probes[i];
        hvx::ProbeJob& j = b->jobs[i];
        j.cube = reinterpret_cast<const float4*>(p.cube);
        j.out = reinterpret_cast<float4*>(p.out_chain);
        j.S = p.face_size;
        j.So = p.out_face_size;
        j.mips = p.out_mips;
        j.n = sample_count;
        j.table = table_off[std::find(mips_seen.begin(), mips_seen.end(), p.out_mips) - mips_seen.begin()];
    }
    if ((e = b->d_jobs.alloc(n_probes)) != hipSuccess) return fail_destroy(bhip(b, e, "probe jobs"));
    if (n_probes && (e = hipMemcpy(b->d_jobs.p, b->jobs.data(), n_probes * sizeof(hvx::ProbeJob),
                                   hipMemcpyHostToDevice)) != hipSuccess)
        return fail_destroy(bhip(b, e, "probe jobs"));
    b->vscratch.resize(n_views);
    for (uint32_t i = 0; i < n_views; ++i) {
        const hvx_view& v = b->views[i];
        const size_t C = (size_t)v.ubo.slice_count_x * v.ubo.slice_count_y * v.ubo.slice_count_z;
        const size_t nb = (C + 1023) / 1024;
        ViewScratch& s = b->vscratch[i];
        if ((e = s.lists.alloc(C * v.cap)) != hipSuccess || (e = s.meta.alloc(2 * C + 2 * nb + 2)) != hipSuccess ||
            (e = s.lvs.alloc(v.n_lights)) != hipSuccess)
            return fail_destroy(bhip(b, e, "view scratch"));
    }
    if ((e = b->d_queues.alloc(hvx::kProbeQueues + 1)) != hipSuccess) return fail_destroy(bhip(b, e, "queues"));
    if ((e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->cfork, hipEventDisableTiming)) != hipSuccess)
        return fail_destroy(bhip(b, e, "streams/events"));
    int st;
    if ((st = upload_items(b)) || (st = upload_views(b))) return fail_destroy(st);
    if ((e = hipDeviceSynchronize()) != hipSuccess) return fail_destroy(bhip(b, e, "batch set-up"));
    *out = b;
    return HVX_OK;
}

int hvx_batch_destroy(hvx_batch* b) {
    if (!b) return HVX_ERR_NULL;
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    release_classes(b);
    if (b->cfork) (void)hipEventDestroy(b->cfork);
    b->d_jobs.release();
    b->d_tables.release();
    b->d_items.release();
    b->d_xcd_begin.release();
    b->d_views.release();
    b->d_queues.release();
    for (ViewScratch& s : b->vscratch) {
        s.lists.release();
        s.meta.release();
        s.lvs.release();
    }
    delete b;
    return HVX_OK;
}

int hvx_batch_set_tuning(hvx_batch* b, const hvx_batch_tuning* t) {
    if (!b || !t) return HVX_ERR_NULL;
    int st;
    if ((st = check_tuning(b, *t))) return st;
    (void)hipDeviceSynchronize();  // the device tables may be in use
    const bool items = t->probe_tile != b->tuning.probe_tile, views = t->cull_shape != b->tuning.cull_shape;
    b->tuning = *t;
    if ((items && (st = upload_items(b))) || (views && (st = upload_views(b)))) return st;
    b->graph_stale = true;
    return HVX_OK;
}

int hvx_batch_get_tuning(const hvx_batch* b, hvx_batch_tuning* t) {
    if (!b || !t) return HVX_ERR_NULL;
    *t = b->tuning;
    return HVX_OK;
}

int hvx_batch_autotune(hvx_batch* b, void* stream) {
    if (!b) return HVX_ERR_NULL;
    hipStream_t s = (hipStream_t)stream;
    hvx_batch_tuning best = b->tuning;
    int st;
    if (!b->probes.empty() && b->classes.empty()) {
        float best_ms = std::numeric_limits<float>::infinity();
        for (uint32_t tile : {64u, 128u, 256u, 512u, 1024u})
            for (uint32_t wgs : {2u, 4u, 8u}) {
                hvx_batch_tuning t = b->tuning;
                t.probe_tile = tile;
                t.probe_wgs_per_cu = wgs;
                if ((st = hvx_batch_set_tuning(b, &t))) return st;
                const float ms = time_ms(b, s, true, 2);
                if (ms < best_ms) {
                    best_ms = ms;
                    best.probe_tile = tile;
                    best.probe_wgs_per_cu = wgs;
                }
            }
        best.probe_ms = best_ms;
    }
    if (!b->views.empty()) {
        float best_ms = std::numeric_limits<float>::infinity();
        for (uint32_t shape = 0; shape < (uint32_t)hvx::kCullShapes; ++shape)
            for (uint32_t wgs : {4u, 8u}) {
                hvx_batch_tuning t = b->tuning;
                t.cull_shape = shape;
                t.cull_wgs_per_cu = wgs;
                if ((st = hvx_batch_set_tuning(b, &t))) return st;
                const float ms = time_ms(b, s, false, 3);
                if (ms < best_ms) {
                    best_ms = ms;
                    best.cull_shape = shape;
                    best.cull_wgs_per_cu = wgs;
                }
            }
        best.cull_ms = best_ms;
    }
    return hvx_batch_set_tuning(b, &best);
}

int hvx_batch_set_contract(hvx_batch* b, int enable, int precision) {
    if (!b) return HVX_ERR_NULL;
    if (enable && precision != HVX_PRECISION_FP16_SCALED && precision != HVX_PRECISION_BF16_SPLIT)
        return bfail(b, HVX_ERR_INVALID_ARG, "unknown contraction precision %d", precision);
    (void)hipDeviceSynchronize();  // the class streams may be in use
    release_classes(b);
    b->graph_stale = true;
    if (!enable) return HVX_OK;
    for (uint32_t i = 0; i < (uint32_t)b->probes.size(); ++i) {
        const hvx_probe& p = b->probes[i];
        auto it = std::find_if(b->classes.begin(), b->classes.end(), [&](const ContractClass& c) {
            return c.S == p.face_size && c.So == p.out_face_size && c.mips == p.out_mips;
        });
        if (it == b->classes.end()) {
            b->classes.emplace_back();
            ContractClass& c = b->classes.back();
            c.S = p.face_size;
            c.So = p.out_face_size;
            c.mips = p.out_mips;
            c.levels = hvx_mip_count(p.face_size);
            int st;
            if ((st = hvx_ibl_contract_plan_create(b->ctx, 1, c.S, c.So, c.mips, precision, &c.plan))) {
                release_classes(b);
                return st;
            }
            it = b->classes.end() - 1;
        }
        it->probes.push_back(i);
    }
    for (ContractClass& c : b->classes) {
        const uint32_t n = (uint32_t)c.probes.size();
        c.stride = (hvx_cube_chain_floats(c.S, c.levels) + 63) & ~(size_t)63;  // 256-byte aligned chains
        std::vector<const float*> src(n);
        std::vector<float*> in(n), out(n);
        int st;
        hipError_t e;
        if ((st = hvx_ibl_contract_plan_reserve(c.plan, n))) {
            release_classes(b);
            return st;
        }
        if ((e = hipMalloc(reinterpret_cast<void**>(&c.chains), c.stride * n * sizeof(float))) != hipSuccess ||
            (e = hipMalloc(reinterpret_cast<void**>(&c.d_src), n * sizeof(void*))) != hipSuccess ||
            (e = hipMalloc(reinterpret_cast<void**>(&c.d_in), n * sizeof(void*))) != hipSuccess ||
            (e = hipMalloc(reinterpret_cast<void**>(&c.d_out), n * sizeof(void*))) != hipSuccess) {
            release_classes(b);
            return bhip(b, e, "contraction class buffers");
        }
        for (uint32_t k = 0; k < n; ++k) {
            src[k] = b->probes[c.probes[k]].cube;
            in[k] = c.chains + (size_t)k * c.stride;
            out[k] = b->probes[c.probes[k]].out_chain;
        }
        if ((e = hipMemcpy(c.d_src, src.data(), n * sizeof(void*), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(c.d_in, in.data(), n * sizeof(void*), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(c.d_out, out.data(), n * sizeof(void*), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&c.done, hipEventDisableTiming)) != hipSuccess) {
            release_classes(b);
            return bhip(b, e, "contraction class set-up");
        }
    }
    // the largest classes first, so their streams start earliest
    std::stable_sort(b->classes.begin(), b->classes.end(),
                     [](const ContractClass& a, const ContractClass& c) { return a.S * a.probes.size() > c.S * c.probes.size(); });
    return HVX_OK;
}

int hvx_batch_run(hvx_batch* b, void* stream) {
    if (!b) return HVX_ERR_NULL;
    const hipError_t e = run_all(b, (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : bhip(b, e, "hvx_batch_run");
}

int hvx_batch_capture(hvx_batch* b, void* stream) {
    if (!b) return HVX_ERR_NULL;
    hipStream_t s = (hipStream_t)stream;
    if (!s) return bfail(b, HVX_ERR_INVALID_ARG, "graph capture needs a non-default stream");
    // Capturing the contraction mode ended in a host-side crash whose cause is
    // not established; until it is, the mode runs eagerly only.
    if (!b->classes.empty())
        return bfail(b, HVX_ERR_UNSUPPORTED, "graph capture of the contraction mode is disabled; use hvx_batch_run");
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bhip(b, e, "hipStreamBeginCapture");
    const hipError_t er = run_all(b, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (er != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return bhip(b, er != hipSuccess ? er : e, "batch capture");
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = g;
    if ((e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0)) != hipSuccess)
        return bhip(b, e, "hipGraphInstantiate");
    b->graph_stale = false;
    return HVX_OK;
}

int hvx_batch_launch(hvx_batch* b, void* stream) {
    if (!b) return HVX_ERR_NULL;
    if (!b->exec || b->graph_stale) return bfail(b, HVX_ERR_INVALID_ARG, "capture the batch first (hvx_batch_capture)");
    const hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    return e == hipSuccess ? HVX_OK : bhip(b, e, "hipGraphLaunch");
}

}  // extern "C"
