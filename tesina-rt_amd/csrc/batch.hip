// Batches of frames of different resident scenes, rendered side by side on
// several streams (fork from the caller's stream, join back), capturable
// into one hipGraph.
#include "rtx_impl.hpp"

struct rtx_batch {
    rtx_ctx* ctx = nullptr;
    std::vector<rtx_renderer*> r;
    std::vector<rtx_scene*> s;
    std::vector<rtx_camera> cam;
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int32_t rtx_batch_create(rtx_ctx* ctx, rtx_renderer* const* renderers, rtx_scene* const* scenes,
                         const rtx_camera* cams, int32_t n, int32_t streams, rtx_batch** out) {
    RTX_REQUIRE(ctx && renderers && scenes && cams && out && n > 0, RTX_EINVAL, "rtx_batch_create: bad argument");
    RTX_REQUIRE(streams >= 1 && streams <= 16, RTX_EINVAL, "rtx_batch_create: streams must be in [1, 16]");
    for (int i = 0; i < n; ++i) {
        RTX_REQUIRE(renderers[i] && scenes[i], RTX_EINVAL, "rtx_batch_create: null renderer or scene");
        RTX_REQUIRE(scenes[i]->built, RTX_ESTATE, "rtx_batch_create: scenes must be built");
        for (int j = 0; j < i; ++j)
            RTX_REQUIRE(renderers[j] != renderers[i], RTX_EINVAL,
                        "rtx_batch_create: a renderer may render one frame of the batch");
    }
    // the batch's own streams give the concurrency; each frame renders on one
    for (int i = 0; i < n; ++i) renderers[i]->single_stream = true;
    rtx_batch* b = new rtx_batch();
    b->ctx = ctx;
    b->r.assign(renderers, renderers + n);
    b->s.assign(scenes, scenes + n);
    b->cam.assign(cams, cams + n);
    int32_t e = RTX_OK;
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) e = RTX_EDEVICE;
    for (int k = 0; k < streams && !e; ++k) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            e = RTX_EDEVICE;
            if (st) (void)hipStreamDestroy(st);
            break;
        }
        b->streams.push_back(st);
        b->join.push_back(ev);
    }
    if (e) {
        rtx_set_error("rtx_batch_create: stream or event creation failed");
        rtx_batch_destroy(b);
        return e;
    }
    *out = b;
    return RTX_OK;
}

int32_t rtx_batch_destroy(rtx_batch* b) {
    RTX_REQUIRE(b, RTX_EINVAL, "rtx_batch_destroy: null");
    for (auto st : b->streams) (void)hipStreamSynchronize(st);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto ev : b->join) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (auto st : b->streams) (void)hipStreamDestroy(st);
    delete b;
    return RTX_OK;
}

int32_t rtx_batch_render(rtx_batch* b, void* stream) {
    RTX_REQUIRE(b, RTX_EINVAL, "rtx_batch_render: null batch");
    hipStream_t st = b->ctx->pick(stream);
    const int K = (int)b->streams.size();
    RTX_CHECK_HIP(hipEventRecord(b->fork, st));
    for (int k = 0; k < K; ++k) RTX_CHECK_HIP(hipStreamWaitEvent(b->streams[k], b->fork, 0));
    ProfScope ps(b->ctx, st, "BATCH");
    for (size_t i = 0; i < b->r.size(); ++i) {
        int32_t e = rtx_render_frame_async(b->r[i], b->s[i], &b->cam[i], 0, b->streams[i % K]);
        if (e) return e;
    }
    for (int k = 0; k < K; ++k) {
        RTX_CHECK_HIP(hipEventRecord(b->join[k], b->streams[k]));
        RTX_CHECK_HIP(hipStreamWaitEvent(st, b->join[k], 0));
    }
    return RTX_OK;
}

int32_t rtx_batch_capture(rtx_batch* b, void* stream) {
    RTX_REQUIRE(b, RTX_EINVAL, "rtx_batch_capture: null batch");
    hipStream_t st = b->ctx->pick(stream);
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    const bool prof = b->ctx->profiling;
    b->ctx->profiling = false;
    RTX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int32_t e = rtx_batch_render(b, st);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &g);
    b->ctx->profiling = prof;
    if (e) {
        if (g) (void)hipGraphDestroy(g);
        return e;
    }
    if (ce != hipSuccess) {
        rtx_set_error(std::string("rtx_batch_capture: ") + hipGetErrorString(ce));
        return RTX_EDEVICE;
    }
    b->graph = g;
    RTX_CHECK_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return RTX_OK;
}

int32_t rtx_batch_replay(rtx_batch* b, void* stream) {
    RTX_REQUIRE(b && b->exec, RTX_ESTATE, "rtx_batch_replay: nothing captured");
    hipStream_t st = b->ctx->pick(stream);
    ProfScope ps(b->ctx, st, "BATCH_GRAPH");
    RTX_CHECK_HIP(hipGraphLaunch(b->exec, st));
    for (auto* r : b->r) r->last_async = true;
    return RTX_OK;
}

int32_t rtx_batch_stats(rtx_batch* b, uint64_t* rays, uint64_t* shadow_rays) {
    RTX_REQUIRE(b, RTX_EINVAL, "rtx_batch_stats: null batch");
    uint64_t ra = 0, sh = 0;
    for (auto* r : b->r) {
        rtx_frame_stats fs{};
        int32_t e = rtx_renderer_last_stats(r, &fs);
        if (e) return e;
        ra += fs.primary_rays + fs.secondary_rays;
        sh += fs.shadow_rays;
    }
    if (rays) *rays = ra;
    if (shadow_rays) *shadow_rays = sh;
    return RTX_OK;
}
