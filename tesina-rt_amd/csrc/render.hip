// Frame orchestration: the plain one-launch-per-stage frame and the
// read-back-free frame.
//
// Reference: Renderer::set_up_frame / update_accelerator /
// render_to_framebuffer / copy_framebuffer / conclude_frame
// (src/rt/renderer.cpp:12-272).  The whole frame is one tile: the reference
// splits it into tiles of CUs * max_work_group * 128 rays, which on MI355X
// already exceeds any configured resolution, and HBM holds worst-case
// secondary bundles (2^bounces * pixels rays).
#include <chrono>

#include "rtx_impl.hpp"

namespace {
// Device counters of a frame (uint64): [0] secondary rays, [1] shadow rays.
__global__ void k_zero_u64(unsigned long long* p, int n) {
    if (threadIdx.x < (unsigned)n) p[threadIdx.x] = 0ull;
}
// Device copy of up to 5 u32 arrays (the render copy of a scene), grid-stride.
struct CopySegs {
    uint32_t* dst[5];
    const uint32_t* src[5];
    uint64_t n[5];
    int count;
};
__global__ __launch_bounds__(256) void k_copy_segs(CopySegs c) {
    for (int k = 0; k < c.count; ++k) {
        const uint64_t n = c.n[k];
        uint32_t* __restrict__ d = c.dst[k];
        const uint32_t* __restrict__ sp = c.src[k];
        if ((n & 3) == 0) {  // 16-byte moves (every buffer is 256-byte aligned)
            uint4* d4 = reinterpret_cast<uint4*>(d);
            const uint4* s4 = reinterpret_cast<const uint4*>(sp);
            for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n / 4; i += (uint64_t)gridDim.x * 256) d4[i] = s4[i];
        } else {
            for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) d[i] = sp[i];
        }
    }
}
}  // namespace

// Allocates render copy k (host-side; do it before frames run concurrently).
int32_t rtx_reserve_render_copy(rtx_scene* s, int k, bool shading_only) {
    RTX_REQUIRE(k >= 0 && k < rtx_scene::COPIES, RTX_EINVAL, "render copy: bad copy index");
    rtx_scene::RenderCopy& rc = s->rc[k];
    int32_t e;
    if ((e = rc.nrm.reserve(s->ctx, (uint64_t)s->V * 16)) || (e = rc.matmap.reserve(s->ctx, (uint64_t)s->T * 4)))
        return e;
    if (shading_only) return RTX_OK;
    if ((e = rc.node.reserve(s->ctx, s->b().node.bytes)) || (e = rc.delta.reserve(s->ctx, s->b().delta.bytes)) ||
        (e = rc.tri.reserve(s->ctx, (uint64_t)s->T * 48)))
        return e;
    return RTX_OK;
}

// Makes render copy k of what the render stages read, on `st` (shading only:
// the normals and material ids).
int32_t rtx_refresh_render_copy(rtx_scene* s, int k, hipStream_t st, bool shading_only) {
    RTX_REQUIRE(k >= 0 && k < rtx_scene::COPIES, RTX_EINVAL, "render copy: bad copy index");
    return rtx_copy_render_state(s, s->rc[k], st, shading_only);
}

// Copies what the render stages read (shading only: the normals and
// material ids) into rc, on `st`.
int32_t rtx_copy_render_state(rtx_scene* s, rtx_scene::RenderCopy& rc, hipStream_t st, bool shading_only) {
    int32_t e;
    DevBuf* src[5] = {&s->nrm, &s->matmap, &s->b().node, &s->b().delta, &s->b().tri};
    DevBuf* dst[5] = {&rc.nrm, &rc.matmap, &rc.node, &rc.delta, &rc.tri};
    const uint64_t bytes[5] = {(uint64_t)s->V * 16, (uint64_t)s->T * 4, s->b().node.bytes, s->b().delta.bytes,
                               (uint64_t)s->T * 48};
    const int n = shading_only ? 2 : 5;
    CopySegs c{};
    for (int j = 0; j < n; ++j) {
        RTX_REQUIRE(src[j]->p && src[j]->bytes >= bytes[j], RTX_ESTATE, "render copy: scene not built");
        if ((e = dst[j]->reserve(s->ctx, bytes[j]))) return e;
        c.dst[j] = dst[j]->as<uint32_t>();
        c.src[j] = src[j]->as<uint32_t>();
        c.n[j] = bytes[j] / 4;
    }
    c.count = n;
    ProfScope ps(s->ctx, st, "render_copy");
    hipLaunchKernelGGL(k_copy_segs, dim3(s->ctx->cus * 4), dim3(256), 0, st, c);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_renderer_create(rtx_ctx* ctx, const rtx_render_params* p, rtx_renderer** out) {
    RTX_REQUIRE(ctx && p && out, RTX_EINVAL, "rtx_renderer_create: null argument");
    RTX_REQUIRE(p->width > 0 && p->height > 0 && p->max_bounces >= 0 && p->max_bounces <= 16, RTX_EINVAL,
                "rtx_renderer_create: bad size or bounce count");
    rtx_renderer* r = new rtx_renderer();
    r->ctx = ctx;
    r->p = *p;
    if (r->p.quad_size <= 0) r->p.quad_size = 32;
    if (r->p.depth <= 0) r->p.depth = 30;
    if (r->p.min_leaf <= 0) r->p.min_leaf = 1;
    const uint64_t P = (uint64_t)p->width * p->height;
    int32_t e = 0;
    // A bounce holds at most twice the rays of the one before, so (P << b) + 2
    // slots never trigger the capacity drop (new_id + 2 >= cap).  A smaller
    // user capacity applies to every secondary bundle, as in the reference.
    for (int b = 0; b <= r->p.max_bounces && !e; ++b) {
        uint64_t cap = b == 0 ? P : (P << b) + 2;
        if (b > 0 && p->ray_capacity && p->ray_capacity < cap) cap = p->ray_capacity;
        if (cap >= (1ull << 30)) cap = (1ull << 30) - 1;
        rtx_rays* rr = nullptr;
        rtx_hits* hh = nullptr;
        if ((e = rtx_rays_create(ctx, cap, &rr))) break;
        r->rays.push_back(rr);
        if ((e = rtx_hits_create(ctx, cap, &hh))) break;
        r->hits.push_back(hh);
    }
    if (!e) e = rtx_fb_create(ctx, p->width, p->height, &r->fb);
    if (!e) e = r->counts.reserve(ctx, 64);
    for (int b = 0; b <= r->p.max_bounces && !e; ++b) {
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            rtx_set_error("rtx_renderer_create: event creation failed");
            e = RTX_EDEVICE;
            break;
        }
        r->ev.push_back(ev);
    }
    if (!e && hipEventCreateWithFlags(&r->ev_join, hipEventDisableTiming) != hipSuccess) {
        r->ev_join = nullptr;
        rtx_set_error("rtx_renderer_create: event creation failed");
        e = RTX_EDEVICE;
    }
    if (!e && (hipEventCreateWithFlags(&r->ev_fork, hipEventDisableTiming) != hipSuccess ||
               hipEventCreateWithFlags(&r->ev_built, hipEventDisableTiming) != hipSuccess)) {
        rtx_set_error("rtx_renderer_create: event creation failed");
        e = RTX_EDEVICE;
    }
    if (!e) {
        // the first primary generation uploads the pixel-sample table (a
        // blocking copy on the null stream, unordered with the non-blocking
        // streams frames run on): do it now, synchronously, not in a frame
        rtx_camera cam{};
        e = rtx_raygen_primary(r->rays[0], &cam, r->p.width, r->p.height, r->p.quad_size, 1, nullptr, 0, 0,
                               ctx->stream);
        if (!e && hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rtx_set_error("rtx_renderer_create: synchronization failed");
            e = RTX_EDEVICE;
        }
    }
    if (e) {
        rtx_renderer_destroy(r);
        return e;
    }
    *out = r;
    return RTX_OK;
}

int32_t rtx_renderer_destroy(rtx_renderer* r) {
    RTX_REQUIRE(r, RTX_EINVAL, "rtx_renderer_destroy: null");
    if (r->side) (void)hipStreamSynchronize(r->side);
    if (r->bstream) (void)hipStreamSynchronize(r->bstream);
    for (auto x : r->rays) rtx_rays_destroy(x);
    for (auto x : r->hits) rtx_hits_destroy(x);
    if (r->fb) rtx_fb_destroy(r->fb);
    if (r->graph_exec) (void)hipGraphExecDestroy(r->graph_exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    for (auto ev : r->ev) (void)hipEventDestroy(ev);
    if (r->ev_join) (void)hipEventDestroy(r->ev_join);
    if (r->side) (void)hipStreamDestroy(r->side);
    if (r->ev_fork) (void)hipEventDestroy(r->ev_fork);
    if (r->ev_built) (void)hipEventDestroy(r->ev_built);
    if (r->bstream) (void)hipStreamDestroy(r->bstream);
    r->counts.release();
    delete r;
    return RTX_OK;
}

rtx_fb* rtx_renderer_fb(rtx_renderer* r) { return r ? r->fb : nullptr; }

int32_t rtx_renderer_set_bounces(rtx_renderer* r, int32_t bounces) {
    RTX_REQUIRE(r, RTX_EINVAL, "rtx_renderer_set_bounces: null renderer");
    RTX_REQUIRE(bounces >= 0 && bounces + 1 <= (int32_t)r->rays.size(), RTX_EINVAL,
                "rtx_renderer_set_bounces: more bounces than the renderer was created for");
    r->p.max_bounces = bounces;
    return RTX_OK;
}

int32_t rtx_renderer_set_pipelined(rtx_renderer* r, int32_t enable) {
    RTX_REQUIRE(r, RTX_EINVAL, "rtx_renderer_set_pipelined: null renderer");
    r->pipelined = enable != 0;
    r->copy_valid = false;
    return RTX_OK;
}

int32_t rtx_renderer_set_animation(rtx_renderer* r, int32_t enable, float t0, float dt) {
    RTX_REQUIRE(r, RTX_EINVAL, "rtx_renderer_set_animation: null renderer");
    r->animate = enable != 0;
    r->anim_t0 = t0;
    r->anim_dt = dt;
    return RTX_OK;
}

// set_up_frame: counters, framebuffer clear, accelerator update.
static int32_t set_up_frame(rtx_renderer* r, rtx_scene* s, int32_t rebuild, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, r->counts.as<unsigned long long>(), 8);
    int32_t e;
    // the dynamic scene's displacement (src/rt/rt-wave.cpp:154-167), then the rebuild
    if (r->animate && (e = rtx_scene_animate(s, r->anim_t0, r->anim_dt, st))) return e;
    if ((e = rtx_fb_clear(r->fb, st))) return e;
    if (rebuild || !s->built) {
        rtx_lbvh_params bp{r->p.depth, r->p.min_leaf};
        if ((e = rtx_lbvh_build(s, &bp, st))) return e;
    }
    return RTX_OK;
}

// The reference's sequence (src/rt/renderer.cpp:113-214): per bounce trace,
// shadow, shade, then generate and read the new count back.
int32_t rtx_render_frame(rtx_renderer* r, rtx_scene* s, const rtx_camera* cam, int32_t rebuild,
                         rtx_frame_stats* stats, void* stream) {
    RTX_REQUIRE(r && s && cam, RTX_EINVAL, "rtx_render_frame: null argument");
    hipStream_t st = r->ctx->pick(stream);
    auto t0 = std::chrono::steady_clock::now();
    const uint64_t P = (uint64_t)r->p.width * r->p.height;
    unsigned long long* cnt = r->counts.as<unsigned long long>();
    int32_t e;
    if ((e = set_up_frame(r, s, rebuild, st))) return e;
    rtx_frame_stats fs{};
    fs.primary_rays = P;
    if ((e = rtx_raygen_primary(r->rays[0], cam, r->p.width, r->p.height, r->p.quad_size, 1, nullptr, 0, P, st)))
        return e;
    if ((e = rtx_trace(s, r->rays[0], r->hits[0], P, st))) return e;
    if ((e = rtx_trace_shadow_impl(s, r->rays[0], r->hits[0], P, (uint64_t*)(cnt + 1), st))) return e;
    if ((e = rtx_shade_impl(s, r->rays[0], r->hits[0], r->fb, P, 1, true, st))) return e;  // one ray a pixel
    uint64_t n = P;
    for (int b = 0; b < r->p.max_bounces; ++b) {
        rtx_rays *in = r->rays[b], *out = r->rays[b + 1];
        rtx_hits* hin = r->hits[b];
        if ((e = rtx_raygen_secondary_impl(s, in, hin, out, n, (uint64_t*)cnt, st))) return e;
        // the reference reads the new count back every bounce
        // (src/rt/secondary-ray-generator.cpp:296-299)
        uint32_t c32 = 0;
        RTX_CHECK_HIP(hipMemcpyAsync(&c32, out->count.p, 4, hipMemcpyDeviceToHost, st));
        RTX_CHECK_HIP(hipStreamSynchronize(st));
        const uint64_t m = c32;
        if (!m) break;
        fs.secondary_rays += m;
        fs.bounces_run++;
        rtx_hits* hout = r->hits[b + 1];
        if ((e = rtx_trace(s, out, hout, m, st))) return e;
        if ((e = rtx_trace_shadow_impl(s, out, hout, m, (uint64_t*)(cnt + 1), st))) return e;
        if ((e = rtx_shade(s, out, hout, r->fb, m, 0, st))) return e;
        n = m;
    }
    if ((e = rtx_fb_resolve(r->fb, st))) return e;
    unsigned long long hc[2] = {0, 0};
    RTX_CHECK_HIP(hipMemcpyAsync(hc, cnt, 16, hipMemcpyDeviceToHost, st));
    RTX_CHECK_HIP(hipStreamSynchronize(st));
    fs.shadow_rays = hc[1];
    auto t1 = std::chrono::steady_clock::now();
    fs.frame_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    r->last = fs;
    r->last_async = false;
    if (stats) *stats = fs;
    return RTX_OK;
}

// Renderer::render_to_framebuffer over a scene of objects (root_count() > 1):
// the reference's sequence with the multi-root trace and shadow kernels
// (src/rt/tracer.cpp:97, 223, 301, 363) and the shared material map of
// Scene::create_bvhs for secondary rays and shading.
int32_t rtx_render_instances(rtx_renderer* r, rtx_instances* in, const rtx_camera* cam, rtx_frame_stats* stats,
                             void* stream) {
    RTX_REQUIRE(r && in && cam, RTX_EINVAL, "rtx_render_instances: null argument");
    hipStream_t st = r->ctx->pick(stream);
    auto t0 = std::chrono::steady_clock::now();
    const uint64_t P = (uint64_t)r->p.width * r->p.height;
    unsigned long long* cnt = r->counts.as<unsigned long long>();
    rtx_scene* sh = nullptr;
    int32_t e;
    if ((e = rtx_instances_shading_scene(in, &sh))) return e;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, cnt, 8);
    if ((e = rtx_fb_clear(r->fb, st))) return e;
    rtx_frame_stats fs{};
    fs.primary_rays = P;
    if ((e = rtx_raygen_primary(r->rays[0], cam, r->p.width, r->p.height, r->p.quad_size, 1, nullptr, 0, P, st)))
        return e;
    if ((e = rtx_trace_instances(in, r->rays[0], r->hits[0], P, st))) return e;
    if ((e = rtx_trace_shadow_instances_impl(in, r->rays[0], r->hits[0], P, (uint64_t*)(cnt + 1), st))) return e;
    if ((e = rtx_shade_impl(sh, r->rays[0], r->hits[0], r->fb, P, 1, true, st))) return e;  // one ray a pixel
    uint64_t n = P;
    for (int b = 0; b < r->p.max_bounces; ++b) {
        rtx_rays *rin = r->rays[b], *rout = r->rays[b + 1];
        if ((e = rtx_raygen_secondary_impl(sh, rin, r->hits[b], rout, n, (uint64_t*)cnt, st))) return e;
        uint32_t c32 = 0;
        RTX_CHECK_HIP(hipMemcpyAsync(&c32, rout->count.p, 4, hipMemcpyDeviceToHost, st));
        RTX_CHECK_HIP(hipStreamSynchronize(st));
        const uint64_t m = c32;
        if (!m) break;
        fs.secondary_rays += m;
        fs.bounces_run++;
        rtx_hits* hout = r->hits[b + 1];
        if ((e = rtx_trace_instances(in, rout, hout, m, st))) return e;
        if ((e = rtx_trace_shadow_instances_impl(in, rout, hout, m, (uint64_t*)(cnt + 1), st))) return e;
        if ((e = rtx_shade(sh, rout, hout, r->fb, m, 0, st))) return e;
        n = m;
    }
    if ((e = rtx_fb_resolve(r->fb, st))) return e;
    unsigned long long hc[2] = {0, 0};
    RTX_CHECK_HIP(hipMemcpyAsync(hc, cnt, 16, hipMemcpyDeviceToHost, st));
    RTX_CHECK_HIP(hipStreamSynchronize(st));
    fs.shadow_rays = hc[1];
    fs.frame_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    r->last = fs;
    r->last_async = false;
    if (stats) *stats = fs;
    return RTX_OK;
}

// Same frame with no host read-back: counts stay on the device.  Bounce b's
// rays depend only on bounce b-1's hits, and its shadow rays and shading on
// nothing later, so after bounce b is traced the main stream goes on to
// generate and trace bounce b+1 while the side stream traces bounce b's
// shadow rays and shades it (a latency-bound small trace then shares the GPU
// with independent work).  Results are identical to the reference sequence:
// shadow rays only set flag bits 8+ that ray generation does not read, and
// shading is an order-free integer sum.
// Streams are created on first use: every stream takes one of the process's
// hardware queues (round-robin at creation), and streams that share a queue
// lose their concurrency.
static int32_t ensure_stream(hipStream_t* st, const char* what) {
    if (*st) return RTX_OK;
    if (hipStreamCreateWithFlags(st, hipStreamNonBlocking) != hipSuccess) {
        *st = nullptr;
        rtx_set_error(std::string("renderer: ") + what + " stream creation failed");
        return RTX_EDEVICE;
    }
    return RTX_OK;
}

// Render stages of the read-back-free frame (after set-up): counts stay on
// the device.  Bounce b's rays depend only on bounce b-1's hits, and its
// shadow rays and shading on nothing later, so after bounce b is traced the
// main stream goes on to generate and trace bounce b+1 while the side stream
// traces bounce b's shadow rays and shades it (a latency-bound small trace
// then shares the GPU with independent work).  Results are identical to the
// reference sequence: shadow rays only set flag bits 8+ that ray generation
// does not read, and shading is an order-free integer sum.
static int32_t render_stages(rtx_renderer* r, rtx_scene* s, const rtx_camera* cam, hipStream_t st) {
    const bool two = !r->single_stream && r->ctx->tune.side_stream != 0;
    int32_t e;
    if (two && (e = ensure_stream(&r->side, "side"))) return e;
    hipStream_t sd = two ? r->side : st;
    const uint64_t P = (uint64_t)r->p.width * r->p.height;
    unsigned long long* cnt = r->counts.as<unsigned long long>();
    ProfScope whole(r->ctx, st, "FRAME_RENDER");
    const int B = r->p.max_bounces;
    if ((e = rtx_raygen_primary(r->rays[0], cam, r->p.width, r->p.height, r->p.quad_size, 1, nullptr, 0, P, st)))
        return e;
    for (int b = 0; b <= B; ++b) {
        const uint64_t nb = b == 0 ? P : RTX_DEVICE_COUNT;
        if (b > 0) {
            const uint64_t n_in = b == 1 ? P : RTX_DEVICE_COUNT;
            if ((e = rtx_raygen_secondary_impl(s, r->rays[b - 1], r->hits[b - 1], r->rays[b], n_in, (uint64_t*)cnt,
                                               st)))
                return e;
        }
        // zero-count bounces run as empty launches (no host decision)
        if ((e = rtx_trace(s, r->rays[b], r->hits[b], nb, st))) return e;
        if (two) {
            RTX_CHECK_HIP(hipEventRecord(r->ev[b], st));
            RTX_CHECK_HIP(hipStreamWaitEvent(sd, r->ev[b], 0));
        }
        if ((e = rtx_trace_shadow_impl(s, r->rays[b], r->hits[b], nb, (uint64_t*)(cnt + 1), sd))) return e;
        // the primary rays have one sample a pixel: no two add to one pixel
        if ((e = rtx_shade_impl(s, r->rays[b], r->hits[b], r->fb, nb, b == 0, b == 0, sd))) return e;
    }
    if (two) {
        RTX_CHECK_HIP(hipEventRecord(r->ev_join, sd));
        RTX_CHECK_HIP(hipStreamWaitEvent(st, r->ev_join, 0));
    }
    return rtx_fb_resolve(r->fb, st);
}

// Pipelined frame (the reference's RendererT, src/rt/renderer-threaded.cpp:
// 452-506: the next scene's accelerator is built on a second queue while the
// last one renders).  The build stream animates and rebuilds the scene for
// the next frame while the main stream renders this frame from the render
// copy made at the end of the previous frame; after both, the copy is
// refreshed.  The image of a frame is therefore the scene as built one frame
// earlier.  The first pipelined frame builds and copies first.
static int32_t render_frame_pipelined(rtx_renderer* r, rtx_scene* s, const rtx_camera* cam, int32_t rebuild,
                                      hipStream_t st) {
    int32_t e;
    const rtx_lbvh_params bp{r->p.depth, r->p.min_leaf};
    if ((e = ensure_stream(&r->bstream, "build"))) return e;
    if (!r->copy_valid || !s->built || s != r->copy_of) {
        if (r->animate && (e = rtx_scene_animate(s, r->anim_t0, r->anim_dt, st))) return e;
        if ((e = rtx_lbvh_build(s, &bp, st))) return e;
        if ((e = rtx_refresh_render_copy(s, 0, st))) return e;
        r->copy_valid = true;
        r->copy_of = s;
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, r->counts.as<unsigned long long>(), 8);
    if ((e = rtx_fb_clear(r->fb, st))) return e;
    RTX_CHECK_HIP(hipEventRecord(r->ev_fork, st));
    RTX_CHECK_HIP(hipStreamWaitEvent(r->bstream, r->ev_fork, 0));
    if (r->animate && (e = rtx_scene_animate(s, r->anim_t0, r->anim_dt, r->bstream))) return e;
    if (rebuild && (e = rtx_lbvh_build(s, &bp, r->bstream))) return e;
    RTX_CHECK_HIP(hipEventRecord(r->ev_built, r->bstream));
    s->use_copy = 0;
    e = render_stages(r, s, cam, st);
    s->use_copy = -1;
    if (e) return e;
    RTX_CHECK_HIP(hipStreamWaitEvent(st, r->ev_built, 0));
    return rtx_refresh_render_copy(s, 0, st);
}

int32_t rtx_render_frame_async(rtx_renderer* r, rtx_scene* s, const rtx_camera* cam, int32_t rebuild, void* stream) {
    RTX_REQUIRE(r && s && cam, RTX_EINVAL, "rtx_render_frame_async: null argument");
    hipStream_t st = r->ctx->pick(stream);
    int32_t e;
    if (r->pipelined) {
        e = render_frame_pipelined(r, s, cam, rebuild, st);
    } else {
        if ((e = set_up_frame(r, s, rebuild, st))) return e;
        e = render_stages(r, s, cam, st);
    }
    if (e) return e;
    r->last_async = true;
    r->last = rtx_frame_stats{};
    r->last.primary_rays = (uint64_t)r->p.width * r->p.height;
    return RTX_OK;
}

int32_t rtx_renderer_last_stats(rtx_renderer* r, rtx_frame_stats* stats) {
    RTX_REQUIRE(r && stats, RTX_EINVAL, "rtx_renderer_last_stats: null argument");
    if (r->last_async) {
        unsigned long long hc[2] = {0, 0};
        RTX_CHECK_HIP(hipStreamSynchronize(r->ctx->stream));
        RTX_CHECK_HIP(hipDeviceSynchronize());
        RTX_CHECK_HIP(hipMemcpy(hc, r->counts.p, 16, hipMemcpyDeviceToHost));
        r->last.secondary_rays = hc[0];
        r->last.shadow_rays = hc[1];
        r->last_async = false;
    }
    *stats = r->last;
    return RTX_OK;
}

// Captures rtx_render_frame_async into a hipGraph (one launch per frame on
// replay).  Every buffer must already exist: render one frame eagerly first.
// The camera, scene and parameters are fixed at capture; per-frame scene
// changes (vertex updates, displacement) are enqueued before each replay.
int32_t rtx_renderer_capture(rtx_renderer* r, rtx_scene* s, const rtx_camera* cam, int32_t rebuild, void* stream) {
    RTX_REQUIRE(r && s && cam, RTX_EINVAL, "rtx_renderer_capture: null argument");
    hipStream_t st = r->ctx->pick(stream);
    if (r->graph_exec) {
        (void)hipGraphExecDestroy(r->graph_exec);
        r->graph_exec = nullptr;
    }
    if (r->graph) {
        (void)hipGraphDestroy(r->graph);
        r->graph = nullptr;
    }
    const bool prof = r->ctx->profiling;
    r->ctx->profiling = false;
    RTX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int32_t e = rtx_render_frame_async(r, s, cam, rebuild, st);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(st, &g);
    r->ctx->profiling = prof;
    if (e) {
        if (g) (void)hipGraphDestroy(g);
        return e;
    }
    if (ce != hipSuccess) {
        rtx_set_error(std::string("rtx_renderer_capture: ") + hipGetErrorString(ce));
        return RTX_EDEVICE;
    }
    r->graph = g;
    RTX_CHECK_HIP(hipGraphInstantiate(&r->graph_exec, g, nullptr, nullptr, 0));
    return RTX_OK;
}

int32_t rtx_renderer_replay(rtx_renderer* r, void* stream) {
    RTX_REQUIRE(r, RTX_EINVAL, "rtx_renderer_replay: null renderer");
    RTX_REQUIRE(r->graph_exec, RTX_ESTATE, "rtx_renderer_replay: nothing captured");
    hipStream_t st = r->ctx->pick(stream);
    ProfScope ps(r->ctx, st, "FRAME_GRAPH");
    RTX_CHECK_HIP(hipGraphLaunch(r->graph_exec, st));
    r->last_async = true;
    return RTX_OK;
}
