// Wavefront renderer: frames of many scenes in flight together, advanced
// one bounce per step, each stage of a step one launch over every frame.
//
// The reference renders one frame of one scene at a time, a launch per
// stage per bounce (Renderer::render_to_framebuffer,
// src/rt/renderer.cpp:113-214); its only overlap is RendererT's build of the
// next scene beside the current render (src/rt/renderer-threaded.cpp:
// 452-506).  Small frames leave most of an MI355X idle, and each bounce
// launch lasts as long as its slowest waves.  Here a step runs, for every
// admitted frame, the next bounce of the reference's sequence: primary rays
// (new frames), closest hit, shadow rays, shading, secondary rays, and the
// resolve of frames that end.  Each stage is one launch over the
// concatenation of the frames' items, cut into chunks that stay inside one
// frame (so a wave reads one scene's BVH), listed grouped by scene;
// traversal waves take chunks from a queue until it drains, so a frame's
// slow chunks overlap other frames' work instead of ending a launch.  Every
// ray runs the single-frame kernels' code, so each frame's image, hit
// records and counts equal its stand-alone frame's.
//
// A frame's rays, hits, framebuffer and counters are carved from one device
// arena at admission (arena.hpp) and returned when the caller releases the
// frame: no allocation while frames run.  Counts stay on the device; the
// host knows each frame's schedule (bounce b at its b-th step), so steps are
// enqueued without read-backs, and a frame is done when the event of its
// last step has completed.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <functional>

#include "arena.hpp"
#include "rtx_impl.hpp"

namespace {

constexpr int RING = 4;  // steps whose tables may be in flight at once

enum FrameState { F_FREE = 0, F_ACTIVE, F_LAST_ENQUEUED, F_DONE };

struct WfFrame {
    int state = F_FREE;
    uint64_t tag = 0;
    rtx_scene* scene = nullptr;
    rtx_camera cam{};
    int w = 0, h = 0, bounces = 0;
    uint64_t user_cap = 0;   // secondary-ray capacity per bounce (0: the full (P << b) + 2)
    rtxwf::FrameLayout L;
    uint64_t off = 0;        // region in the arena
    int next = 0;            // bounce run by the next step
    int64_t done_step = -1;  // step after which the frame is complete
    rtx_frame_stats stats{};
};

struct StepSlot {
    hipEvent_t ev = nullptr;
    int64_t step = -1;           // step that last used this slot
    std::vector<int> finishing;  // frames that step completes, in fin order
};


}  // namespace

struct rtx_wavefront {
    rtx_ctx* ctx = nullptr;
    int max_frames = 0, quad = 32;
    DevBuf arena_mem;
    rtxwf::Arena arena;
    std::vector<WfFrame> f;
    DevBuf dtab, dctr, dfin;  // per ring slot: segment table, queue counters, counters of finishing frames
    WfSeg* htab = nullptr;               // pinned staging of the tables
    unsigned long long* hfin = nullptr;  // pinned copy of the finishing frames' counters
    StepSlot slot[RING];
    int64_t steps = 0;      // steps enqueued
    int64_t observed = -1;  // newest step seen complete
    int auto_share = 100;   // persistent traversal share when the wf_share tune is 0
    std::deque<int> done_q; // frames complete and not yet returned by poll
};

static char* frame_base(rtx_wavefront* wf, const WfFrame& fr) { return static_cast<char*>(wf->arena_mem.p) + fr.off; }

template <class T>
static T* at(char* base, uint64_t off) {
    return reinterpret_cast<T*>(base + off);
}

// Handles completed steps up to `upto`, oldest first (blocking: waits for
// them): frames finished by a step take their counters and become done.
static int32_t observe(rtx_wavefront* wf, int64_t upto, bool blocking) {
    for (int64_t s = wf->observed + 1; s <= upto && s < wf->steps; ++s) {
        StepSlot& sl = wf->slot[s % RING];
        const hipError_t q = blocking ? hipEventSynchronize(sl.ev) : hipEventQuery(sl.ev);
        if (q == hipErrorNotReady) return RTX_OK;
        if (q != hipSuccess) {
            rtx_set_error(std::string("wavefront: step failed: ") + hipGetErrorString(q));
            return RTX_EDEVICE;
        }
        for (size_t j = 0; j < sl.finishing.size(); ++j) {
            WfFrame& fr = wf->f[sl.finishing[j]];
            const unsigned long long* c = wf->hfin + ((size_t)(s % RING) * RTX_WF_MAX_SEGS + j) * 4;
            fr.stats.primary_rays = fr.L.P;
            fr.stats.secondary_rays = c[0];
            fr.stats.shadow_rays = c[1];
            fr.stats.bounces_run = (uint32_t)c[2];
            fr.state = F_DONE;
            wf->done_q.push_back(sl.finishing[j]);
        }
        wf->observed = s;
    }
    return RTX_OK;
}

uint64_t rtx_wavefront_frame_bytes(int32_t width, int32_t height, int32_t bounces, uint64_t ray_capacity) {
    if (width <= 0 || height <= 0 || bounces < 0 || bounces > 16) return 0;
    return rtxwf::align_up(rtxwf::FrameLayout((uint64_t)width * height, bounces, ray_capacity).bytes);
}

int32_t rtx_wavefront_create(rtx_ctx* ctx, const rtx_wavefront_params* p, rtx_wavefront** out) {
    RTX_REQUIRE(ctx && p && out, RTX_EINVAL, "rtx_wavefront_create: null argument");
    RTX_REQUIRE(p->max_frames >= 1 && p->max_frames <= RTX_WF_MAX_SEGS, RTX_EINVAL,
                "rtx_wavefront_create: max_frames must be in [1, 256]");
    const uint64_t bytes = p->arena_bytes ? p->arena_bytes : (8ull << 30);
    rtx_wavefront* wf = new rtx_wavefront();
    wf->ctx = ctx;
    wf->max_frames = p->max_frames;
    wf->quad = p->quad_size > 0 ? p->quad_size : 32;
    wf->f.resize(p->max_frames);
    int32_t e = RTX_OK;
    if ((e = wf->arena_mem.reserve(ctx, bytes)) || (e = wf->dtab.reserve(ctx, sizeof(WfSeg) * RTX_WF_MAX_SEGS * RING)) ||
        (e = wf->dctr.reserve(ctx, (size_t)WF_Q_COUNT * 4 * RING)) || (e = wf->dfin.reserve(ctx, 32 * RTX_WF_MAX_SEGS * RING))) {
        rtx_wavefront_destroy(wf);
        return e;
    }
    wf->arena.reset(bytes);
    if (hipHostMalloc((void**)&wf->htab, sizeof(WfSeg) * RTX_WF_MAX_SEGS * RING, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void**)&wf->hfin, 32 * RTX_WF_MAX_SEGS * RING, hipHostMallocDefault) != hipSuccess) {
        wf->htab = nullptr;
        rtx_set_error("rtx_wavefront_create: pinned host allocation failed");
        rtx_wavefront_destroy(wf);
        return RTX_ENOMEM;
    }
    for (auto& sl : wf->slot)
        if (hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
            sl.ev = nullptr;
            rtx_set_error("rtx_wavefront_create: event creation failed");
            rtx_wavefront_destroy(wf);
            return RTX_EDEVICE;
        }
    *out = wf;
    return RTX_OK;
}

int32_t rtx_wavefront_destroy(rtx_wavefront* wf) {
    RTX_REQUIRE(wf, RTX_EINVAL, "rtx_wavefront_destroy: null");
    for (auto& sl : wf->slot)
        if (sl.ev) {
            (void)hipEventSynchronize(sl.ev);
            (void)hipEventDestroy(sl.ev);
        }
    if (wf->htab) (void)hipHostFree(wf->htab);
    if (wf->hfin) (void)hipHostFree(wf->hfin);
    wf->arena_mem.release();
    wf->dtab.release();
    wf->dctr.release();
    wf->dfin.release();
    delete wf;
    return RTX_OK;
}

int32_t rtx_wavefront_add(rtx_wavefront* wf, rtx_scene* s, const rtx_camera* cam, int32_t width, int32_t height,
                          int32_t bounces, uint64_t ray_capacity, uint64_t tag, int32_t* frame) {
    RTX_REQUIRE(wf && s && cam && frame, RTX_EINVAL, "rtx_wavefront_add: null argument");
    RTX_REQUIRE(width > 0 && height > 0 && (uint64_t)width * height < (1ull << 30), RTX_EINVAL,
                "rtx_wavefront_add: bad frame size");
    RTX_REQUIRE(bounces >= 0 && bounces <= 16, RTX_EINVAL, "rtx_wavefront_add: bounces must be in [0, 16]");
    RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_wavefront_add: scene has no BVH (build it first)");
    RTX_REQUIRE(s->mats.p && s->lights.p, RTX_ESTATE, "rtx_wavefront_add: scene has no materials or lights");
    RTX_REQUIRE(s->use_set < 0 && s->use_copy < 0, RTX_ESTATE, "rtx_wavefront_add: scene is inside a frame pipeline");
    const rtxwf::FrameLayout L((uint64_t)width * height, bounces, ray_capacity);
    RTX_REQUIRE(rtxwf::align_up(L.bytes) <= wf->arena.size(), RTX_EINVAL,
                "rtx_wavefront_add: the frame needs more memory than the whole arena");
    int id = -1;
    for (int i = 0; i < wf->max_frames; ++i)
        if (wf->f[i].state == F_FREE) {
            id = i;
            break;
        }
    RTX_REQUIRE(id >= 0, RTX_ENOMEM, "rtx_wavefront_add: max_frames frames are in flight or unreleased");
    const uint64_t off = wf->arena.alloc(L.bytes);
    RTX_REQUIRE(off != rtxwf::NONE, RTX_ENOMEM, "rtx_wavefront_add: the arena has no room now (release frames)");
    WfFrame& fr = wf->f[id];
    fr = WfFrame{};
    fr.state = F_ACTIVE;
    fr.tag = tag;
    fr.scene = s;
    fr.cam = *cam;
    fr.w = width;
    fr.h = height;
    fr.bounces = bounces;
    fr.user_cap = ray_capacity;
    fr.L = L;
    fr.off = off;
    *frame = id;
    return RTX_OK;
}

// The segment of frame `fr` for its next bounce.
static WfSeg make_seg(rtx_wavefront* wf, const WfFrame& fr, uint32_t fin) {
    rtx_scene* s = fr.scene;
    const SceneView sv = scene_view(s);
    const rtxwf::FrameLayout& L = fr.L;
    char* b = frame_base(wf, fr);
    const int bn = fr.next, in = bn & 1, out = in ^ 1;
    WfSeg g{};
    g.node = sv.node;
    g.delta = sv.delta;
    g.tri = sv.tri;
    g.nrm = sv.nrm;
    g.uv = s->uv.as<float2>();
    g.mats = s->mats.as<rtx_material>();
    g.matmap = sv.matmap;
    g.lights = s->lights.as<rtx_lights>();
    g.tex = s->tex.p ? s->tex.as<uchar4>() : nullptr;
    g.tex_tab = s->tex_tab.as<TexEnt>();
    g.ntex = s->ntex;
    g.cube = s->cube ? 1 : 0;
    g.ori4 = at<float4>(b, L.ray[in][0]);
    g.dir4 = at<float4>(b, L.ray[in][1]);
    g.pixel = at<int32_t>(b, L.ray[in][2]);
    g.contrib = at<float>(b, L.ray[in][3]);
    g.count = at<uint32_t>(b, L.count[in]);
    g.o_ori4 = at<float4>(b, L.ray[out][0]);
    g.o_dir4 = at<float4>(b, L.ray[out][1]);
    g.o_pixel = at<int32_t>(b, L.ray[out][2]);
    g.o_contrib = at<float>(b, L.ray[out][3]);
    g.o_count = at<uint32_t>(b, L.count[out]);
    g.status = at<uint32_t>(b, L.status);
    g.flags = at<uint32_t>(b, L.flags);
    g.t = at<float>(b, L.t);
    g.id = at<int32_t>(b, L.id);
    g.bary = at<float2>(b, L.bary);
    g.n4 = at<float4>(b, L.n4);
    g.p4 = at<float4>(b, L.p4);
    g.uv2 = at<float2>(b, L.uv2);
    g.list = at<uint32_t>(b, L.list);
    g.list_count = at<uint32_t>(b, L.list_count);
    g.accum = at<uint32_t>(b, L.accum);
    g.rgba = at<uchar4>(b, L.rgba);
    g.stats = at<unsigned long long>(b, L.stats);
    g.cam = fr.cam;
    g.w = fr.w;
    g.h = fr.h;
    g.quad = wf->quad;
    g.o_cap = (uint32_t)rtxwf::bounce_capacity(L.P, bn + 1, fr.user_cap);
    g.primary = bn == 0;
    g.gen = bn < fr.bounces;
    g.resolve = bn == fr.bounces;
    g.fin = fin;
    return g;
}

// One step: every active frame runs its next bounce.  `finishing` (optional)
// receives the frames this step completes.
int32_t rtx_wavefront_step_impl(rtx_wavefront* wf, hipStream_t st, int32_t* pending, std::vector<int>* finishing,
                                const std::function<int32_t(hipStream_t, const std::vector<int>&)>& before_done) {
    std::vector<int> act;
    for (int i = 0; i < wf->max_frames; ++i)
        if (wf->f[i].state == F_ACTIVE) act.push_back(i);
    if (finishing) finishing->clear();
    if (act.empty()) {
        if (pending) *pending = 0;
        return RTX_OK;
    }
    // frames of one scene next to each other: consecutive chunks read one BVH
    std::stable_sort(act.begin(), act.end(), [wf](int a, int b) { return wf->f[a].scene < wf->f[b].scene; });
    const int k = (int)(wf->steps % RING);
    StepSlot& sl = wf->slot[k];
    if (sl.step >= 0) {  // the slot's last step must be complete and its counters taken
        int32_t e = observe(wf, sl.step, true);
        if (e) return e;
    }
    WfSeg* T = wf->htab + (size_t)k * RTX_WF_MAX_SEGS;
    uint64_t px_new = 0, px_done = 0, max_chunks = 0, max_pair_chunks = 0;
    bool gen = false, deep = false;
    std::vector<int> fin;
    const uint32_t n = (uint32_t)act.size();
    for (uint32_t i = 0; i < n; ++i) {
        const WfFrame& fr = wf->f[act[i]];
        const bool last = fr.next == fr.bounces;
        T[i] = make_seg(wf, fr, last ? (uint32_t)fin.size() : 0u);
        if (fr.next == 0) px_new += fr.L.P;
        const uint64_t cap_in = rtxwf::bounce_capacity(fr.L.P, fr.next, fr.user_cap);
        max_chunks += (cap_in + RTX_WAVE - 1) / RTX_WAVE;
        max_pair_chunks += (cap_in * (uint64_t)fr.scene->host_lights.count + RTX_WAVE - 1) / RTX_WAVE;
        if (last) {
            px_done += fr.L.P;
            fin.push_back(act[i]);
        } else {
            gen = true;
        }
        if (fr.scene->depth >= 32) deep = true;
    }
    rtx_ctx* c = wf->ctx;
    WfSeg* D = wf->dtab.as<WfSeg>() + (size_t)k * RTX_WF_MAX_SEGS;
    uint32_t* ctr = wf->dctr.as<uint32_t>() + (size_t)WF_Q_COUNT * k;
    unsigned long long* dfin = wf->dfin.as<unsigned long long>() + (size_t)k * RTX_WF_MAX_SEGS * 4;
    // the traversal form: every form gives the same records; the sibling-pair
    // form cannot follow the MAX_LEVELS rule of trees of 32 levels, and the
    // persistent and stackless forms have no segmented kernels
    int var = (int)c->tune.trace_variant;
    if (var > 2 || (var == 2 && deep)) var = 1;
    const int tb = (int)c->tune.trace_block;
    const uint32_t coop = c->tune.coop_leaf > 0 ? (uint32_t)c->tune.coop_leaf : 0xffffffffu;
    const int share = c->tune.wf_share > 0 ? (int)c->tune.wf_share : wf->auto_share;
    RTX_CHECK_HIP(hipMemcpyAsync(D, T, sizeof(WfSeg) * n, hipMemcpyHostToDevice, st));
    {
        ProfScope ps(c, st, "wf_init");
        rtx_wf_init(D, n, ctr, st);
    }
    if (px_new) {
        ProfScope ps(c, st, "wf_raygen");
        rtx_wf_raygen(D, n, px_new, c->cus, st);
    }
    {
        ProfScope ps(c, st, "wf_trace");
        rtx_wf_trace(D, n, ctr, var, tb, (int)c->tune.stack_lds, coop, (uint32_t)c->tune.wf_grain,
                     (int)c->tune.wf_sched, max_chunks, c->cus, share, st);
    }
    {
        ProfScope ps(c, st, "wf_shadow");
        // shadow rays in any child order unless a scene's tree has 32 levels
        rtx_wf_shadow(D, n, ctr, var, tb, (int)c->tune.stack_lds, coop, (uint32_t)c->tune.wf_grain,
                      (int)c->tune.wf_sched, max_pair_chunks, c->cus, share, c->tune.shadow_order != 0 || deep,
                      st);
    }
    {
        ProfScope ps(c, st, "wf_shade");
        rtx_wf_shade(D, n, c->cus, st);
    }
    if (gen) {
        ProfScope ps(c, st, "wf_secondary");
        rtx_wf_secondary(D, n, ctr, c->cus, st);
    }
    if (px_done) {
        ProfScope ps(c, st, "wf_resolve");
        rtx_wf_resolve(D, n, px_done, dfin, c->cus, st);
        RTX_CHECK_HIP(hipMemcpyAsync(wf->hfin + (size_t)k * RTX_WF_MAX_SEGS * 4, dfin, 32 * fin.size(),
                                     hipMemcpyDeviceToHost, st));
    }
    RTX_CHECK_HIP(hipGetLastError());
    if (before_done && !fin.empty()) {  // the caller's copies of the finishing frames, covered by the step's event
        int32_t e = before_done(st, fin);
        if (e) return e;
    }
    RTX_CHECK_HIP(hipEventRecord(sl.ev, st));
    sl.step = wf->steps;
    sl.finishing = fin;
    for (int id : act) {
        WfFrame& fr = wf->f[id];
        if (fr.next == fr.bounces) {
            fr.state = F_LAST_ENQUEUED;
            fr.done_step = wf->steps;
        } else {
            fr.next++;
        }
    }
    ++wf->steps;
    if (pending) {
        int32_t p = 0;
        for (auto& fr : wf->f) p += fr.state == F_ACTIVE;
        *pending = p;
    }
    if (finishing) *finishing = fin;
    return RTX_OK;
}

int32_t rtx_wavefront_step(rtx_wavefront* wf, void* stream, int32_t* pending) {
    RTX_REQUIRE(wf, RTX_EINVAL, "rtx_wavefront_step: null");
    return rtx_wavefront_step_impl(wf, wf->ctx->pick(stream), pending, nullptr, nullptr);
}

int32_t rtx_wavefront_run(rtx_wavefront* wf, void* stream) {
    RTX_REQUIRE(wf, RTX_EINVAL, "rtx_wavefront_run: null");
    hipStream_t st = wf->ctx->pick(stream);
    for (int32_t pending = 1; pending > 0;) {
        int32_t e = rtx_wavefront_step_impl(wf, st, &pending, nullptr, nullptr);
        if (e) return e;
    }
    return RTX_OK;
}

int32_t rtx_wavefront_sync(rtx_wavefront* wf) {
    RTX_REQUIRE(wf, RTX_EINVAL, "rtx_wavefront_sync: null");
    return observe(wf, wf->steps - 1, true);
}

int32_t rtx_wavefront_poll(rtx_wavefront* wf, int32_t* frames, uint64_t* tags, uint32_t cap, uint32_t* n) {
    RTX_REQUIRE(wf && n && (frames || cap == 0), RTX_EINVAL, "rtx_wavefront_poll: null argument");
    int32_t e = observe(wf, wf->steps - 1, false);
    if (e) return e;
    uint32_t k = 0;
    while (k < cap && !wf->done_q.empty()) {
        const int id = wf->done_q.front();
        wf->done_q.pop_front();
        frames[k] = id;
        if (tags) tags[k] = wf->f[id].tag;
        ++k;
    }
    *n = k;
    return RTX_OK;
}

// A frame whose last step is enqueued, waited for.
static int32_t finished_frame(rtx_wavefront* wf, int32_t frame, const char* what, WfFrame** out) {
    RTX_REQUIRE(wf && frame >= 0 && frame < wf->max_frames, RTX_EINVAL, std::string(what) + ": bad frame");
    WfFrame& fr = wf->f[frame];
    RTX_REQUIRE(fr.state == F_LAST_ENQUEUED || fr.state == F_DONE, RTX_ESTATE,
                std::string(what) + ": the frame has not run its last bounce (step or run first)");
    if (fr.state == F_LAST_ENQUEUED) {
        int32_t e = observe(wf, fr.done_step, true);
        if (e) return e;
    }
    *out = &fr;
    return RTX_OK;
}

int32_t rtx_wavefront_frame_stats(rtx_wavefront* wf, int32_t frame, rtx_frame_stats* stats) {
    WfFrame* fr = nullptr;
    int32_t e = finished_frame(wf, frame, "rtx_wavefront_frame_stats", &fr);
    if (e) return e;
    if (stats) {
        *stats = fr->stats;
        stats->node_count = 0;
    }
    return RTX_OK;
}

int32_t rtx_wavefront_read_accum(rtx_wavefront* wf, int32_t frame, uint32_t* rgb3) {
    WfFrame* fr = nullptr;
    int32_t e = finished_frame(wf, frame, "rtx_wavefront_read_accum", &fr);
    if (e) return e;
    RTX_REQUIRE(rgb3, RTX_EINVAL, "rtx_wavefront_read_accum: null output");
    RTX_CHECK_HIP(hipMemcpy(rgb3, frame_base(wf, *fr) + fr->L.accum, fr->L.P * 12, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_wavefront_read_rgba8(rtx_wavefront* wf, int32_t frame, uint8_t* rgba) {
    WfFrame* fr = nullptr;
    int32_t e = finished_frame(wf, frame, "rtx_wavefront_read_rgba8", &fr);
    if (e) return e;
    RTX_REQUIRE(rgba, RTX_EINVAL, "rtx_wavefront_read_rgba8: null output");
    RTX_CHECK_HIP(hipMemcpy(rgba, frame_base(wf, *fr) + fr->L.rgba, fr->L.P * 4, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_wavefront_release(rtx_wavefront* wf, int32_t frame) {
    RTX_REQUIRE(wf && frame >= 0 && frame < wf->max_frames, RTX_EINVAL, "rtx_wavefront_release: bad frame");
    WfFrame& fr = wf->f[frame];
    RTX_REQUIRE(fr.state == F_DONE, RTX_ESTATE, "rtx_wavefront_release: the frame is not complete (poll or read it)");
    wf->arena.release(fr.off);
    for (auto it = wf->done_q.begin(); it != wf->done_q.end(); ++it)
        if (*it == frame) {
            wf->done_q.erase(it);
            break;
        }
    fr = WfFrame{};
    return RTX_OK;
}

int32_t rtx_wavefront_memory(rtx_wavefront* wf, uint64_t* used, uint64_t* capacity) {
    RTX_REQUIRE(wf, RTX_EINVAL, "rtx_wavefront_memory: null");
    if (used) *used = wf->arena.used();
    if (capacity) *capacity = wf->arena.size();
    return RTX_OK;
}

// ------------------------------------------------- for the request server --
uint64_t rtx_wf_arena_size(const rtx_wavefront* wf) { return wf->arena.size(); }
int64_t rtx_wf_steps_in_flight(const rtx_wavefront* wf) { return wf->steps - 1 - wf->observed; }

void rtx_wf_set_auto_share(rtx_wavefront* wf, int percent) { wf->auto_share = std::min(100, std::max(1, percent)); }
const void* rtx_wf_accum_dev(rtx_wavefront* wf, int32_t frame) {
    return frame_base(wf, wf->f[frame]) + wf->f[frame].L.accum;
}
