// Request server over resident scenes: open-loop arrivals, continuous
// batching on the wavefront renderer, HBM-budget admission, device-side
// completion times.  No counterpart in the reference, whose apps render one
// frame of one scene from one thread (src/rt/renderer.cpp:113-214).
//
// A dispatcher thread owns every GPU call of the server.  Requests whose
// arrival time has come join a first-come-first-served admission queue
// (arena.hpp: Admission): a request is admitted when its rays, hits and
// framebuffer (rtx_wavefront_frame_bytes) fit the request arena and fewer
// than `workers` requests are in flight; otherwise it waits at the head, and
// one larger than the whole arena is refused.  While requests are in flight
// the dispatcher keeps up to `steps_ahead` wavefront steps enqueued per group: each
// step advances every request of the group by one bounce in shared
// launches, so requests that arrive while others run join at the next step,
// whatever their scene, size or depth, up to `srv_step_pixels` new primary
// pixels a group's step (the others wait at the head for the next step, so
// that a burst of arrivals does not lengthen one step for every request in
// it).  Requests are spread over `groups`
// wavefront renderers, each stepping on a stream of its own with its share
// of the arena and of the requests in flight, so that the launch tails of
// one group overlap the work of the others.  A request's latency is done_ns - arrival_ns,
// where done_ns is when the dispatcher observed the event of its last step
// (accumulators copied out included).
//
// `plain` mode is the per-request baseline: one request at a time with the
// reference's frame and its per-bounce read-backs (rtx_render_frame) on a
// renderer per resolution class.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <queue>
#include <thread>
#include <unordered_map>

#include "arena.hpp"
#include "rtx_impl.hpp"

uint64_t rtx_scene_device_bytes(const rtx_scene* s);

namespace {

struct Later {
    bool operator()(const rtx_request& a, const rtx_request& b) const { return a.arrival_ns > b.arrival_ns; }
};

// A request between its arrival and its completion.
struct Live {
    rtx_request req{};
    int32_t group = -1;
    int32_t frame = -1;
    int64_t start_ns = 0;
    void* staging = nullptr;  // pinned copy of the accumulators (requests with out_accum)
};

}  // namespace

struct rtx_server {
    rtx_ctx* ctx = nullptr;
    rtx_server_params p{};
    // plain mode
    rtx_renderer* plain_r[RTX_SERVER_MAX_CLASSES] = {};
    // wavefront mode: groups of requests, each a wavefront renderer stepping
    // on a stream of its own (one group's launch tails overlap another's work)
    std::vector<rtx_wavefront*> wf;
    std::vector<hipStream_t> gstream;
    hipStream_t stream = nullptr;
    std::vector<rtx_scene*> scenes;
    uint64_t workers_bytes = 0, scenes_bytes = 0, budget = 0;
    uint64_t step_px_cap = 0;  // the srv_step_pixels tune when the server was created
    std::mutex mu;
    std::condition_variable cv;
    std::priority_queue<rtx_request, std::vector<rtx_request>, Later> pending;  // by arrival
    std::deque<rtx_completion> completions;
    uint64_t submitted = 0, finished = 0;
    bool stop = false;
    std::thread thread;
    void run_plain();
    void run_wavefront();
};

int64_t rtx_now_ns(void) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

static int class_of(const rtx_server* srv, int w, int h) {
    for (int c = 0; c < srv->p.classes; ++c)
        if (srv->p.class_w[c] == w && srv->p.class_h[c] == h) return c;
    return -1;
}

static rtx_completion completion_of(const rtx_request& rq, int32_t status) {
    rtx_completion c{};
    c.tag = rq.tag;
    c.arrival_ns = rq.arrival_ns;
    c.status = status;
    c.worker = -1;
    return c;
}

// The per-request baseline: the reference's frame, one request at a time.
void rtx_server::run_plain() {
    for (;;) {
        rtx_request rq{};
        bool have = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (!pending.empty() && pending.top().arrival_ns <= rtx_now_ns()) {
                    rq = pending.top();
                    pending.pop();
                    have = true;
                    break;
                }
                if (stop && pending.empty()) return;
                int64_t until = rtx_now_ns() + 2000000;
                if (!pending.empty()) until = std::min(until, pending.top().arrival_ns);
                const int64_t wait = until - rtx_now_ns();
                if (wait > 20000)
                    cv.wait_for(lk, std::chrono::nanoseconds(wait - 10000));
                else
                    std::this_thread::yield();
            }
        }
        if (!have) continue;
        rtx_completion c = completion_of(rq, RTX_OK);
        const int cls = class_of(this, rq.width, rq.height);
        rtx_scene* sc = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (rq.scene >= 0 && rq.scene < (int)scenes.size()) sc = scenes[rq.scene];
        }
        c.start_ns = rtx_now_ns();
        if (cls < 0 || !sc || rq.bounces < 0 || rq.bounces > p.max_bounces) {
            c.status = RTX_EINVAL;
        } else {
            rtx_renderer* R = plain_r[cls];
            rtx_frame_stats fs{};
            int32_t e = rtx_renderer_set_bounces(R, rq.bounces);
            if (!e) e = rtx_render_frame(R, sc, &rq.camera, 0, &fs, stream);
            if (!e && rq.out_accum)
                e = hipMemcpy(rq.out_accum, R->fb->accum.p, (size_t)rq.width * rq.height * 12,
                              hipMemcpyDeviceToHost) == hipSuccess ? RTX_OK : RTX_EDEVICE;
            c.rays = fs.primary_rays + fs.secondary_rays;
            c.shadow_rays = fs.shadow_rays;
            c.status = e;
            c.worker = cls;
        }
        c.done_ns = rtx_now_ns();
        std::lock_guard<std::mutex> lk(mu);
        completions.push_back(c);
        ++finished;
        cv.notify_all();
    }
}

void rtx_server::run_wavefront() {
    const int G = (int)wf.size();
    rtxwf::Admission adm;
    std::unordered_map<uint64_t, Live> live;  // by the server's request sequence number
    uint64_t seq = 0;
    size_t inflight = 0;
    std::vector<size_t> g_inflight(G, 0);
    std::vector<int32_t> g_active(G, 0);  // admitted requests with steps still to enqueue, per group
    std::vector<rtx_completion> done;
    const uint64_t arena = rtx_wf_arena_size(wf[0]);
    const int64_t steps_ahead = p.steps_ahead > 0 ? p.steps_ahead : 1;
    std::vector<uint64_t> g_newpx(G, 0);  // primary pixels admitted into each group's next step
    auto finish = [&](const rtx_completion& c) { done.push_back(c); };
    for (;;) {
        const int64_t now = rtx_now_ns();
        bool stopping = false;
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!pending.empty() && pending.top().arrival_ns <= now) {
                const rtx_request rq = pending.top();
                pending.pop();
                const bool ok = rq.scene >= 0 && rq.scene < (int)scenes.size() && rq.bounces >= 0 &&
                                rq.bounces <= p.max_bounces && rq.width > 0 && rq.height > 0;
                if (!ok) {
                    rtx_completion c = completion_of(rq, RTX_EINVAL);
                    c.start_ns = c.done_ns = now;
                    finish(c);
                    continue;
                }
                Live L;
                L.req = rq;
                live.emplace(seq, L);
                adm.push(seq, rtx_wavefront_frame_bytes(rq.width, rq.height, rq.bounces, 0));
                ++seq;
            }
            stopping = stop && pending.empty();
        }
        // admission: first come first served, by arena bytes, into the group
        // with the fewest requests in flight that has room
        std::vector<uint64_t> admitted, rejected;
        auto try_add = [&](const rtxwf::Admission::Item& it) {
            Live& L = live.at(it.id);
            rtx_scene* sc = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                sc = scenes[L.req.scene];
            }
            std::vector<int> order(G);
            for (int g = 0; g < G; ++g) order[g] = g;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return g_inflight[a] < g_inflight[b]; });
            for (int g : order) {
                if (g_inflight[g] >= (size_t)p.workers / G) continue;
                const uint64_t px = (uint64_t)L.req.width * (uint64_t)L.req.height;
                if (step_px_cap && g_newpx[g] > 0 && g_newpx[g] + px > step_px_cap) continue;
                int32_t fr = -1;
                const int32_t e = rtx_wavefront_add(wf[g], sc, &L.req.camera, L.req.width, L.req.height,
                                                    L.req.bounces, 0, it.id, &fr);
                if (e == RTX_ENOMEM) continue;  // this group has no room now
                if (e) {
                    rtx_completion c = completion_of(L.req, e);
                    c.start_ns = c.done_ns = rtx_now_ns();
                    finish(c);
                    live.erase(it.id);
                    return true;  // leaves the queue (not in flight)
                }
                L.group = g;
                L.frame = fr;
                L.start_ns = rtx_now_ns();
                ++g_inflight[g];
                g_newpx[g] += px;
                ++g_active[g];
                if (L.req.out_accum &&
                    hipHostMalloc(&L.staging, (size_t)L.req.width * L.req.height * 12, hipHostMallocDefault) !=
                        hipSuccess)
                    L.staging = nullptr;
                return true;
            }
            return false;  // no group has room now: wait at the head
        };
        adm.admit(arena, inflight, (size_t)p.workers, try_add, admitted, rejected);
        for (uint64_t id : admitted)
            if (live.count(id)) ++inflight;
        for (uint64_t id : rejected) {
            Live& L = live.at(id);
            rtx_completion c = completion_of(L.req, RTX_ENOMEM);
            c.start_ns = c.done_ns = rtx_now_ns();
            finish(c);
            live.erase(id);
        }
        // steps: each group keeps up to STEPS_AHEAD enqueued while its requests need them
        for (int g = 0; g < G; ++g) {
            if (g_active[g] <= 0 || rtx_wf_steps_in_flight(wf[g]) >= steps_ahead) continue;
            auto copy_out = [&](hipStream_t st, const std::vector<int>& fin) -> int32_t {
                for (auto& kv : live) {
                    Live& L = kv.second;
                    if (!L.staging || L.group != g) continue;
                    if (std::find(fin.begin(), fin.end(), L.frame) == fin.end()) continue;
                    RTX_CHECK_HIP(hipMemcpyAsync(L.staging, rtx_wf_accum_dev(wf[g], L.frame),
                                                 (size_t)L.req.width * L.req.height * 12, hipMemcpyDeviceToHost, st));
                }
                return RTX_OK;
            };
            int32_t pend = 0;
            g_newpx[g] = 0;
            const int32_t e = rtx_wavefront_step_impl(wf[g], gstream[g], &pend, nullptr, copy_out);
            if (e) {  // a failed step fails every request of the group
                for (auto it = live.begin(); it != live.end();) {
                    if (it->second.group == g) {
                        rtx_completion c = completion_of(it->second.req, e);
                        c.start_ns = it->second.start_ns;
                        c.done_ns = rtx_now_ns();
                        finish(c);
                        --inflight;
                        it = live.erase(it);
                    } else {
                        ++it;
                    }
                }
                g_inflight[g] = 0;
                pend = 0;
            }
            g_active[g] = pend;
        }
        // completions
        for (int g = 0; g < G; ++g) {
            int32_t fr[64];
            uint64_t tags[64];
            uint32_t nf = 0;
            if (rtx_wavefront_poll(wf[g], fr, tags, 64, &nf) != RTX_OK) continue;
            for (uint32_t i = 0; i < nf; ++i) {
                auto it = live.find(tags[i]);
                if (it == live.end()) continue;
                Live& L = it->second;
                rtx_completion c = completion_of(L.req, RTX_OK);
                c.start_ns = L.start_ns;
                c.done_ns = rtx_now_ns();
                rtx_frame_stats fs{};
                c.status = rtx_wavefront_frame_stats(wf[g], L.frame, &fs);
                c.rays = fs.primary_rays + fs.secondary_rays;
                c.shadow_rays = fs.shadow_rays;
                c.worker = g * RTX_WF_MAX_FRAMES + L.frame;
                if (L.req.out_accum) {
                    if (L.staging)
                        memcpy(L.req.out_accum, L.staging, (size_t)L.req.width * L.req.height * 12);
                    else
                        c.status = RTX_ENOMEM;
                }
                if (L.staging) (void)hipHostFree(L.staging);
                (void)rtx_wavefront_release(wf[g], L.frame);
                --inflight;
                --g_inflight[g];
                finish(c);
                live.erase(it);
            }
        }
        if (!done.empty()) {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& c : done) completions.push_back(c);
            finished += done.size();
            done.clear();
            cv.notify_all();
        }
        if (stopping && live.empty()) break;
        // wait: spin while steps are in flight (completion times matter),
        // otherwise sleep until the next arrival or a submission
        if (!live.empty()) {
            std::this_thread::yield();
        } else {
            std::unique_lock<std::mutex> lk(mu);
            int64_t until = rtx_now_ns() + 2000000;  // 2 ms
            if (!pending.empty()) until = std::min(until, pending.top().arrival_ns);
            const int64_t wait = until - rtx_now_ns();
            if (wait > 20000 && !stop) cv.wait_for(lk, std::chrono::nanoseconds(wait - 10000));
        }
    }
}

int32_t rtx_server_create(rtx_ctx* ctx, const rtx_server_params* p, rtx_server** out) {
    RTX_REQUIRE(ctx && p && out, RTX_EINVAL, "rtx_server_create: null argument");
    RTX_REQUIRE(p->max_bounces >= 0 && p->max_bounces <= 16, RTX_EINVAL, "rtx_server_create: bad max_bounces");
    RTX_REQUIRE(p->classes >= 0 && p->classes <= RTX_SERVER_MAX_CLASSES, RTX_EINVAL,
                "rtx_server_create: bad class count");
    RTX_REQUIRE(!p->plain || p->classes >= 1, RTX_EINVAL, "rtx_server_create: plain mode needs resolution classes");
    RTX_REQUIRE(p->plain || (p->workers >= 1 && p->workers <= RTX_WF_MAX_FRAMES), RTX_EINVAL,
                "rtx_server_create: workers (requests in flight) must be in [1, 256]");
    for (int c = 0; c < p->classes; ++c)
        RTX_REQUIRE(p->class_w[c] > 0 && p->class_h[c] > 0, RTX_EINVAL, "rtx_server_create: bad class size");
    rtx_server* srv = new rtx_server();
    srv->ctx = ctx;
    srv->step_px_cap = (uint64_t)ctx->tune.srv_step_pixels;
    srv->p = *p;
    if (srv->p.plain) srv->p.workers = 1;
    if (srv->p.queue_cap <= 0) srv->p.queue_cap = 65536;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    srv->budget = p->hbm_budget ? p->hbm_budget : (uint64_t)(0.9 * (double)free_b);
    const uint64_t before = ctx->bytes_in_use;
    int32_t e = RTX_OK;
    // (each stream takes one of the process's hardware queues: the plain
    // server renders on one, the wavefront groups on one each)
    if (srv->p.plain && hipStreamCreateWithFlags(&srv->stream, hipStreamNonBlocking) != hipSuccess) {
        srv->stream = nullptr;
        rtx_set_error("rtx_server_create: stream creation failed");
        e = RTX_EDEVICE;
    }
    if (!e && srv->p.plain) {
        for (int c = 0; c < srv->p.classes && !e; ++c) {
            rtx_render_params rp{};
            rp.width = p->class_w[c];
            rp.height = p->class_h[c];
            rp.max_bounces = p->max_bounces;
            e = rtx_renderer_create(ctx, &rp, &srv->plain_r[c]);
        }
    } else if (!e) {
        const int G = p->groups > 0 ? p->groups : 2;
        const uint64_t arena = p->arena_bytes ? p->arena_bytes : srv->budget / 2;
        if (G > 8 || p->workers < G || p->steps_ahead < 0 || p->steps_ahead > 4) {
            rtx_set_error("rtx_server_create: groups must be in [1, 8] and at most workers, steps_ahead in [0, 4]");
            e = RTX_EINVAL;
        } else if (arena >= srv->budget) {
            rtx_set_error("rtx_server_create: the request arena exceeds the HBM budget");
            e = RTX_ENOMEM;
        }
        for (int g = 0; g < G && !e; ++g) {
            rtx_wavefront_params wp{};
            wp.arena_bytes = arena / G;
            wp.max_frames = (p->workers + G - 1) / G;
            rtx_wavefront* w = nullptr;
            hipStream_t gs = nullptr;
            if ((e = rtx_wavefront_create(ctx, &wp, &w))) break;
            // several groups: each persistent traversal launch leaves room on
            // every CU for the other groups' short launches (shading,
            // secondary generation), which otherwise wait for it to drain
            if (G > 1) rtx_wf_set_auto_share(w, 65);
            srv->wf.push_back(w);
            if (hipStreamCreateWithFlags(&gs, hipStreamNonBlocking) != hipSuccess) {
                rtx_set_error("rtx_server_create: stream creation failed");
                e = RTX_EDEVICE;
                break;
            }
            srv->gstream.push_back(gs);
        }
    }
    if (!e && hipDeviceSynchronize() != hipSuccess) e = RTX_EDEVICE;
    srv->workers_bytes = ctx->bytes_in_use - before;
    if (!e && srv->workers_bytes > srv->budget) {
        rtx_set_error("rtx_server_create: the request memory alone exceeds the HBM budget");
        e = RTX_ENOMEM;
    }
    if (e) {
        rtx_server_destroy(srv);
        return e;
    }
    srv->thread = std::thread([srv] {
        if (srv->p.plain)
            srv->run_plain();
        else
            srv->run_wavefront();
    });
    *out = srv;
    return RTX_OK;
}

int32_t rtx_server_destroy(rtx_server* srv) {
    RTX_REQUIRE(srv, RTX_EINVAL, "rtx_server_destroy: null");
    if (srv->thread.joinable()) {
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->stop = true;
        }
        srv->cv.notify_all();
        srv->thread.join();
    }
    if (srv->stream) (void)hipStreamSynchronize(srv->stream);
    for (auto* r : srv->plain_r)
        if (r) rtx_renderer_destroy(r);
    for (auto* w : srv->wf) rtx_wavefront_destroy(w);
    for (auto gs : srv->gstream) {
        (void)hipStreamSynchronize(gs);
        (void)hipStreamDestroy(gs);
    }
    if (srv->stream) (void)hipStreamDestroy(srv->stream);
    delete srv;
    return RTX_OK;
}

int32_t rtx_server_add_scene(rtx_server* srv, rtx_scene* s, int32_t* scene_id) {
    RTX_REQUIRE(srv && s && scene_id, RTX_EINVAL, "rtx_server_add_scene: null argument");
    RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_server_add_scene: scene has no BVH (build it first)");
    const uint64_t b = rtx_scene_device_bytes(s);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->workers_bytes + srv->scenes_bytes + b > srv->budget) {
        rtx_set_error("rtx_server_add_scene: scene of " + std::to_string(b) + " bytes exceeds the HBM budget");
        return RTX_ENOMEM;
    }
    srv->scenes.push_back(s);
    srv->scenes_bytes += b;
    *scene_id = (int32_t)srv->scenes.size() - 1;
    return RTX_OK;
}

int32_t rtx_server_memory(rtx_server* srv, uint64_t* workers_bytes, uint64_t* scenes_bytes) {
    RTX_REQUIRE(srv, RTX_EINVAL, "rtx_server_memory: null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (workers_bytes) *workers_bytes = srv->workers_bytes;
    if (scenes_bytes) *scenes_bytes = srv->scenes_bytes;
    return RTX_OK;
}

int32_t rtx_server_submit(rtx_server* srv, const rtx_request* reqs, uint32_t n) {
    RTX_REQUIRE(srv && (reqs || n == 0), RTX_EINVAL, "rtx_server_submit: null argument");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        RTX_REQUIRE(srv->submitted - srv->finished + n <= (uint64_t)srv->p.queue_cap, RTX_ENOMEM,
                    "rtx_server_submit: queue full");
        for (uint32_t i = 0; i < n; ++i) srv->pending.push(reqs[i]);
        srv->submitted += n;
    }
    srv->cv.notify_all();
    return RTX_OK;
}

int32_t rtx_server_poll(rtx_server* srv, rtx_completion* out, uint32_t cap, uint32_t* n) {
    RTX_REQUIRE(srv && n && (out || cap == 0), RTX_EINVAL, "rtx_server_poll: null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    uint32_t k = 0;
    while (k < cap && !srv->completions.empty()) {
        out[k++] = srv->completions.front();
        srv->completions.pop_front();
    }
    *n = k;
    return RTX_OK;
}

int32_t rtx_server_drain(rtx_server* srv, int64_t timeout_ms) {
    RTX_REQUIRE(srv, RTX_EINVAL, "rtx_server_drain: null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    const bool ok = srv->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms),
                                     [srv] { return srv->finished >= srv->submitted; });
    RTX_REQUIRE(ok, RTX_ESTATE, "rtx_server_drain: timed out");
    return RTX_OK;
}
