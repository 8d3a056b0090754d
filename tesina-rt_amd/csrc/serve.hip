// Request server over resident scenes: open-loop arrivals, concurrent
// workers (one stream and one preallocated renderer per resolution class
// each), HBM-budget admission of scenes, device-side completion times.
//
// A dispatcher thread owns every GPU call of the server: it moves requests
// whose arrival time has come into the ready queue, starts ready requests on
// free workers (a worker freed by a completion takes the next request at
// once, so requests of different scenes, sizes and depths run side by side),
// and polls the workers' completion events.  A request's latency is
// done_ns - arrival_ns, where done_ns is when the dispatcher observed its
// frame (accumulators and ray counts included) complete on the device.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <queue>
#include <thread>

#include "rtx_impl.hpp"

uint64_t rtx_scene_device_bytes(const rtx_scene* s);

namespace {

struct Worker {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    rtx_renderer* r[RTX_SERVER_MAX_CLASSES] = {};
    unsigned long long* counts_host = nullptr;  // pinned: [0] secondary, [1] shadow
    bool busy = false;
    rtx_request req{};
    int cls = 0;
    int64_t start_ns = 0;
};

struct Later {
    bool operator()(const rtx_request& a, const rtx_request& b) const { return a.arrival_ns > b.arrival_ns; }
};

}  // namespace

struct rtx_server {
    rtx_ctx* ctx = nullptr;
    rtx_server_params p{};
    std::vector<Worker> workers;
    std::vector<rtx_scene*> scenes;
    uint64_t workers_bytes = 0, scenes_bytes = 0, budget = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::priority_queue<rtx_request, std::vector<rtx_request>, Later> pending;  // by arrival
    std::deque<rtx_completion> completions;
    uint64_t submitted = 0, finished = 0;
    bool stop = false;
    std::thread thread;
    void run();
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

void rtx_server::run() {
    std::deque<rtx_request> ready;
    std::vector<rtx_completion> done_batch;
    for (;;) {
        const int64_t now = rtx_now_ns();
        bool stopping = false;
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!pending.empty() && pending.top().arrival_ns <= now) {
                ready.push_back(pending.top());
                pending.pop();
            }
            stopping = stop && pending.empty();
        }
        // completions
        int inflight = 0;
        for (size_t wi = 0; wi < workers.size(); ++wi) {
            Worker& w = workers[wi];
            if (!w.busy) continue;
            const hipError_t q = hipEventQuery(w.done);
            if (q == hipErrorNotReady) {
                ++inflight;
                continue;
            }
            rtx_completion c{};
            c.tag = w.req.tag;
            c.arrival_ns = w.req.arrival_ns;
            c.start_ns = w.start_ns;
            c.done_ns = rtx_now_ns();
            c.rays = (uint64_t)w.req.width * w.req.height + w.counts_host[0];
            c.shadow_rays = w.counts_host[1];
            c.status = q == hipSuccess ? RTX_OK : RTX_EDEVICE;
            c.worker = (int32_t)wi;
            done_batch.push_back(c);
            w.busy = false;
        }
        // start ready requests on free workers
        for (size_t wi = 0; wi < workers.size() && !ready.empty(); ++wi) {
            Worker& w = workers[wi];
            if (w.busy) continue;
            const rtx_request rq = ready.front();
            ready.pop_front();
            rtx_completion c{};
            c.tag = rq.tag;
            c.arrival_ns = rq.arrival_ns;
            c.worker = (int32_t)wi;
            const int cls = class_of(this, rq.width, rq.height);
            int nscenes = 0;
            {
                std::lock_guard<std::mutex> lk(mu);
                nscenes = (int)scenes.size();
            }
            if (cls < 0 || rq.scene < 0 || rq.scene >= nscenes || rq.bounces < 0 || rq.bounces > p.max_bounces) {
                c.status = RTX_EINVAL;
                c.start_ns = c.done_ns = rtx_now_ns();
                done_batch.push_back(c);
                continue;
            }
            rtx_renderer* R = w.r[cls];
            rtx_scene* sc = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                sc = scenes[rq.scene];
            }
            int32_t e = rtx_renderer_set_bounces(R, rq.bounces);
            w.start_ns = rtx_now_ns();
            if (p.plain) {
                // the per-request baseline: the reference's frame with its
                // per-bounce read-backs, one request at a time
                rtx_frame_stats fs{};
                if (!e) e = rtx_render_frame(R, sc, &rq.camera, 0, &fs, w.stream);
                if (!e && rq.out_accum)
                    e = hipMemcpy(rq.out_accum, R->fb->accum.p, (size_t)rq.width * rq.height * 12,
                                  hipMemcpyDeviceToHost) == hipSuccess ? RTX_OK : RTX_EDEVICE;
                c.start_ns = w.start_ns;
                c.done_ns = rtx_now_ns();
                c.rays = fs.primary_rays + fs.secondary_rays;
                c.shadow_rays = fs.shadow_rays;
                c.status = e;
                done_batch.push_back(c);
                continue;
            }
            if (!e) e = rtx_render_frame_async(R, sc, &rq.camera, 0, w.stream);
            if (!e && hipMemcpyAsync(w.counts_host, R->counts.p, 16, hipMemcpyDeviceToHost, w.stream) != hipSuccess)
                e = RTX_EDEVICE;
            if (!e && rq.out_accum &&
                hipMemcpyAsync(rq.out_accum, R->fb->accum.p, (size_t)rq.width * rq.height * 12,
                               hipMemcpyDeviceToHost, w.stream) != hipSuccess)
                e = RTX_EDEVICE;
            if (!e && hipEventRecord(w.done, w.stream) != hipSuccess) e = RTX_EDEVICE;
            if (e) {
                c.status = e;
                c.start_ns = w.start_ns;
                c.done_ns = rtx_now_ns();
                done_batch.push_back(c);
                continue;
            }
            w.req = rq;
            w.cls = cls;
            w.busy = true;
            ++inflight;
        }
        if (!done_batch.empty()) {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& c : done_batch) completions.push_back(c);
            finished += done_batch.size();
            done_batch.clear();
            cv.notify_all();
        }
        if (stopping && ready.empty() && inflight == 0) break;
        // wait: spin while frames are in flight (completion times matter),
        // otherwise sleep until the next arrival or a submission
        if (inflight > 0 && (ready.empty() || inflight == (int)workers.size())) {
            std::this_thread::yield();
        } else if (ready.empty()) {
            std::unique_lock<std::mutex> lk(mu);
            int64_t until = rtx_now_ns() + 2000000;  // 2 ms
            if (!pending.empty()) until = std::min(until, pending.top().arrival_ns);
            const int64_t wait = until - rtx_now_ns();
            if (wait > 20000 && !stop)
                cv.wait_for(lk, std::chrono::nanoseconds(wait - 10000));
        }
    }
}

int32_t rtx_server_create(rtx_ctx* ctx, const rtx_server_params* p, rtx_server** out) {
    RTX_REQUIRE(ctx && p && out, RTX_EINVAL, "rtx_server_create: null argument");
    RTX_REQUIRE(p->workers >= 1 && p->workers <= 16, RTX_EINVAL, "rtx_server_create: workers must be in [1, 16]");
    RTX_REQUIRE(p->max_bounces >= 0 && p->max_bounces <= 16, RTX_EINVAL, "rtx_server_create: bad max_bounces");
    RTX_REQUIRE(p->classes >= 1 && p->classes <= RTX_SERVER_MAX_CLASSES, RTX_EINVAL,
                "rtx_server_create: bad class count");
    for (int c = 0; c < p->classes; ++c)
        RTX_REQUIRE(p->class_w[c] > 0 && p->class_h[c] > 0, RTX_EINVAL, "rtx_server_create: bad class size");
    rtx_server* srv = new rtx_server();
    srv->ctx = ctx;
    srv->p = *p;
    if (srv->p.plain) srv->p.workers = 1;
    if (srv->p.queue_cap <= 0) srv->p.queue_cap = 65536;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    srv->budget = p->hbm_budget ? p->hbm_budget : (uint64_t)(0.9 * (double)free_b);
    const uint64_t before = ctx->bytes_in_use;
    int32_t e = 0;
    srv->workers.resize(srv->p.workers);
    for (auto& w : srv->workers) {
        if (hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&w.counts_host, 16, hipHostMallocDefault) != hipSuccess) {
            rtx_set_error("rtx_server_create: stream, event or pinned buffer creation failed");
            e = RTX_EDEVICE;
            break;
        }
        for (int c = 0; c < srv->p.classes && !e; ++c) {
            rtx_render_params rp{};
            rp.width = p->class_w[c];
            rp.height = p->class_h[c];
            rp.max_bounces = p->max_bounces;
            if ((e = rtx_renderer_create(ctx, &rp, &w.r[c]))) break;
            w.r[c]->single_stream = true;  // one stream per worker; workers give the concurrency
            // first raygen uploads the pixel-sample table; do it now, not in a request
            rtx_camera cam{};
            e = rtx_raygen_primary(w.r[c]->rays[0], &cam, rp.width, rp.height, 32, 1, nullptr, 0, 0, w.stream);
        }
        if (e) break;
    }
    if (!e && hipDeviceSynchronize() != hipSuccess) e = RTX_EDEVICE;
    srv->workers_bytes = ctx->bytes_in_use - before;
    if (!e && srv->workers_bytes > srv->budget) {
        rtx_set_error("rtx_server_create: workers alone exceed the HBM budget");
        e = RTX_ENOMEM;
    }
    if (e) {
        rtx_server_destroy(srv);
        return e;
    }
    srv->thread = std::thread([srv] { srv->run(); });
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
    for (auto& w : srv->workers) {
        if (w.stream) (void)hipStreamSynchronize(w.stream);
        for (auto* r : w.r)
            if (r) rtx_renderer_destroy(r);
        if (w.counts_host) (void)hipHostFree(w.counts_host);
        if (w.done) (void)hipEventDestroy(w.done);
        if (w.stream) (void)hipStreamDestroy(w.stream);
    }
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
