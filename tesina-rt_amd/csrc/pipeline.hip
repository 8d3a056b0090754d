// Frame pipeline of one scene with several frames in flight (see rtx.h).
#include "rtx_impl.hpp"

struct rtx_pipeline {
    rtx_ctx* ctx = nullptr;
    rtx_scene* s = nullptr;
    int depth = 2;
    rtx_lbvh_params bp{30, 1};
    std::vector<rtx_renderer*> r;     // slot renderers
    std::vector<hipStream_t> st;      // slot streams
    std::vector<hipEvent_t> sorted;   // slot's frame: sort phase done (its tree phase may start)
    std::vector<hipEvent_t> copied;   // slot's shading copy holds its frame's normals and material ids
    std::vector<hipEvent_t> done;     // slot's frame finished (its buffer set and copy are free again)
    std::vector<int> reads;           // buffer set each slot's last frame renders from (-1: none yet)
    hipStream_t bst = nullptr;        // build stream
    bool animate = false;
    float t0 = 0.f, dt = 0.f;
    uint64_t next = 0;
};

int32_t rtx_pipeline_create(rtx_ctx* ctx, rtx_scene* s, const rtx_render_params* p, int32_t depth,
                            rtx_pipeline** out) {
    RTX_REQUIRE(ctx && s && p && out, RTX_EINVAL, "rtx_pipeline_create: null argument");
    RTX_REQUIRE(depth >= 1 && depth <= rtx_scene::COPIES, RTX_EINVAL, "rtx_pipeline_create: depth must be in [1, 8]");
    rtx_pipeline* pl = new rtx_pipeline();
    pl->ctx = ctx;
    pl->s = s;
    pl->depth = depth;
    pl->bp = rtx_lbvh_params{p->depth > 0 ? p->depth : 30, p->min_leaf > 0 ? p->min_leaf : 1};
    int32_t e = RTX_OK;
    if (hipStreamCreateWithFlags(&pl->bst, hipStreamNonBlocking) != hipSuccess) {
        pl->bst = nullptr;
        e = RTX_EDEVICE;
    }
    for (int k = 0; k < depth && !e; ++k) {
        rtx_renderer* r = nullptr;
        if ((e = rtx_renderer_create(ctx, p, &r))) break;
        // slots overlap whole frames with each other; one stream per slot keeps
        // the stream count (and hardware-queue sharing) low
        r->single_stream = true;
        pl->r.push_back(r);
        hipStream_t q = nullptr;
        hipEvent_t a = nullptr, b = nullptr, c = nullptr;
        if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&a, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c, hipEventDisableTiming) != hipSuccess) {
            e = RTX_EDEVICE;
            if (q) (void)hipStreamDestroy(q);
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            break;
        }
        pl->st.push_back(q);
        pl->copied.push_back(a);
        pl->done.push_back(b);
        pl->sorted.push_back(c);
        pl->reads.push_back(-1);
    }
    if (e) {
        if (e == RTX_EDEVICE) rtx_set_error("rtx_pipeline_create: stream or event creation failed");
        rtx_pipeline_destroy(pl);
        return e;
    }
    *out = pl;
    return RTX_OK;
}

int32_t rtx_pipeline_destroy(rtx_pipeline* pl) {
    RTX_REQUIRE(pl, RTX_EINVAL, "rtx_pipeline_destroy: null");
    if (pl->bst) (void)hipStreamSynchronize(pl->bst);
    for (auto q : pl->st) (void)hipStreamSynchronize(q);
    for (auto* r : pl->r) rtx_renderer_destroy(r);
    for (auto ev : pl->copied) (void)hipEventDestroy(ev);
    for (auto ev : pl->sorted) (void)hipEventDestroy(ev);
    for (auto ev : pl->done) (void)hipEventDestroy(ev);
    for (auto q : pl->st) (void)hipStreamDestroy(q);
    if (pl->bst) (void)hipStreamDestroy(pl->bst);
    delete pl;
    return RTX_OK;
}

int32_t rtx_pipeline_set_animation(rtx_pipeline* pl, int32_t enable, float t0, float dt) {
    RTX_REQUIRE(pl, RTX_EINVAL, "rtx_pipeline_set_animation: null");
    pl->animate = enable != 0;
    pl->t0 = t0;
    pl->dt = dt;
    return RTX_OK;
}

// Frame n uses slot k = n mod depth: buffer set k, shading copy k and stream
// k.  The build stream runs, in frame order, each frame's animation and the
// sort phase of its build (which permutes the scene's index buffers in place
// and so chains the builds), then copies the normals and material ids the
// frame shades with.  Slot k's stream runs the tree phase of the build into
// set k and renders the frame from it.  Set k and copy k are rewritten only
// after frame n - depth, their last user, is done.
int32_t rtx_pipeline_frame(rtx_pipeline* pl, const rtx_camera* cam, int32_t rebuild, uint64_t* frame) {
    RTX_REQUIRE(pl && cam, RTX_EINVAL, "rtx_pipeline_frame: null argument");
    rtx_scene* s = pl->s;
    const uint64_t n = pl->next;
    const int k = (int)(n % (uint64_t)pl->depth);
    const bool build = rebuild || !s->built;
    int32_t e = RTX_OK;
    if (n == 0) {  // every slot's buffers, before frames run side by side (no allocation after this)
        const int keep = s->cur;
        for (int j = 0; j < pl->depth && !e; ++j) {
            s->cur = j;
            e = rtx_lbvh_reserve(s, pl->bp.depth);
            if (!e) e = rtx_reserve_render_copy(s, j, true);
        }
        s->cur = keep;
        if (e) return e;
    }
    if (pl->animate && (e = rtx_scene_animate(s, pl->t0, pl->dt, pl->bst))) return e;
    if (n >= (uint64_t)pl->depth) RTX_CHECK_HIP(hipStreamWaitEvent(pl->bst, pl->done[k], 0));
    int set = s->cur;
    if (build) {
        // frames without a rebuild read the current set from their own slots
        for (int j = 0; j < pl->depth; ++j)
            if (j != k && pl->reads[j] == k) RTX_CHECK_HIP(hipStreamWaitEvent(pl->bst, pl->done[j], 0));
        s->cur = k;
        if ((e = rtx_lbvh_build_split(s, &pl->bp, pl->bst, pl->sorted[k], pl->st[k]))) return e;
        set = k;
    }
    if ((e = rtx_refresh_render_copy(s, k, pl->bst, true))) return e;
    RTX_CHECK_HIP(hipEventRecord(pl->copied[k], pl->bst));
    RTX_CHECK_HIP(hipStreamWaitEvent(pl->st[k], pl->copied[k], 0));
    s->use_set = set;
    s->use_copy = k;
    e = rtx_render_frame_async(pl->r[k], s, cam, 0, pl->st[k]);
    s->use_set = -1;
    s->use_copy = -1;
    if (e) return e;
    RTX_CHECK_HIP(hipEventRecord(pl->done[k], pl->st[k]));
    pl->reads[k] = set;
    pl->next = n + 1;
    if (frame) *frame = n;
    return RTX_OK;
}

static int32_t slot_of(rtx_pipeline* pl, uint64_t frame, int* k) {
    RTX_REQUIRE(frame < pl->next && frame + (uint64_t)pl->depth >= pl->next, RTX_ESTATE,
                "rtx_pipeline: frame not in flight or already overwritten");
    *k = (int)(frame % (uint64_t)pl->depth);
    return RTX_OK;
}

int32_t rtx_pipeline_wait(rtx_pipeline* pl, uint64_t frame) {
    RTX_REQUIRE(pl, RTX_EINVAL, "rtx_pipeline_wait: null");
    int k = 0;
    int32_t e = slot_of(pl, frame, &k);
    if (e) return e;
    RTX_CHECK_HIP(hipEventSynchronize(pl->done[k]));
    return RTX_OK;
}

rtx_fb* rtx_pipeline_fb(rtx_pipeline* pl, uint64_t frame) {
    int k = 0;
    if (!pl || slot_of(pl, frame, &k)) return nullptr;
    return pl->r[k]->fb;
}

int32_t rtx_pipeline_stats(rtx_pipeline* pl, uint64_t frame, rtx_frame_stats* stats) {
    RTX_REQUIRE(pl && stats, RTX_EINVAL, "rtx_pipeline_stats: null argument");
    int k = 0;
    int32_t e = slot_of(pl, frame, &k);
    if (e) return e;
    RTX_CHECK_HIP(hipEventSynchronize(pl->done[k]));
    return rtx_renderer_last_stats(pl->r[k], stats);
}
