// Internal declarations of the rtx library (gfx950).
//
// Device data layout (HBM):
//   scene : pos4[V] (x,y,z,0)   nrm4[V] (normalized vertex normal, 0)   uv2[V]
//           vnrm4[V] (vertex normal as given, read back for parity)
//           idx[3T] (permuted in place by each build)   matmap[T]   orig[T]
//   bvh   : node[N]  32 B = {lo.xyz, a | hi.xyz, b}; internal: a = left child
//                    (right = a+1, children are adjacent), b = 0;
//                    leaf: a = first triangle, b = triangle count (>= 1)
//           delta[N] 16 B = right.lo - left.lo (near-child test of
//                    src/kernel/trace-bvh.cl:392-409, precomputed at refit)
//           parent[N], tri[3T] 48 B per sorted triangle {v0,i0 | v1,i1 | v2,i2}
//   rays  : ori4 (origin, tMin)  dir4 (dir, tMax)  pixel  contrib   (SoA)
//   hits  : flags t id bary2 n4 p4 uv2                               (SoA)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <functional>
#include <vector>

#include "../../include/rtx.h"

#define RTX_WAVE 64

// ------------------------------------------------------------ error path --
void rtx_set_error(const std::string& msg);
#define RTX_CHECK_HIP(expr)                                                              \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            rtx_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return RTX_EDEVICE;                                                          \
        }                                                                                \
    } while (0)
#define RTX_REQUIRE(cond, code, msg)   \
    do {                               \
        if (!(cond)) {                 \
            rtx_set_error(msg);        \
            return (code);             \
        }                              \
    } while (0)

// ---------------------------------------------------------- device buffer --
struct rtx_ctx;
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    rtx_ctx* ctx = nullptr;
    // Grows to at least `n` bytes; contents are not kept.
    int32_t reserve(rtx_ctx* c, size_t n);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// --------------------------------------------------------------- objects --
struct KernelTiming {
    std::string name;
    hipEvent_t start, stop;
};

// Tuning knobs (set through rtx_tune_set; picked by the autotuner).
struct rtx_tuning {
    int64_t trace_variant = 1;  // 0 single node, 1 + early delta load, 2 sibling pairs, 3 persistent + ray queue,
                                // 4 stackless walk by parent pointers
    int64_t trace_block = 128;  // 64, 128 or 256 threads per traversal block
    int64_t xcds = 1;           // ray chunks = L2 domains (MI355X: 8 XCDs); 1 = plain grid order
    int64_t sort_bits = 8;      // LBVH radix sort digit width: 8 or 4 bits a pass
    int64_t sort_items = 16;    // LBVH radix sort: pairs per thread of a tile (16 or 32)
    int64_t build_grid = 8;     // LBVH hierarchy and refit levels: persistent blocks per CU (1..16)
    int64_t stack_lds = 32;     // traversal stack entries per lane kept in LDS: 8, 16 or 32 (all)
    int64_t shadow_order = 0;   // 1: shadow rays visit children in the reference's near-child order everywhere;
                                // 0: in index order (no delta fetch) wherever the order cannot change the result
    int64_t coop_leaf = 8;      // leaves of >= this many triangles are tested by the whole wave (0 = off)
    int64_t refill = 16;        // variant 3: idle lanes of a wave that trigger a queue refill
    int64_t side_stream = 1;    // renderers trace shadows and shade each bounce on a second stream
    int64_t inst_lane_roots = 2;  // instances, single-node forms: 1 each lane walks its own roots in one loop,
                                  // 2 the same over the roots whose world box its ray meets, 0 wave-uniform
    int64_t wf_grain = 1;         // wavefront traversal: 64-ray chunks a wave takes per queue atomic
    int64_t srv_step_pixels = 2500000;  // server: new primary pixels a group's step takes at most (0: no
                                        // cap; a step always takes at least one request)
    int64_t wf_share = 0;         // wavefront persistent traversal: percent of the blocks that fit on a CU
                                  // (1..100); 0 = automatic: 100 for a lone renderer, 65 for each of a
                                  // server's groups when there are several
    int64_t wf_sched = 1;         // wavefront traversal: 0 one chunk per wave in launch order, 1 persistent waves
};

struct rtx_ctx {
    int device = 0;
    rtx_tuning tune;
    hipStream_t stream = nullptr;
    int cus = 256;
    uint64_t bytes_in_use = 0;
    bool profiling = false;
    void* build_scratch = nullptr;  // this device's copy of the allocation-free build's scratch array
    std::vector<KernelTiming> pending;
    std::vector<hipEvent_t> event_pool;
    struct Acc {
        std::string name;
        uint64_t count;
        double ms;
    };
    std::vector<Acc> acc;
    hipStream_t pick(void* s) const { return s ? static_cast<hipStream_t>(s) : stream; }
};

// Brackets one launch with events when profiling is on.
struct ProfScope {
    rtx_ctx* ctx;
    hipStream_t st;
    int slot = -1;
    ProfScope(rtx_ctx* c, hipStream_t s, const char* name);
    ~ProfScope();
};

// One set of LBVH buffers: the build's outputs (nodes, near-child deltas,
// triangle records, sorted codes) and its scratch.
struct BuildSet {
    DevBuf node, delta, parent, tri, codes, perm;
    DevBuf gbox;        // 6 x int32 ordered-float keys
    DevBuf levels;      // uint32: [0..64) level offsets, [64..128) level counts, [128] node count
    // build scratch
    DevBuf bstart, bend;          // per-node build state: triangle range
    DevBuf keys_alt, vals_alt;    // radix ping-pong
    DevBuf centroid, partials;    // per-triangle centroids, per-block AABB partials
    DevBuf sort_hist;             // [4][256] digit histograms -> exclusive offsets
    DevBuf sort_status;           // look-back words, per pass x tile x digit
    DevBuf tickets;               // tile tickets for every ordered kernel
    DevBuf hier_status;           // look-back words, per level x tile
    DevBuf fat;                   // node ids of the leaves bounded by one wave each (kept for refit)
    // Node i lives at byte 32 + 32 i of `node`: left children have odd ids, so
    // every sibling pair (i, i + 1) fills one 64-byte aligned chunk.
    float4* nodes() const { return reinterpret_cast<float4*>(static_cast<char*>(node.p) + 32); }
    template <class F>
    void each(F f) {
        DevBuf* all[] = {&node,     &delta,    &parent,   &tri,       &codes,       &perm,    &gbox,
                         &levels,   &bstart,   &bend,     &keys_alt,  &vals_alt,    &centroid, &partials,
                         &sort_hist, &sort_status, &tickets, &hier_status, &fat};
        for (DevBuf* d : all) f(*d);
    }
};

struct rtx_scene {
    rtx_ctx* ctx = nullptr;
    uint32_t V = 0, T = 0, M = 0;
    DevBuf pos, nrm, uv;
    DevBuf vnrm;        // vertex normals as given (nrm holds their normalization)
    DevBuf idx, idx_alt, matmap, matmap_alt, orig, orig_alt;
    DevBuf mats, lights, waves, stage;
    DevBuf clock;       // uint32 frame number of the device-side animation
    // Images of the shading stage: RGBA8 texels and a TexEnt per image; the
    // material textures, then (when the cubemap is on) its six faces.
    DevBuf tex, tex_tab;
    int ntex = 0;
    bool cube = false;
    std::vector<uint8_t> host_tex, host_cube;  // the images as given (the atlas is rebuilt from both)
    std::vector<int32_t> host_tex_dims, host_cube_dims;
    int nwaves = 0;     // wave table entries set by rtx_scene_set_waves
    rtx_lights host_lights{};
    // BVH
    bool built = false;
    bool sah = false;   // the tree came from the CPU SAH builder (no level layout: no refit)
    int depth = 30, min_leaf = 1;
    // Render copies of what the render stages read (pipelined frames: the
    // next frame's build writes the build buffers while frames in flight
    // render from copies).  use_copy selects one for the launches being
    // enqueued (-1: the build's own buffers).
    struct RenderCopy {
        DevBuf node, delta, tri, nrm, matmap;
    };
    static constexpr int COPIES = 8;
    RenderCopy rc[COPIES];
    int use_copy = -1;
    // Frames in flight: the launches being enqueued read the nodes, deltas
    // and triangle records of set use_set, and the normals and material ids
    // of copy use_copy (-1: none).
    int use_set = -1;
    // LBVH buffer sets: builds write, and frames read, set `cur`.  Frames in
    // flight give every slot a set of its own, so that the tree stages of one
    // frame's build run while the next frame sorts (only the permutation of
    // the index buffer chains builds).
    BuildSet sets[COPIES];
    int cur = 0;
    BuildSet& b() { return sets[cur]; }
    const BuildSet& b() const { return sets[cur]; }
    float4* nodes() const { return sets[cur].nodes(); }
    // An accelerator copy the launches being enqueued read instead of all of
    // the above (rtx_trace_with); null: none.
    const struct rtx_bvh* ext = nullptr;
};

// A copy of a scene's accelerator and the per-vertex data a trace reads
// (rtx_bvh_copy), with the scene state it was taken from.
struct rtx_bvh {
    rtx_ctx* ctx = nullptr;
    rtx_scene::RenderCopy copy;
    const rtx_scene* of = nullptr;
    uint32_t V = 0, T = 0;
    int depth = 30;
    bool valid = false;
};

// One image of the shading stage's atlas: first texel, width, height.
struct TexEnt {
    uint64_t off;
    int32_t w, h;
};

// Buffers the render stages (trace, shadow, secondary generation, shading)
// read from a scene: the build's own, or the render copy.
struct SceneView {
    const float4* node;
    const float4* delta;
    const float4* tri;
    const float4* nrm;
    const int32_t* matmap;
    const uint32_t* parent;  // node parents (the stackless walk); null for a render copy
};
inline SceneView scene_view(const rtx_scene* s) {
    if (s->ext) {
        const rtx_scene::RenderCopy& c = s->ext->copy;
        return SceneView{reinterpret_cast<const float4*>(static_cast<const char*>(c.node.p) + 32),
                         c.delta.as<float4>(), c.tri.as<float4>(), c.nrm.as<float4>(), c.matmap.as<int32_t>(),
                         nullptr};
    }
    if (s->use_set >= 0) {
        const BuildSet& b = s->sets[s->use_set];
        const bool c = s->use_copy >= 0;
        return SceneView{b.nodes(), b.delta.as<float4>(), b.tri.as<float4>(),
                         c ? s->rc[s->use_copy].nrm.as<float4>() : s->nrm.as<float4>(),
                         c ? s->rc[s->use_copy].matmap.as<int32_t>() : s->matmap.as<int32_t>(),
                         b.parent.as<uint32_t>()};
    }
    if (s->use_copy >= 0) {
        const rtx_scene::RenderCopy& c = s->rc[s->use_copy];
        return SceneView{reinterpret_cast<const float4*>(static_cast<const char*>(c.node.p) + 32),
                         c.delta.as<float4>(), c.tri.as<float4>(), c.nrm.as<float4>(), c.matmap.as<int32_t>(),
                         nullptr};
    }
    return SceneView{s->nodes(), s->b().delta.as<float4>(), s->b().tri.as<float4>(), s->nrm.as<float4>(),
                     s->matmap.as<int32_t>(), s->b().parent.as<uint32_t>()};
}

struct rtx_rays {
    rtx_ctx* ctx = nullptr;
    uint64_t cap = 0;
    DevBuf ori4, dir4, pixel, contrib;
    DevBuf count;  // uint32[4]: [0] ray count, [1] scan total, [2] first dropped offset, [3] ticket
    DevBuf scan_status;  // look-back words of the generation that writes this bundle
    DevBuf psamp;        // pixel-sample table of primary generation (spp x 3 floats)
    std::vector<float> psamp_host;  // last table uploaded (uploads happen only on change)
};

struct rtx_hits {
    rtx_ctx* ctx = nullptr;
    uint64_t cap = 0;
    DevBuf flags, t, id, bary, n4, p4, uv2;
    DevBuf list, list_count;     // indices of rays with a hit (filled by the trace kernel)
    bool list_valid = false;     // list matches the current hit records...
    uint64_t list_for = 0;       // ...of this many rays (or RTX_DEVICE_COUNT)
};

// One root of a multi-root traversal: a built scene's buffers as the
// traversal reads them, where its triangles start in the shared numbering,
// and the object-to-world transform with its inverse (row-major 4x4, the
// BVHRoot of include/rt/multi-bvh.hpp:9-14).  176 bytes.
struct InstRoot {
    const float4* node;
    const float4* delta;
    const float4* tri;
    const float4* nrm;
    const float2* uv;
    uint32_t base;
    uint32_t ntri;
    float tr[16];
    float tr_inv[16];
};

// World-space box of a root for the lane-roots kernels' culling: the root
// node's box under tr, inflated at use by k * (2 |o| + s) (k < 0: never
// culled).
struct RootBox {
    float lo[3];
    float k;
    float hi[3];
    float s;
};

// Objects of a scene (Scene::create_bvhs / update_bvh_roots,
// src/rt/scene.cpp:596-700): instances of built scenes under transforms.
// `table` is the host image of the device root table; `shipped` is what the
// device holds (rewritten chunk by chunk, in stream order, when they differ).
struct rtx_instances {
    bool deep = false;  // some instance's tree has 32 levels (set by each root refresh)
    rtx_ctx* ctx = nullptr;
    int n = 0;
    std::vector<rtx_scene*> scene;
    std::vector<float> tr, tr_inv;  // 16 floats per instance
    std::vector<InstRoot> table, shipped;
    DevBuf roots;
    DevBuf hit_root;  // uint32 per ray: the root of its closest hit (an order hint for shadow rays)
    DevBuf wbox;      // RootBox per root (culling of the lane-roots kernels)
    rtx_lights lights{};
    // Shading of frames over the instances: a scene holding only the
    // materials (one per distinct scene: its first instance's), the material
    // id of every triangle in the shared numbering, and the lights.
    std::vector<rtx_material> inst_mats;
    rtx_scene* shading = nullptr;
    std::vector<uint32_t> shading_bases;  // the numbering the shading scene was built for
};

struct rtx_fb {
    rtx_ctx* ctx = nullptr;
    int w = 0, h = 0;
    DevBuf accum, rgba;
};

struct rtx_renderer {
    rtx_ctx* ctx = nullptr;
    rtx_render_params p{};
    std::vector<rtx_rays*> rays;  // bounce b's rays: [0] primary, capacity (pixels << b) + 2 (or the user cap)
    std::vector<rtx_hits*> hits;  // bounce b's hit records
    rtx_fb* fb = nullptr;
    rtx_frame_stats last{};
    DevBuf counts;  // uint64 device ray counters of the current frame (see render.hip)
    bool last_async = false;
    // Animation applied at the start of every frame (rtx_renderer_set_animation).
    bool animate = false;
    float anim_t0 = 0.f, anim_dt = 0.f;
    // Side stream of the read-back-free frame: each bounce's shadow rays and
    // shading run there while the main stream generates and traces the next
    // bounce.  ev[b] marks "bounce b traced" on the main stream, ev_join the
    // end of the side stream's work.  Default priority, non-blocking.
    hipStream_t side = nullptr;
    bool single_stream = false;  // render every stage on the caller's stream (frame pipeline slots)
    std::vector<hipEvent_t> ev;
    hipEvent_t ev_join = nullptr;
    // Pipelined frames: the next frame's animation and build run on bstream
    // while this frame renders from the scene's render copy.
    bool pipelined = false;
    bool copy_valid = false;
    const rtx_scene* copy_of = nullptr;
    hipStream_t bstream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_built = nullptr;
    hipGraph_t graph = nullptr;          // captured read-back-free frame
    hipGraphExec_t graph_exec = nullptr;
};

// ------------------------------------------------ ordered (look-back) scan --
// Status word: bits 31..30 flag (0 empty, 1 aggregate, 2 inclusive prefix),
// bits 29..0 value.
#define RTX_LB_AGG 0x40000000u
#define RTX_LB_INC 0x80000000u
#define RTX_LB_VAL 0x3fffffffu

// Fill kernel (used instead of hipMemsetAsync so that every entry point can
// be captured into a hipGraph).
int32_t rtx_fill_u32(rtx_ctx* ctx, void* p, uint32_t value, uint64_t n, hipStream_t st);
// Internal variants used by the renderer: device-count stats accumulation.
int32_t rtx_trace_shadow_impl(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, uint64_t* dev_shadow_count,
                              hipStream_t st);
// Shadow rays of several (rays, hits) bundles in one launch.
int32_t rtx_trace_shadow_multi(rtx_scene* s, rtx_rays* const* rays, rtx_hits* const* hits, const uint64_t* counts,
                               int nseg, uint64_t* dev_shadow_count, hipStream_t st);
// Makes render copy k of a scene (what the render stages read), on `st`.
// Render copy k of what the render stages read: all of it, or (shading
// only) the normals and material ids, whose buffers the next frame's
// animation and permutation overwrite in place.
int32_t rtx_shade_impl(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, rtx_fb* fb, uint64_t count, int32_t primary,
                       bool unique_pixels, void* stream);
int32_t rtx_refresh_render_copy(rtx_scene* s, int k, hipStream_t st, bool shading_only = false);
int32_t rtx_copy_render_state(rtx_scene* s, rtx_scene::RenderCopy& rc, hipStream_t st, bool shading_only);
int32_t rtx_reserve_render_copy(rtx_scene* s, int k, bool shading_only = false);
// Build into set s->cur with the sort phase on sort_stream and the tree phase
// on tree_stream (which waits for `sorted`, recorded after the sort phase).
int32_t rtx_lbvh_build_split(rtx_scene* s, const rtx_lbvh_params* p, hipStream_t sort_stream, hipEvent_t sorted,
                             hipStream_t tree_stream);
// Allocates set s->cur for builds of `depth` levels (host side).
int32_t rtx_lbvh_reserve(rtx_scene* s, int depth);
// Looks up this device's copy of the allocation-free build's scratch array
// (lbvh.hip); called once by rtx_init with the device current.
int32_t rtx_lbvh_scratch_init(rtx_ctx* c);
int32_t rtx_trace_shadow_instances_impl(rtx_instances* in, rtx_rays* rays, rtx_hits* hits, uint64_t count,
                                        uint64_t* dev_shadow_count, hipStream_t st);
// The CPU SAH builder (tesina-rt_amd/rt/sah.cpp).
int32_t rtx_sah_construct(const float* pos3, uint32_t V, const uint32_t* idx3, uint32_t T,
                          std::vector<rtx_node_ref>& out, std::vector<uint32_t>& order, uint32_t* max_depth);
int32_t rtx_raygen_secondary_impl(rtx_scene* s, rtx_rays* in, rtx_hits* hits, rtx_rays* out, uint64_t count_in,
                                  uint64_t* dev_total_acc, hipStream_t st);

static inline uint32_t rtx_div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ------------------------------------------------------------ wavefront --
// One frame's share of a wavefront step: every frame in flight advances one
// bounce per step, and each stage of the step is one launch over the rays of
// every frame.  The step's table of segments lives in device memory; a wave
// (or block) always works inside one segment, so the scene's pointers are
// uniform across it.
constexpr int RTX_WF_MAX_SEGS = 256;
struct WfSeg {
    // scene (resident, built)
    const float4* node;
    const float4* delta;
    const float4* tri;
    const float4* nrm;
    const float2* uv;
    const rtx_material* mats;
    const int32_t* matmap;
    const rtx_lights* lights;
    const uchar4* tex;
    const TexEnt* tex_tab;
    int32_t ntex, cube;
    // rays of this bounce (count: [0] rays, [1] scan total, [2] first dropped
    // offset, as rtx_rays::count) and of the next one
    float4* ori4;
    float4* dir4;
    int32_t* pixel;
    float* contrib;
    uint32_t* count;
    float4* o_ori4;
    float4* o_dir4;
    int32_t* o_pixel;
    float* o_contrib;
    uint32_t* o_count;
    uint32_t* status;  // look-back words of this step's secondary generation (one per 256 input rays)
    // hit records of this bounce's rays, and the list of rays with a hit
    uint32_t* flags;
    float* t;
    int32_t* id;
    float2* bary;
    float4* n4;
    float4* p4;
    float2* uv2;
    uint32_t* list;
    uint32_t* list_count;
    // framebuffer and the frame's device counters ([0] secondary rays,
    // [1] shadow rays, [2] bounces run)
    uint32_t* accum;
    uchar4* rgba;
    unsigned long long* stats;
    rtx_camera cam;
    int32_t w, h, quad;
    uint32_t o_cap;    // capacity of the next bounce's rays (the renderer's (P << (b + 1)) + 2)
    uint32_t primary;  // bounce 0: primary rays and the framebuffer clear
    uint32_t gen;      // secondary rays are generated this step
    uint32_t resolve;  // last bounce: the framebuffer is resolved to RGBA8
    uint32_t fin;      // resolving frames: slot of its counters in the step's output
};
// Counters of one step in flight, set by the step's first kernel: the trace
// and shadow launches' queues (8 counters each, 64 bytes apart) and the
// secondary generation's ticket (zeroed), and the trace launch's table of
// chunk ends per segment with its total (in-order trace form).
enum {
    WF_Q_TRACE = 0,
    WF_Q_SHADOW = 128,
    WF_Q_SEC = 256,
    WF_Q_ZERO = 272,  // words [0, WF_Q_ZERO) are zeroed
    WF_TAB = 272,
    WF_TOTAL = WF_TAB + RTX_WF_MAX_SEGS,
    WF_STAB = WF_TOTAL + 16,  // the shadow launch's table (set by the plan kernel after the trace)
    WF_STOTAL = WF_STAB + RTX_WF_MAX_SEGS,
    WF_Q_COUNT = WF_STOTAL + 16
};

// Launchers of the segmented stage kernels (trace.hip, shade.hip,
// secondary.hip).  `any_*` says whether some segment needs the stage.
void rtx_wf_init(const WfSeg* segs, uint32_t nseg, uint32_t* ctr, hipStream_t st);
void rtx_wf_raygen(const WfSeg* segs, uint32_t nseg, uint64_t pixels, int cus, hipStream_t st);
// sched 1: persistent waves over per-XCD queues; 0: one chunk per wave in
// launch order over a grid of `max_chunks` (an upper bound of the chunks).
void rtx_wf_trace(const WfSeg* segs, uint32_t nseg, uint32_t* ctr, int var, int tb, int sr, uint32_t coop,
                  uint32_t grain, int sched, uint64_t max_chunks, int cus, int share, hipStream_t st);
void rtx_wf_shadow(const WfSeg* segs, uint32_t nseg, uint32_t* ctr, int var, int tb, int sr, uint32_t coop,
                   uint32_t grain, int sched, uint64_t max_chunks, int cus, int share, bool ordered,
                   hipStream_t st);
void rtx_wf_shade(const WfSeg* segs, uint32_t nseg, int cus, hipStream_t st);
void rtx_wf_secondary(const WfSeg* segs, uint32_t nseg, uint32_t* ctr, int cus, hipStream_t st);
// Also copies the counters of each resolving frame to fin_out[S.fin] (4 x u64).
void rtx_wf_resolve(const WfSeg* segs, uint32_t nseg, uint64_t pixels, unsigned long long* fin_out, int cus,
                    hipStream_t st);

// Wavefront internals for the request server (wavefront.hip).
// before_done(stream, finishing frames) runs after the step's kernels and
// before its completion event is recorded.
int32_t rtx_wavefront_step_impl(rtx_wavefront* wf, hipStream_t st, int32_t* pending, std::vector<int>* finishing,
                                const std::function<int32_t(hipStream_t, const std::vector<int>&)>& before_done);
uint64_t rtx_wf_arena_size(const rtx_wavefront* wf);
int64_t rtx_wf_steps_in_flight(const rtx_wavefront* wf);
// Percent of the resident blocks per CU that this renderer's persistent
// traversal launches take when the wf_share tune is automatic (0).
void rtx_wf_set_auto_share(rtx_wavefront* wf, int percent);
const void* rtx_wf_accum_dev(rtx_wavefront* wf, int32_t frame);
