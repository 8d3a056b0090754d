/*
 * rtx.h — C ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int32 status codes
 * (0 = ok, negative = error; rtx_last_error() gives the text, thread-local).
 * Nothing here ever calls exit().  Every entry point that launches GPU work
 * takes a `stream` (a hipStream_t passed as void*; NULL = the context's own
 * stream) and never synchronizes unless its comment says "synchronizes".
 *
 * Each declaration cites the reference interface it replaces; paths are
 * relative to the Tesina-RT source tree.
 */
#ifndef RTX_H
#define RTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status ---- */
enum rtx_status {
    RTX_OK = 0,
    RTX_EINVAL = -1,     /* bad argument or handle */
    RTX_EDEVICE = -2,    /* HIP runtime error */
    RTX_ENOMEM = -3,     /* device allocation failed or budget exceeded */
    RTX_ESTATE = -4,     /* call out of order (e.g. trace before build) */
    RTX_EUNSUPPORTED = -5
};

/* Pass as a ray count to use the count held on the device by the bundle
 * (set by rtx_raygen_primary / rtx_raygen_secondary); lets a frame run with
 * no host read-back, e.g. inside a captured hipGraph. */
#define RTX_DEVICE_COUNT ((uint64_t)-1)

typedef struct rtx_ctx rtx_ctx;
typedef struct rtx_scene rtx_scene;
typedef struct rtx_rays rtx_rays;
typedef struct rtx_hits rtx_hits;
typedef struct rtx_fb rtx_fb;
typedef struct rtx_renderer rtx_renderer;
typedef struct rtx_server rtx_server;
typedef struct rtx_batch rtx_batch;
typedef struct rtx_pipeline rtx_pipeline;
typedef struct rtx_instances rtx_instances;
typedef struct rtx_obj rtx_obj;
typedef struct rtx_bvh rtx_bvh;

/* ---------------------------------------------------- plain data types -- */

/* material_cl, 32 bytes (include/rt/material.hpp:57-73). texture < 0 = none. */
typedef struct rtx_material {
    float diffuse[4];
    float shininess;
    float reflectiveness;
    float refractive_index;
    int32_t texture;
} rtx_material;

/* Light types: SPOT and DIR as include/rt/light.hpp:69-72; POINT is this
 * build's extension (shadow ray toward pos with tMax = distance). */
enum { RTX_LIGHT_SPOT = 0, RTX_LIGHT_DIR = 1, RTX_LIGHT_POINT = 2 };
#define RTX_MAX_LIGHTS 8

typedef struct rtx_light {
    int32_t type;
    float pos[3];
    float dir[3];   /* DIR: direction the light travels (normalized by the caller
                       as Scene::set_dir_light does, src/rt/scene.cpp:987-1003) */
    float color[3];
    float angle;    /* SPOT: cosine cut-off */
    float radius;
} rtx_light;

/* lights_cl generalised to several lights (include/rt/light.hpp:86-92). */
typedef struct rtx_lights {
    float ambient[3];
    int32_t count;
    rtx_light l[RTX_MAX_LIGHTS];
} rtx_lights;

/* Camera basis as Camera::set produces it (src/rt/camera.cpp:31-59): dir is
 * unit length, right and up are scaled by tan(FOV/2)*aspect and tan(FOV/2). */
typedef struct rtx_camera {
    float pos[3], dir[3], right[3], up[3];
} rtx_camera;

/* RendererConfig knobs on the LBVH path (include/rt/renderer-config.hpp:7-36). */
typedef struct rtx_lbvh_params {
    int32_t depth;     /* bvh_depth: Morton bits, clamped to [1,32] (reference [1,64]) */
    int32_t min_leaf;  /* bvh_min_leaf_size, clamped to [1,256] */
} rtx_lbvh_params;

/* Reference BVHNode layout, 48 bytes (include/rt/bvh.hpp:15-74), produced by
 * rtx_read_nodes_ref for parity checks. */
typedef struct rtx_node_ref {
    float hi[4];
    float lo[4];
    uint32_t a;        /* l_child or start_index */
    uint32_t b;        /* r_child or end_index */
    uint32_t parent;
    int8_t split_axis; /* never written by the LBVH (0 here) */
    int8_t leaf;       /* 0 = internal, nonzero = leaf */
    uint8_t pad[2];
} rtx_node_ref;

/* FrameStats (include/rt/frame-stats.hpp:5-64) of one frame: ray counts as
 * Renderer counts them (src/rt/renderer.cpp:137, 184-185) plus shadow rays,
 * and per-stage device times in ms (filled when profiling is on). */
enum {
    RTX_ST_BVH_BUILD = 0, RTX_ST_PRIM_RAY_GEN, RTX_ST_SEC_RAY_GEN, RTX_ST_PRIM_TRACE, RTX_ST_PRIM_SHADOW_TRACE,
    RTX_ST_SEC_TRACE, RTX_ST_SEC_SHADOW_TRACE, RTX_ST_SHADE, RTX_ST_FB_CLEAR, RTX_ST_FB_COPY, RTX_ST_COUNT
};
typedef struct rtx_frame_stats {
    uint64_t primary_rays;
    uint64_t secondary_rays;
    uint64_t shadow_rays;
    uint32_t bounces_run;
    uint32_t node_count;
    double stage_ms[RTX_ST_COUNT];
    double frame_ms;
} rtx_frame_stats;

/* Renderer settings (Renderer::configure_from_*, src/rt/renderer.cpp:279-348). */
typedef struct rtx_render_params {
    int32_t width, height;
    int32_t quad_size;      /* prim_ray_quad_size, default 32 */
    int32_t max_bounces;
    int32_t depth;          /* bvh_depth */
    int32_t min_leaf;       /* bvh_min_leaf_size */
    uint64_t ray_capacity;  /* secondary-ray capacity per bundle; 0 = 2^max_bounces * pixels */
} rtx_render_params;

/* ------------------------------------------------------------- context -- */
/* Replaces CLInfo::initialize + DeviceInterface::initialize
 * (src/cl-gl/opencl-init.cpp:61-145, src/gpu/interface.cpp:27-38). */
int32_t rtx_init(int32_t device, rtx_ctx** out);
int32_t rtx_shutdown(rtx_ctx* ctx);
const char* rtx_last_error(void);
int32_t rtx_abi_version(void);
/* hipStream_t of the context (as void*). */
void* rtx_stream(rtx_ctx* ctx);
/* DeviceInterface::finish_commands (src/gpu/interface.cpp:273-293); synchronizes. */
int32_t rtx_sync(rtx_ctx* ctx, void* stream);
/* Per-kernel timing with hipEvents around every launch (FrameStats timing,
 * src/rt/renderer.cpp:441-446, without forced per-stage finishes). */
int32_t rtx_profile_enable(rtx_ctx* ctx, int32_t on);
/* 1 when per-kernel timing is on, 0 when off or ctx is null. */
int32_t rtx_profile_enabled(rtx_ctx* ctx);
/* Synchronizes; writes "name count total_ms\n" lines; returns bytes needed. */
int64_t rtx_profile_report(rtx_ctx* ctx, char* buf, uint64_t cap);
int32_t rtx_profile_reset(rtx_ctx* ctx);
/* Device memory held by the library for this context, bytes. */
uint64_t rtx_mem_in_use(rtx_ctx* ctx);
/* Tuning knobs, set by the autotuner (no reference counterpart: the
 * reference fixes its work-group sizes per device, src/gpu/function.cpp:322-385):
 *   "trace_variant": 0 single-node traversal, 1 single node with the near-child
 *                    delta loaded early, 2 sibling-pair traversal, 3 persistent
 *                    waves fed from a ray queue (single-node steps);
 *   "refill":        variant 3: idle lanes of a wave that trigger a refill
 *                    from the queue (1..64);
 *   "trace_block":   64, 128 or 256 threads per traversal work-group;
 *   "xcds":          number of L2 domains the traversal grid is split over
 *                    (work-group i runs on XCD i mod xcds; each XCD gets one
 *                    contiguous, screen-coherent chunk of the rays; 1 = plain
 *                    grid order);
 *   "coop_leaf":     leaves of at least this many triangles are tested by
 *                    the whole wave, 64 triangles per round (0 = never);
 *   "side_stream":   1: a renderer traces each bounce's shadow rays and shades
 *                    it on a second stream while the next bounce is generated
 *                    and traced (frames then capture as forked graphs); 0: all
 *                    stages on the caller's stream (linear graphs).
 *   "inst_lane_roots": instances under the single-node forms (0, 1): 1: each
 *                    lane moves on to its next root as soon as its walk of the
 *                    last one ends, in one loop; 2 (default): the same over
 *                    only the roots whose world-space box (the root box under
 *                    the transform, widened far beyond rounding) the ray
 *                    meets, affine roots only, 64 roots per wave-uniform
 *                    chunk; 0: the wave
 *                    walks the roots one after the other, every lane waiting
 *                    for the wave's slowest walk of each root.
 *   "wf_grain":      wavefront renderer: consecutive 64-ray chunks a
 *                    traversal wave takes per queue atomic (1..64).
 *   "wf_sched":      wavefront closest-hit and shadow launches: 0 one 64-ray
 *                    chunk per wave in launch order (the faster for one large
 *                    frame); 1 (default) persistent waves over per-XCD queues
 *                    (the faster for many frames).
 *   "wf_share":      wavefront persistent launches (wf_sched 1): percent
 *                    (1..100) of the blocks that fit on a CU that they take;
 *                    0 (default) = 100 for a lone renderer and 65 for each of
 *                    a server's groups when it has several, so that one
 *                    group's short launches find room beside another group's
 *                    traversal.
 *   "srv_step_pixels": request server: primary pixels of new requests one
 *                    group's step takes at most (default 2500000; 0 = no
 *                    cap); a step always takes at least one request, and the
 *                    rest wait at the head of the queue for the next step;
 *                    read when a server is created.
 *   "shadow_order":  0 (default): shadow rays visit children in index order
 *                    with no near-child delta fetched wherever the order cannot
 *                    change the result (trees below 32 levels: the any-hit
 *                    result is whether some leaf whose boxes pass at the fixed
 *                    tMax holds a hit); 1: the reference's near-child order
 *                    everywhere;
 *   "stack_lds":     traversal stack entries per lane kept in LDS by the
 *                    single-node and sibling-pair forms (scene and wavefront
 *                    launches): 32 (the whole MAX_LEVELS stack), 16 or 8 (a
 *                    ring of the top levels; deeper levels spill to the
 *                    lane's private memory), trading LDS for resident waves.
 *   "sort_bits":     LBVH radix sort digit width: 8 (default, 4 passes of
 *                    30-bit codes) or 4 (8 passes, longer scatter runs).
 *   "sort_items":    LBVH radix sort pairs per thread of a tile: 16 (default)
 *                    or 32.
 *   "build_grid":    LBVH hierarchy and refit level launches: persistent
 *                    blocks per CU (1..16, default 8).
 * All settings give identical results. */
int32_t rtx_tune_set(rtx_ctx* ctx, const char* key, int64_t value);
/* Current value, or -1 for an unknown key. */
int64_t rtx_tune_get(rtx_ctx* ctx, const char* key);

/* --------------------------------------------------------------- scene -- */
/* Scene::initialize / destroy (src/rt/scene.cpp:163-189). */
int32_t rtx_scene_create(rtx_ctx* ctx, rtx_scene** out);
int32_t rtx_scene_destroy(rtx_scene* s);
/* Vertices SoA: pos V*3 floats, nrm V*3 (may be NULL: normals = 0), uv V*2
 * (may be NULL). Scene::transfer_aggregate_mesh_to_device
 * (src/rt/scene.cpp:237-321); blocking upload. */
int32_t rtx_scene_set_vertices(rtx_scene* s, const float* pos, const float* nrm, const float* uv, uint32_t V);
/* Same from the reference's 80-byte Vertex array (include/rt/mesh.hpp:23-30). */
int32_t rtx_scene_set_vertices_ref(rtx_scene* s, const void* vertex80, uint32_t V);
/* Positions only, asynchronous on `stream` (the per-frame vertex update of
 * Scene::update_aggregate_mesh_vertices, src/rt/scene.cpp:955-985).
 * host_pos: V*3 floats; must stay valid until the stream reaches the copy. */
int32_t rtx_scene_update_positions(rtx_scene* s, const float* host_pos, void* stream);
/* Triangles: idx 3T vertex ids, matmap T material ids (NULL = 0). */
int32_t rtx_scene_set_triangles(rtx_scene* s, const uint32_t* idx, const int32_t* matmap, uint32_t T);
int32_t rtx_scene_set_materials(rtx_scene* s, const rtx_material* mats, uint32_t M);
/* Scene::set_dir_light / set_spot_light / set_ambient_light (src/rt/scene.cpp:987-1040). */
int32_t rtx_scene_set_lights(rtx_scene* s, const rtx_lights* lights);
/* TextureAtlas::load_texture (src/rt/texture-atlas.cpp:25-60) with the images
 * decoded by the caller: `count` RGBA8 images concatenated (texel (x, y) of an
 * image at byte 4*(x + y*width), as create_tex_gl_from_file lays it out,
 * src/cl-gl/opengl-init.cpp:143-200), dims = width, height per image.  A
 * material with texture k >= 0 takes its diffuse colour from image k sampled
 * at the hit's uv (linear filter, repeat, normalized coordinates,
 * src/kernel/ray-shader.cl:145-153, 187-240); an index without an image
 * samples black.  count 0 removes the images. */
#define RTX_MAX_TEXTURES 64
int32_t rtx_scene_set_textures(rtx_scene* s, const uint8_t* rgba, const int32_t* dims, int32_t count);
/* Cubemap::initialize and its enable flag (src/rt/cubemap.cpp:4-57): six RGBA8
 * faces +x, -x, +y, -y, +z, -z concatenated, dims = width, height per face.  A
 * ray without a hit then adds the colour of the cubemap in its direction (linear
 * filter, clamp to edge) times its contribution (src/kernel/ray-shader.cl:275-307).
 * rgba NULL disables the cubemap: a miss adds nothing. */
int32_t rtx_scene_set_cubemap(rtx_scene* s, const uint8_t* rgba, const int32_t* dims);
uint32_t rtx_scene_triangle_count(rtx_scene* s);

/* Travelling-wave displacement of a heightfield, on the device: for every
 * vertex y = sum_k a_k sin(kappa_k (d_k . (x,z)) - omega_k t + phi_k), with
 * analytic normals.  waves: n*6 floats (a, kappa, dx, dz, omega, phi).
 * The dynamic-scene pattern of src/rt/rt-wave.cpp:139-270 (mangle kernel,
 * src/kernel/mangler.cl:13-32). */
int32_t rtx_scene_displace_waves(rtx_scene* s, const float* waves, int32_t n, float t, void* stream);
/* Device-resident form of the same animation, for captured frames (no
 * counterpart in the reference, whose rt-wave loop passes t from the host
 * every frame): set_waves uploads the wave table once (blocking) and resets
 * the device frame clock to 0; animate displaces at t = t0 + dt * clock and
 * then advances the clock by one, both on `stream`. */
int32_t rtx_scene_set_waves(rtx_scene* s, const float* waves, int32_t n);
int32_t rtx_scene_set_clock(rtx_scene* s, uint32_t frame);
int32_t rtx_scene_get_clock(rtx_scene* s, uint32_t* frame);
int32_t rtx_scene_animate(rtx_scene* s, float t0, float dt, void* stream);

/* ----------------------------------------------------------------- LBVH -- */
/* BVHBuilder::build_lbvh (src/rt/bvh-builder.cpp:201-741): prim AABB ->
 * global AABB -> Morton -> radix sort -> permute indices/material map in
 * place -> level-ordered hierarchy -> AABB refit.  Asynchronous. */
int32_t rtx_lbvh_build(rtx_scene* s, const rtx_lbvh_params* p, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving or animation loops):
 * rtx_lbvh_build with the same kernels and the same tree, calling no
 * allocator.  Its scratch is a fixed array of RTX_LBVH_NOALLOC_SCRATCH_BYTES
 * that each GPU holds until the process ends; scenes of more than
 * RTX_LBVH_NOALLOC_MAX_TRIS triangles are rejected (RTX_EINVAL).  The outputs
 * are the scene's tree buffers, which must already be reserved for its
 * triangle count by an earlier rtx_lbvh_build (else RTX_ESTATE).  Every call
 * on one GPU shares that array: calls that overlap on one GPU (other streams,
 * other threads) race.  Asynchronous. */
#define RTX_LBVH_NOALLOC_MAX_TRIS 1048576u
#define RTX_LBVH_NOALLOC_SCRATCH_BYTES 43549184u
int32_t rtx_lbvh_build_noalloc(rtx_scene* s, const rtx_lbvh_params* p, void* stream);
/* BVHBuilder::refit_lbvh (src/rt/bvh-builder.cpp:743-850): AABBs from the
 * current (already permuted) vertices, topology kept.  Asynchronous. */
int32_t rtx_lbvh_refit(rtx_scene* s, void* stream);
/* Node count and per-iteration output counts OC[0..63] (scene.bvh_node_count,
 * bvh_level_node_count; src/rt/bvh-builder.cpp:614-625, 733-736).  Synchronizes. */
int32_t rtx_lbvh_info(rtx_scene* s, uint32_t* node_count, uint32_t* level_counts64);
/* The CPU SAH BVH, the reference's alternative accelerator (BVH::construct_and_map,
 * src/rt/bvh.cpp:7-175, via Scene::create_aggregate_bvh): built on the host
 * from the scene's current triangle order (synchronizes), then the index
 * buffer, material map and upload ids are permuted into its order and every
 * traversal form traces it.  A later rtx_lbvh_build replaces it; refit is
 * LBVH-only.  rtx_lbvh_info / rtx_read_nodes_ref read it back. */
int32_t rtx_sah_build(rtx_scene* s, void* stream);
/* The same builder on host arrays only (pos3 V x 3, idx3 T x 3), for
 * comparison: nodes in the reference layout (cap >= 2T - 1), order =
 * sorted position -> input triangle. */
int32_t rtx_sah_build_host(const float* pos3, uint32_t V, const uint32_t* idx3, uint32_t T, rtx_node_ref* nodes,
                           uint32_t cap, uint32_t* count, uint32_t* order);

/* Parity read-backs (synchronize).  Arrays are in the current sorted order. */
int32_t rtx_read_codes(rtx_scene* s, uint32_t* codes, uint32_t T);      /* sorted Morton codes */
int32_t rtx_read_perm(rtx_scene* s, uint32_t* perm, uint32_t T);        /* sorted pos -> input pos of the last build */
int32_t rtx_read_indices(rtx_scene* s, uint32_t* idx3, uint32_t T);     /* permuted index buffer */
int32_t rtx_read_matmap(rtx_scene* s, int32_t* matmap, uint32_t T);     /* permuted material map */
int32_t rtx_read_orig_ids(rtx_scene* s, uint32_t* orig, uint32_t T);    /* sorted pos -> triangle id at upload */
int32_t rtx_read_global_bbox(rtx_scene* s, float* lo3, float* hi3);
/* Current vertex positions and vertex normals (xyz, V*3 floats each; either
 * may be NULL), as a vertex buffer holds them (hits use their normalization,
 * src/kernel/trace-bvh.cl:222-224): the displaced geometry that the oracle
 * renders in dynamic-scene parity checks. */
int32_t rtx_read_vertices(rtx_scene* s, float* pos3, float* nrm3, uint32_t V);
/* Nodes in the reference's 48-byte layout, canonical level order. */
int32_t rtx_read_nodes_ref(rtx_scene* s, rtx_node_ref* out, uint32_t cap, uint32_t* count);

/* ------------------------------------------------------- ray bundles ---- */
/* RayBundle::initialize / HitBundle::initialize (src/rt/ray.cpp:48-66,
 * 120-140).  Storage is SoA on the device: ori4 = (origin, tMin),
 * dir4 = (direction, tMax), pixel, contribution. */
int32_t rtx_rays_create(rtx_ctx* ctx, uint64_t capacity, rtx_rays** out);
int32_t rtx_rays_destroy(rtx_rays* r);
uint64_t rtx_rays_capacity(rtx_rays* r);
/* Device-held ray count; synchronizes. */
int32_t rtx_rays_count(rtx_rays* r, uint64_t* count);
int32_t rtx_rays_read(rtx_rays* r, float* ori4, float* dir4, int32_t* pixel, float* contrib, uint64_t n);
int32_t rtx_rays_write(rtx_rays* r, const float* ori4, const float* dir4, const int32_t* pixel,
                       const float* contrib, uint64_t n);

int32_t rtx_hits_create(rtx_ctx* ctx, uint64_t capacity, rtx_hits** out);
int32_t rtx_hits_destroy(rtx_hits* h);
/* SampleTraceInfo (src/kernel/trace-bvh.cl:62-74) SoA: flags bit0 hit, bit1
 * inverse_n, bit 8+l shadowed by light l; t; id (sorted triangle position,
 * -1 = miss); bary (u, v); n4; p4 (hit point); uv2.  Any pointer may be NULL. */
int32_t rtx_hits_read(rtx_hits* h, uint32_t* flags, float* t, int32_t* id, float* bary2, float* n4, float* p4,
                      float* uv2, uint64_t n);
int32_t rtx_hits_write(rtx_hits* h, const uint32_t* flags, const float* t, const int32_t* id, const float* n4,
                       const float* p4, uint64_t n);

/* PrimaryRayGenerator::generate (src/rt/primary-ray-generator.cpp:58-165;
 * kernel src/kernel/primary-ray-generator.cl:28-103): rays for samples
 * [offset, offset+count) in quad order; psamples = spp triples (ox, oy,
 * contribution), NULL = one sample {0,0,1}.  Sets the bundle's device count. */
int32_t rtx_raygen_primary(rtx_rays* rays, const rtx_camera* cam, int32_t width, int32_t height, int32_t quad_size,
                           int32_t spp, const float* psamples, uint64_t offset, uint64_t count, void* stream);

/* Tracer::trace (src/rt/tracer.cpp:140-269; src/kernel/trace-bvh.cl:190-481):
 * closest hit + hit completion for `count` rays (or RTX_DEVICE_COUNT). */
int32_t rtx_trace(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream);
/* Tracer::shadow_trace (src/rt/tracer.cpp:273-352;
 * src/kernel/shadow-trace-bvh.cl:160-418): one shadow ray per hit and light. */
int32_t rtx_trace_shadow(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream);

/* A copy of a built scene's accelerator: the reference's bvh_mem
 * DeviceMemory, a copy of the scene's node array that its threaded renderer
 * keeps (SceneGPUCopy, src/rt/scene.cpp:99-128) so that frames trace it
 * while the scene's own BVH is rebuilt.  Here it holds everything a trace
 * reads beside the rays: the node array, near-child deltas and triangle
 * records (in BVH order), the vertex normals and the material map. */
int32_t rtx_bvh_create(rtx_ctx* ctx, rtx_bvh** out);
int32_t rtx_bvh_destroy(rtx_bvh* b);
/* Copies scene s's current accelerator into b on the stream (a device copy;
 * b keeps it until the next copy).  s must be built. */
int32_t rtx_bvh_copy(rtx_bvh* b, rtx_scene* s, void* stream);
/* Tracer::trace(Scene&, DeviceMemory& bvh_mem, ...) (include/rt/tracer.hpp:21,
 * src/rt/tracer.cpp:90-135): rtx_trace of scene s against the copy b instead
 * of the scene's own accelerator (b must be a copy of s: same triangle and
 * vertex counts).  Not thread-safe for one scene (as every call on it). */
int32_t rtx_trace_with(rtx_scene* s, rtx_bvh* b, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream);
/* Tracer::shadow_trace(Scene&, DeviceMemory& bvh_mem, ...)
 * (include/rt/tracer.hpp:26): rtx_trace_shadow against the copy b. */
int32_t rtx_trace_shadow_with(rtx_scene* s, rtx_bvh* b, rtx_rays* rays, rtx_hits* hits, uint64_t count,
                              void* stream);

/* ------------------------------------------------------------ OBJ ingest -- */
/* ModelOBJ::import (src/rt/obj-loader.cpp:163-227) with its MTL reader
 * (:1124-1290), host only: positions, normals and uvs of the deduplicated
 * vertices, triangles in file order (polygons fanned), and material groups.
 * rebuild_normals = the reference's rebuildNormals.  Errors: unreadable
 * file, a face with fewer than 3 corners, an index out of range. */
typedef struct rtx_obj_group {
    uint32_t first_triangle, triangle_count;  /* ModelMesh startIndex / 3, triangleCount */
    int32_t material;                         /* index in the model's material list */
    float ambient[4], diffuse[4], specular[4];
    float shininess, alpha;
    char name[128], color_map[256], bump_map[256];
} rtx_obj_group;
int32_t rtx_obj_load(const char* path, int32_t rebuild_normals, rtx_obj** out);
int32_t rtx_obj_destroy(rtx_obj* m);
int32_t rtx_obj_info(rtx_obj* m, uint32_t* vertices, uint32_t* triangles, uint32_t* groups);
/* ModelOBJ::get_aggregate_mesh (:1354-1390): every vertex and triangle. */
int32_t rtx_obj_read(rtx_obj* m, float* pos3, float* nrm3, float* uv2, uint32_t* idx3);
/* Group g (ModelOBJ::ModelMesh and its Material), in the loader's order:
 * opaque first (by alpha), equal alpha in file order. */
int32_t rtx_obj_group_info(rtx_obj* m, uint32_t g, rtx_obj_group* out);
/* ModelOBJ::get_meshes (:1292-1352) for group g: its triangles with vertices
 * renumbered in order of first use; *vertices = their count (call with
 * NULL arrays to size them; idx3 holds triangle_count x 3).  The reference
 * also leaves triangle_count zero triangles ahead of them (a resize followed
 * by push_back); they are not produced. */
int32_t rtx_obj_read_group(rtx_obj* m, uint32_t g, uint32_t* vertices, float* pos3, float* nrm3, float* uv2,
                           uint32_t* idx3);

/* ------------------------------------------------------------ instances -- */
/* Objects of a scene over per-mesh BVHs (Scene::create_bvhs and
 * Scene::update_bvh_roots, src/rt/scene.cpp:596-700; BVHRoot / MultiBVH,
 * include/rt/multi-bvh.hpp:9-19): instance i is built scene scenes[i] under
 * the row-major 4x4 object-to-world transform tr[16i..16i+15] and its
 * inverse tr_inv.  A scene may back several instances.  Triangle ids are
 * shared: the distinct scenes in order of first use take consecutive blocks
 * (rtx_instances_tri_base gives each instance's first id).  The scenes'
 * current BVHs are used at each trace call, so a scene may be rebuilt (its
 * geometry animated) between calls; one stream at a time uses a set. */
#define RTX_MAX_INSTANCES 4096
int32_t rtx_instances_create(rtx_ctx* ctx, rtx_scene* const* scenes, const float* tr, const float* tr_inv, int32_t n,
                             rtx_instances** out);
int32_t rtx_instances_destroy(rtx_instances* in);
/* New transforms for every instance (Object moves + update_bvh_roots); the
 * next trace call on a stream ships them, in that stream's order. */
int32_t rtx_instances_set_transforms(rtx_instances* in, const float* tr, const float* tr_inv);
/* Lights of the shadow rays (Scene::set_dir_light / set_spot_light). */
int32_t rtx_instances_set_lights(rtx_instances* in, const rtx_lights* lights);
int32_t rtx_instances_tri_base(rtx_instances* in, uint32_t* base);
/* Tracer::trace with the multi-root kernel (src/rt/tracer.cpp:97, 223;
 * trace_multi + complete_transformed_hit_info, src/kernel/trace-bvh.cl:123-183,
 * 419-463): every root traced from the ray's tMax in its own frame; t is the
 * world distance to the nearest hit, the first root wins ties; the normal is
 * tr times the interpolated object normal, facing the ray. */
int32_t rtx_trace_instances(rtx_instances* in, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream);
/* Tracer::shadow_trace with shadow_trace_multi (src/rt/tracer.cpp:301, 363;
 * src/kernel/shadow-trace-bvh.cl:317-373): one ray per hit and light, tMin
 * 0.01, tested against every root; a spot light lights, untraced, a point
 * with dot(toward the light, spot dir) < angle, as that kernel does. */
int32_t rtx_trace_shadow_instances(rtx_instances* in, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream);
/* Object::mat for every instance.  As Scene::create_bvhs (src/rt/scene.cpp:652-670)
 * builds the material map, a scene's triangles take the material of the
 * first instance that uses it (later instances of a scene share it). */
int32_t rtx_instances_set_materials(rtx_instances* in, const rtx_material* per_instance);
/* The scene that shades the instances' hits and spawns their secondary rays
 * (the shared material map and materials of Scene::create_bvhs,
 * src/rt/scene.cpp:652-670, with the set's lights).  It holds no geometry;
 * textures and the cubemap are set on it with rtx_scene_set_textures /
 * rtx_scene_set_cubemap and stay across calls.  Owned by the instance set
 * (destroyed with it); rtx_instances_set_materials must have been called.
 * Refreshed by each call when the scenes' triangle counts changed. */
int32_t rtx_instances_shading_scene(rtx_instances* in, rtx_scene** out);

/* SecondaryRayGenerator::generate, scan variant (src/rt/secondary-ray-generator.cpp:57-145,
 * 248-340; src/kernel/secondary-ray-generator.cl:43-181): Whitted children in
 * parent order, refraction before reflection; a parent whose children would
 * reach rays_out's capacity emits none.  The output count stays on the
 * device; if count_out is not NULL it is also read back (synchronizes), as
 * the reference does. */
int32_t rtx_raygen_secondary(rtx_scene* s, rtx_rays* rays_in, rtx_hits* hits, rtx_rays* rays_out, uint64_t count_in,
                             uint64_t* count_out, void* stream);

/* ---------------------------------------------------------- framebuffer -- */
/* FrameBuffer::initialize / clear / copy (src/rt/framebuffer.cpp:118-165;
 * src/kernel/framebuffer.cl:21-54): uint32 r,g,b accumulators x1e6 and an
 * RGBA8 image. */
int32_t rtx_fb_create(rtx_ctx* ctx, int32_t width, int32_t height, rtx_fb** out);
int32_t rtx_fb_destroy(rtx_fb* fb);
int32_t rtx_fb_clear(rtx_fb* fb, void* stream);
int32_t rtx_fb_resolve(rtx_fb* fb, void* stream);
int32_t rtx_fb_read_accum(rtx_fb* fb, uint32_t* rgb3);
int32_t rtx_fb_read_rgba8(rtx_fb* fb, uint8_t* rgba);

/* RayShader::shade (src/rt/ray-shader.cpp:33-154; src/kernel/ray-shader.cl:116-459):
 * ambient + Lambert cos^2 + Phong^8 per unshadowed light, times
 * (1 - reflectiveness), floor(clamp(c*val)*1e6) added to the pixel with
 * integer atomics (order-free).  `primary` is accepted for API parity. */
int32_t rtx_shade(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, rtx_fb* fb, uint64_t count, int32_t primary,
                  void* stream);

/* ------------------------------------------------------------- renderer -- */
/* Renderer (src/rt/renderer.cpp:12-272): owns two ray bundles, a hit bundle
 * and a framebuffer sized from params. */
int32_t rtx_renderer_create(rtx_ctx* ctx, const rtx_render_params* p, rtx_renderer** out);
int32_t rtx_renderer_destroy(rtx_renderer* r);
rtx_fb* rtx_renderer_fb(rtx_renderer* r);
/* One frame, plain form: set_up_frame (clear + build if `rebuild`) +
 * render_to_framebuffer + copy_framebuffer, one launch per stage and the
 * reference's per-bounce count read-back.  Synchronizes at the end; fills
 * stats if not NULL. */
int32_t rtx_render_frame(rtx_renderer* r, rtx_scene* s, const rtx_camera* cam, int32_t rebuild,
                         rtx_frame_stats* stats, void* stream);
/* One frame with no host read-back (device-held counts), asynchronous; the
 * form that rtx_render_graph captures. */
int32_t rtx_render_frame_async(rtx_renderer* r, rtx_scene* s, const rtx_camera* cam, int32_t rebuild,
                               void* stream);
/* Ray counts of the last frame (synchronizes). */
int32_t rtx_renderer_last_stats(rtx_renderer* r, rtx_frame_stats* stats);
/* Renderer::render_to_framebuffer over a scene of objects (root_count() > 1):
 * the reference's per-bounce sequence with the multi-root trace and shadow
 * rays, secondary rays and shading through the instances' material map. */
int32_t rtx_render_instances(rtx_renderer* r, rtx_instances* in, const rtx_camera* cam, rtx_frame_stats* stats,
                             void* stream);
/* Captures one rtx_render_frame_async into a hipGraph kept by the renderer
 * (no counterpart in the reference, which launches every stage from the
 * host).  Render one frame eagerly first so that every buffer exists.  The
 * camera and scene handles are fixed at capture; per-frame vertex updates or
 * displacement are enqueued on the stream before each replay. */
int32_t rtx_renderer_capture(rtx_renderer* r, rtx_scene* s, const rtx_camera* cam, int32_t rebuild, void* stream);
/* Launches the captured frame; asynchronous. */
int32_t rtx_renderer_replay(rtx_renderer* r, void* stream);
/* Dynamic scene (src/rt/rt-wave.cpp:139-270): when enabled, every frame
 * (plain, read-back-free or captured) starts with rtx_scene_animate(s, t0,
 * dt) before the LBVH rebuild. */
int32_t rtx_renderer_set_animation(rtx_renderer* r, int32_t enable, float t0, float dt);
/* Pipelined read-back-free frames (RendererT, src/rt/renderer-threaded.cpp:
 * 452-506): while a frame renders, the next frame's animation and LBVH build
 * run on a second stream; the image of each frame shows the scene as built
 * one frame earlier.  Takes effect at the next rtx_render_frame_async /
 * rtx_renderer_capture (capture after one eager pipelined frame). */
int32_t rtx_renderer_set_pipelined(rtx_renderer* r, int32_t enable);

/* Bounces of the next frames, at most the max_bounces the renderer was
 * created with (the reference's RendererConfig::max_bounces, per frame). */
int32_t rtx_renderer_set_bounces(rtx_renderer* r, int32_t bounces);

/* ----------------------------------------------------------- wavefront -- */
/* Frames of many scenes rendered together (no counterpart in the reference,
 * which renders one frame of one scene at a time: Renderer::
 * render_to_framebuffer, src/rt/renderer.cpp:113-214).  Each step advances
 * every admitted frame by one bounce of that sequence -- primary rays (new
 * frames), closest hit, shadow rays, shading, secondary rays, resolve (frames
 * that end) -- with one launch per stage over all frames' rays: persistent
 * waves take 64-ray chunks of one frame each from a queue, frames grouped
 * by scene.  Each frame's accumulators, RGBA8 and counts equal its
 * stand-alone frame's (rtx_render_frame with rebuild = 0) bit for bit.
 * A frame's rays, hits and framebuffer come from a device arena allocated
 * once at creation: admission charges rtx_wavefront_frame_bytes() against
 * it, and frames never allocate.  One thread drives a wavefront. */
typedef struct rtx_wavefront rtx_wavefront;
#define RTX_WF_MAX_FRAMES 256
typedef struct rtx_wavefront_params {
    uint64_t arena_bytes;  /* device memory for frames in flight; 0 = 8 GiB */
    int32_t max_frames;    /* frames admitted and not released at once, 1..RTX_WF_MAX_FRAMES */
    int32_t quad_size;     /* prim_ray_quad_size, 0 = 32 */
} rtx_wavefront_params;
int32_t rtx_wavefront_create(rtx_ctx* ctx, const rtx_wavefront_params* p, rtx_wavefront** out);
/* Waits for enqueued steps, then frees. */
int32_t rtx_wavefront_destroy(rtx_wavefront* wf);
/* Arena bytes of a frame of width x height pixels and `bounces` bounces
 * (two ray buffers sized like the stand-alone renderer's largest bundles,
 * hit records, framebuffer, counters); 0 for bad sizes. */
uint64_t rtx_wavefront_frame_bytes(int32_t width, int32_t height, int32_t bounces, uint64_t ray_capacity);
/* Admits a frame of a built scene (one sample per pixel).  ray_capacity:
 * secondary rays per bounce at most, as rtx_render_params::ray_capacity (0 =
 * the full (pixels << bounce) + 2).  RTX_ENOMEM: no room in the arena now,
 * or max_frames unreleased frames (release frames and retry); RTX_EINVAL:
 * larger than the whole arena.  No GPU work. */
int32_t rtx_wavefront_add(rtx_wavefront* wf, rtx_scene* s, const rtx_camera* cam, int32_t width, int32_t height,
                          int32_t bounces, uint64_t ray_capacity, uint64_t tag, int32_t* frame);
/* Enqueues one step on `stream` (asynchronous; waits only when the step
 * table it reuses is still in flight).  *pending = frames that still need
 * steps. */
int32_t rtx_wavefront_step(rtx_wavefront* wf, void* stream, int32_t* pending);
/* Steps until every admitted frame has enqueued its last bounce (asynchronous). */
int32_t rtx_wavefront_run(rtx_wavefront* wf, void* stream);
/* Waits for every enqueued step (synchronizes). */
int32_t rtx_wavefront_sync(rtx_wavefront* wf);
/* Frames completed on the device since the last poll (non-blocking). */
int32_t rtx_wavefront_poll(rtx_wavefront* wf, int32_t* frames, uint64_t* tags, uint32_t cap, uint32_t* n);
/* Counts (primary, secondary, shadow rays, bounces run) and images of a
 * frame whose last bounce is enqueued; they wait for it. */
int32_t rtx_wavefront_frame_stats(rtx_wavefront* wf, int32_t frame, rtx_frame_stats* stats);
int32_t rtx_wavefront_read_accum(rtx_wavefront* wf, int32_t frame, uint32_t* rgb3);
int32_t rtx_wavefront_read_rgba8(rtx_wavefront* wf, int32_t frame, uint8_t* rgba);
/* Returns a complete frame's memory to the arena. */
int32_t rtx_wavefront_release(rtx_wavefront* wf, int32_t frame);
int32_t rtx_wavefront_memory(rtx_wavefront* wf, uint64_t* used, uint64_t* capacity);

/* ------------------------------------------------------------- serving -- */
/* Request server over resident scenes (no counterpart in the reference,
 * whose apps render one scene from one thread).  A dispatcher thread of the
 * library admits requests at their arrival time, first come first served,
 * when their rays, hits and framebuffer (rtx_wavefront_frame_bytes) fit the
 * request arena and fewer than `workers` are in flight (a request larger
 * than the arena is refused with RTX_ENOMEM), and renders them with the
 * wavefront renderer: every step advances all requests in flight by one
 * bounce in shared launches, so requests that arrive while others run join
 * at the next step whatever their scene, resolution or depth (no rebuild:
 * scenes are resident and built).  Completion is timestamped when the
 * device has finished the request.  Scenes are admitted against an HBM
 * budget that also holds the arena.  With `plain` set the server is the
 * per-request baseline: one request at a time, the plain frame
 * (rtx_render_frame) on a renderer per resolution class. */
#define RTX_SERVER_MAX_CLASSES 4
typedef struct rtx_server_params {
    int32_t workers;          /* requests in flight at most (1..RTX_WF_MAX_FRAMES); plain: ignored */
    int32_t max_bounces;      /* largest request depth (1..16) */
    int32_t classes;          /* plain mode: resolution classes (1..RTX_SERVER_MAX_CLASSES); else unused */
    int32_t class_w[RTX_SERVER_MAX_CLASSES], class_h[RTX_SERVER_MAX_CLASSES];
    uint64_t hbm_budget;      /* bytes for the request memory + resident scenes; 0 = 90% of free device memory */
    int32_t plain;            /* 1: baseline, one request at a time with the plain frame */
    int32_t queue_cap;        /* submissions held at most (0 = 65536) */
    uint64_t arena_bytes;     /* request arena (part of the budget); 0 = half the budget */
    int32_t groups;           /* wavefront renderers on streams of their own, sharing the arena and the
                                 requests in flight (1..8); 0 = 2 */
    int32_t steps_ahead;      /* wavefront steps a group keeps enqueued (1..4): fewer, the sooner a new
                                 request joins; more, the fewer idle gaps between steps; 0 = 1 */
} rtx_server_params;

typedef struct rtx_request {
    int32_t scene;            /* id from rtx_server_add_scene */
    int32_t width, height;    /* any size (plain mode: one of the resolution classes) */
    int32_t bounces;          /* 0..max_bounces */
    rtx_camera camera;
    int64_t arrival_ns;       /* rtx_now_ns() clock; not started before it */
    uint64_t tag;             /* returned with the completion */
    void* out_accum;          /* optional host copy of the accumulators (w*h*3 uint32), may be NULL */
} rtx_request;

typedef struct rtx_completion {
    uint64_t tag;
    int64_t arrival_ns, start_ns, done_ns;  /* start: admitted; done: completion observed on the device */
    uint64_t rays;                          /* primary + secondary rays */
    uint64_t shadow_rays;
    int32_t status;                         /* RTX_OK or the error of this request */
    int32_t worker;                         /* wavefront frame slot (plain: resolution class) */
} rtx_completion;

/* Monotonic clock of the server, nanoseconds. */
int64_t rtx_now_ns(void);
int32_t rtx_server_create(rtx_ctx* ctx, const rtx_server_params* p, rtx_server** out);
/* Stops the dispatcher after the queued requests finish, then frees. */
int32_t rtx_server_destroy(rtx_server* srv);
/* Makes a built scene resident: RTX_ENOMEM when it does not fit the budget
 * next to the workers and the scenes already resident. */
int32_t rtx_server_add_scene(rtx_server* srv, rtx_scene* s, int32_t* scene_id);
/* Device bytes held by the workers and by the resident scenes. */
int32_t rtx_server_memory(rtx_server* srv, uint64_t* workers_bytes, uint64_t* scenes_bytes);
/* Queues n requests (thread-safe); RTX_ENOMEM when the queue is full. */
int32_t rtx_server_submit(rtx_server* srv, const rtx_request* reqs, uint32_t n);
/* Moves up to cap finished requests into out; *n = how many. */
int32_t rtx_server_poll(rtx_server* srv, rtx_completion* out, uint32_t cap, uint32_t* n);
/* Waits until every submitted request has finished (or timeout_ms passed:
 * RTX_ESTATE). */
int32_t rtx_server_drain(rtx_server* srv, int64_t timeout_ms);

/* ------------------------------------------------------- frame pipeline -- */
/* Frames of one scene with `depth` (1..8) frames in flight: the reference's
 * RendererT (src/rt/renderer-threaded.cpp:452-506) overlaps the next build
 * with the last render; this goes further.  Frame N uses slot N mod depth:
 * its own renderer, stream, LBVH buffer set and copy of the normals and
 * material ids.  A build stream runs, frame after frame, the animation (if
 * enabled) and the sort phase of each build (codes, sort, the in-place
 * permutation that chains one build to the next), then copies the normals
 * and material ids; the slot's stream builds the tree into the slot's set and
 * renders the frame from it.  A slot's set and copy are rewritten only once
 * frame N - depth has finished with them.  So builds and renders of
 * consecutive frames overlap.  Each frame's image equals the sequential
 * run's.  The scene must not be driven by another renderer while the
 * pipeline runs; afterwards its current buffer set holds the last frame's
 * build.  Each slot renders on one stream (depth + 1 streams with the build
 * stream); which depth is fastest depends on how those streams share the
 * process's hardware queues (GPU_MAX_HW_QUEUES), so time a few depths. */
int32_t rtx_pipeline_create(rtx_ctx* ctx, rtx_scene* s, const rtx_render_params* p, int32_t depth,
                            rtx_pipeline** out);
int32_t rtx_pipeline_destroy(rtx_pipeline* pl);
/* Animation at the start of every frame (rtx_scene_animate on the build stream). */
int32_t rtx_pipeline_set_animation(rtx_pipeline* pl, int32_t enable, float t0, float dt);
/* Enqueues the next frame (asynchronous); *frame = its index (0, 1, ...). */
int32_t rtx_pipeline_frame(rtx_pipeline* pl, const rtx_camera* cam, int32_t rebuild, uint64_t* frame);
/* Waits for `frame`; its framebuffer and stats stay valid until frame+depth
 * is enqueued. */
int32_t rtx_pipeline_wait(rtx_pipeline* pl, uint64_t frame);
rtx_fb* rtx_pipeline_fb(rtx_pipeline* pl, uint64_t frame);
int32_t rtx_pipeline_stats(rtx_pipeline* pl, uint64_t frame, rtx_frame_stats* stats);

/* ------------------------------------------------------------ batches -- */
/* A batch of frames of different scenes rendered together (no counterpart
 * in the reference, which renders one scene per process): frame i renders
 * scenes[i] from cams[i] with renderers[i] (resident scenes, no rebuild).
 * The frames are spread round-robin over `streams` internal streams forked
 * from and joined back to the caller's stream, so that small frames run side
 * by side; the whole batch can be captured into one hipGraph.  Handles must
 * outlive the batch. */
int32_t rtx_batch_create(rtx_ctx* ctx, rtx_renderer* const* renderers, rtx_scene* const* scenes,
                         const rtx_camera* cams, int32_t n, int32_t streams, rtx_batch** out);
int32_t rtx_batch_destroy(rtx_batch* b);
/* Enqueues every frame of the batch on `stream` (asynchronous). */
int32_t rtx_batch_render(rtx_batch* b, void* stream);
/* Captures rtx_batch_render into a hipGraph (render the batch eagerly once
 * first), and launches it. */
int32_t rtx_batch_capture(rtx_batch* b, void* stream);
int32_t rtx_batch_replay(rtx_batch* b, void* stream);
/* Sum of the frames' ray counts of the last batch (synchronizes). */
int32_t rtx_batch_stats(rtx_batch* b, uint64_t* rays, uint64_t* shadow_rays);

/* --------------------------------------------------------- host helpers -- */
/* Camera::set (src/rt/camera.cpp:31-59); host only. */
int32_t rtx_camera_set(rtx_camera* cam, const float pos[3], const float dir[3], const float up[3], float fov,
                       float aspect, int32_t right_handed);

#ifdef __cplusplus
}
#endif
#endif /* RTX_H */
