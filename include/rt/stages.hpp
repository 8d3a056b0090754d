// C++ stage classes of the ray-tracing hot path, with the names, methods and
// 0 / -1 return convention of the reference's classes, implemented over the
// C ABI (rtx.h) so that the reference's Renderer-level code drops in.
//
// Reference interfaces mirrored (paths relative to the Tesina-RT tree):
//   DeviceInterface        include/gpu/interface.hpp:20-80 (device context only)
//   Camera                 include/rt/camera.hpp, src/rt/camera.cpp:31-59
//   RendererConfig         include/rt/renderer-config.hpp:7-36
//   FrameStats, rt_stage   include/rt/frame-stats.hpp:5-64
//   Scene (geometry, materials, lights, accelerator state)  include/rt/scene.hpp
//   RayBundle, HitBundle   include/rt/ray.hpp:63-102
//   FrameBuffer            include/rt/framebuffer.hpp:14-47
//   Log                    include/misc/log.hpp:9-58
//   BVHBuilder             include/rt/bvh-builder.hpp:16-28
//   PrimaryRayGenerator    include/rt/primary-ray-generator.hpp:19-54
//   Tracer                 include/rt/tracer.hpp:12-70
//   SecondaryRayGenerator  include/rt/secondary-ray-generator.hpp
//   RayShader              include/rt/ray-shader.hpp:14-35
//   Renderer               include/rt/renderer.hpp, src/rt/renderer.cpp:12-272
//
// Differences by design: device memory lives behind opaque rtx handles
// instead of DeviceMemory ids (the accelerator copy the reference traces
// from a DeviceMemory is a BVHMemory here); FrameBuffer::copy resolves into a device RGBA8
// image (read with read_rgba8) instead of a GL texture; no OpenGL or OpenCL;
// TextureAtlas and Cubemap take decoded RGBA8 images instead of file names.
#pragma once
#ifndef RT_STAGES_HPP
#define RT_STAGES_HPP

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../rtx.h"

struct vec3 {
    float x, y, z;
};

// Log (include/misc/log.hpp:9-58): text to stdout unless silent, and to a
// file once initialize() opened one.
class Log {
public:
    // Opens the file; returns true when it could not be opened (as the
    // reference's does).
    bool initialize(const std::string& file = "log") {
        enabled = true;
        silent = false;
        m_file.open(file.c_str(), std::fstream::out);
        return !m_file.is_open();
    }
    void out(const std::string& str) { m_file << str << std::endl; }
    std::ostream& o() { return m_file; }
    bool enabled = false;
    bool silent = false;

private:
    std::fstream m_file;
};

template <typename T>
Log& operator<<(Log& l, const T& t) {
    if (!l.silent) std::cout << t;
    if (l.enabled) l.o() << t;
    return l;
}

// Process-wide device context (the reference's DeviceInterface/CLInfo
// singletons, src/gpu/interface.cpp:27-38).
class DeviceInterface {
public:
    static DeviceInterface* instance();
    int32_t initialize(int device = 0);
    bool good() const { return m_ctx != nullptr; }
    rtx_ctx* ctx() const { return m_ctx; }
    void finish_commands();
    int32_t shutdown();

private:
    rtx_ctx* m_ctx = nullptr;
};

struct pixel_sample_cl {
    float ox, oy, contribution;
};

class Camera {
public:
    Camera();
    Camera(const vec3& position, const vec3& direction, const vec3& up, float fov, float aspect,
           bool right_handed = false);
    void set(const vec3& position, const vec3& direction, const vec3& up, float fov, float aspect,
             bool right_handed = false);
    vec3 pos, dir, right, up;
    float fov, aspect;
    bool right_handed;
    rtx_camera basis() const;
};

struct RendererConfig {
    RendererConfig();
    int use_lbvh;
    int bvh_refit_only;
    int bvh_depth;
    int bvh_min_leaf_size;
    int sec_ray_use_atomics;
    int sec_ray_use_disc;
    int prim_ray_quad_size;
    int prim_ray_use_zcurve;
    double tile_to_cores_ratio;
};

enum rt_stage {
    BVH_BUILD = 0,
    PRIM_RAY_GEN,
    SEC_RAY_GEN,
    PRIM_TRACE,
    PRIM_SHADOW_TRACE,
    SEC_TRACE,
    SEC_SHADOW_TRACE,
    SHADE,
    FB_CLEAR,
    FB_COPY,
    STAGE_COUNT
};

struct FrameStats {
    double get_stage_time(rt_stage s) const { return stage_times[s]; }
    double get_stage_mean_time(rt_stage s) const { return acc_frames ? stage_acc_times[s] / acc_frames : 0; }
    double get_frame_time() const { return frame_time; }
    double get_mean_frame_time() const { return acc_frames ? frame_acc_time / acc_frames : 0; }
    size_t get_ray_count() const { return total_ray_count; }
    size_t get_secondary_ray_count() const { return total_sec_ray_count; }
    void clear_times() {
        for (int i = 0; i < STAGE_COUNT; ++i) stage_times[i] = 0;
    }
    void clear_mean_times() {
        acc_frames = 0;
        frame_acc_time = 0;
        for (int i = 0; i < STAGE_COUNT; ++i) stage_acc_times[i] = 0;
    }
    double stage_times[STAGE_COUNT] = {};
    double stage_acc_times[STAGE_COUNT] = {};
    uint32_t acc_frames = 0;
    double frame_time = 0;
    double frame_acc_time = 0;
    size_t total_ray_count = 0;
    size_t total_sec_ray_count = 0;
};

enum AcceleratorType { LBVH_ACCELERATOR, SAH_BVH_ACCELERATOR, KDTREE_ACCELERATOR };

typedef int32_t texture_id;

// TextureAtlas (include/rt/texture-atlas.hpp:14-40; src/rt/texture-atlas.cpp:25-60)
// with the images decoded by the caller: RGBA8, texel (x, y) at byte
// 4*(x + y*width), as create_tex_gl_from_file lays them out.  Materials refer
// to an image by its id; the images reach the device before the next shading.
class TextureAtlas {
public:
    int32_t initialize() { return 0; }
    int32_t destroy();
    texture_id load_texture_rgba(const uint8_t* rgba, uint32_t width, uint32_t height);
    size_t texture_count() const { return m_dims.size() / 2; }
    static const texture_id invalid_tex_id = -1;

private:
    friend class Scene;
    std::vector<uint8_t> m_rgba;
    std::vector<int32_t> m_dims;
    bool m_dirty = false;
};

// Cubemap (include/rt/cubemap.hpp; src/rt/cubemap.cpp:4-57) from six decoded
// RGBA8 faces, +x, -x, +y, -y, +z, -z; `enabled` as the reference's flag
// (set by initialize).  Misses take its colour when enabled.
class Cubemap {
public:
    int32_t initialize_rgba(const uint8_t* const faces[6], const uint32_t dims[12]);
    int32_t destroy();
    bool enabled = false;

private:
    friend class Scene;
    std::vector<uint8_t> m_rgba;
    std::vector<int32_t> m_dims;
    uint64_t m_version = 0;
};

class Scene {
public:
    Scene();
    ~Scene();
    int32_t initialize();
    int32_t destroy();
    bool valid() const { return m_h != nullptr; }
    bool ready() const { return m_built || m_bvhs_built; }
    // Geometry (the aggregate mesh of Scene::transfer_aggregate_mesh_to_device,
    // src/rt/scene.cpp:237-321): pos/nrm V x 3, uv V x 2, idx T x 3.
    int32_t set_geometry(const float* pos, const float* nrm, const float* uv, uint32_t V, const uint32_t* idx,
                         const int32_t* matmap, uint32_t T);
    int32_t set_vertices_ref(const void* vertex80, uint32_t V);  // 80-byte Vertex array
    int32_t set_triangles(const uint32_t* idx, const int32_t* matmap, uint32_t T);
    int32_t update_vertex_positions(const float* pos);
    int32_t set_materials(const std::vector<rtx_material>& mats);
    // Lights (src/rt/scene.cpp:987-1040).  set_dir_light / set_spot_light
    // replace the light list with one light, as the reference's single slot.
    int32_t set_dir_light(const vec3& dir, const vec3& color);
    int32_t set_spot_light(const vec3& pos, const vec3& dir, const vec3& color, float angle, float radius);
    int32_t set_ambient_light(const vec3& color);
    int32_t add_point_light(const vec3& pos, const vec3& color);  // extension of this build
    int32_t set_lights(const rtx_lights& lights);                  // the whole light record
    size_t triangle_count() const { return m_T; }
    AcceleratorType get_accelerator_type() const { return LBVH_ACCELERATOR; }
    rtx_scene* handle() const { return m_h; }

    // Objects over per-mesh BVHs (Scene::add_object, create_bvhs,
    // update_bvh_roots, root_count: src/rt/scene.cpp:596-700).  Meshes come
    // from add_mesh (in place of load_obj_file); an object is a mesh under
    // Object::geom's transform, given as its matrices (row-major 4x4,
    // GeometricProperties::getTransformMatrix / getTransformMatrixInv).  The
    // per-mesh trees are LBVHs.  With objects and no aggregate mesh,
    // root_count() is the object count, and the Tracer takes the multi-root
    // kernels when it exceeds 1 (src/rt/tracer.cpp:97, 223); one object is
    // traced as the single-root kernel traces it: its mesh, untransformed.
    int32_t add_mesh(const float* pos, const float* nrm, const float* uv, uint32_t V, const uint32_t* idx,
                     uint32_t T);  // mesh id, or -1
    // Scene::load_obj_file_as_aggregate / load_obj_file (src/rt/scene.cpp:395-470)
    // through the library's OBJ loader: the whole model as one mesh, or one
    // mesh per material group.
    int32_t load_obj_file_as_aggregate(const std::string& filename);  // mesh id, or -1
    std::vector<int32_t> load_obj_file(const std::string& filename);
    int32_t add_object(int32_t mesh);  // object id, or -1
    int32_t set_object_transform(int32_t object, const float* tr16, const float* tr_inv16);
    int32_t set_object_material(int32_t object, const rtx_material& mat);  // Object::mat
    // Scene::create_aggregate_mesh (src/rt/scene.cpp:495-539): every object's
    // mesh with its vertices transformed on the host (M (p, 1) and M (n, 0),
    // homogenized), one material per object, as this scene's geometry; the
    // BVH builder then builds it and root_count() is 1.
    int32_t create_aggregate_mesh();
    int32_t create_bvhs(int depth = 30, int min_leaf = 1);
    int32_t update_bvh_roots();
    size_t object_count() const { return m_obj_mesh.size(); }
    size_t root_count() const { return m_built ? 1 : m_obj_mesh.size(); }
    bool aggregate_built() const { return m_built; }
    rtx_instances* instances() const { return m_inst; }
    rtx_scene* mesh_handle(int32_t mesh) const { return m_meshes[(size_t)mesh]; }
    int32_t object_mesh(int32_t object) const { return m_obj_mesh[(size_t)object]; }

    // The scene whose materials shade this scene's hits and spawn its
    // secondary rays: the aggregate mesh's own, or, for objects over per-mesh
    // BVHs, one material per mesh, its first object's (Scene::create_bvhs,
    // src/rt/scene.cpp:652-670).  nullptr on error.
    rtx_scene* shading_handle();
    // The atlas's images and the cubemap `cm` (enabled or not) on the device
    // for the shading stage; uploads only what changed.
    int32_t sync_images(const Cubemap& cm);

    Camera camera;
    TextureAtlas texture_atlas;
    Cubemap cubemap;
    uint32_t bvh_node_count = 0;
    uint32_t bvh_level_node_count[64] = {};

private:
    friend class BVHBuilder;
    int32_t push_lights();
    rtx_scene* m_img_target = nullptr;   // the scene that holds the uploaded images
    const Cubemap* m_cube_on = nullptr;  // the cubemap on the device, and its version
    uint64_t m_cube_version = 0;
    rtx_scene* m_h = nullptr;
    rtx_lights m_lights{};
    size_t m_T = 0, m_V = 0;
    bool m_built = false;
    std::vector<rtx_scene*> m_meshes;
    struct HostMesh {
        std::vector<float> pos, nrm, uv;
        std::vector<uint32_t> idx;
    };
    std::vector<HostMesh> m_host;  // the meshes as given (create_aggregate_mesh reads them)
    std::vector<int32_t> m_obj_mesh;
    std::vector<rtx_material> m_obj_mat;
    std::vector<float> m_tr, m_tr_inv;
    rtx_instances* m_inst = nullptr;
    bool m_bvhs_built = false;
};

class RayBundle {
public:
    RayBundle() = default;
    ~RayBundle();
    int32_t initialize(size_t rays);
    int32_t resize(size_t rays);
    bool valid() const { return m_h != nullptr; }
    int32_t count() const { return (int32_t)m_count; }
    rtx_rays* handle() const { return m_h; }

private:
    rtx_rays* m_h = nullptr;
    size_t m_count = 0;
};

class HitBundle {
public:
    HitBundle() = default;
    ~HitBundle();
    int32_t initialize(size_t sz);
    int32_t resize(size_t sz);
    bool valid() const { return m_h != nullptr; }
    int32_t count() const { return (int32_t)m_count; }
    rtx_hits* handle() const { return m_h; }

private:
    rtx_hits* m_h = nullptr;
    size_t m_count = 0;
};


class FrameBuffer {
public:
    FrameBuffer() = default;
    ~FrameBuffer();
    int32_t initialize(size_t sz[2]);
    int32_t resize(size_t sz[2]);
    bool initialized() const { return m_h != nullptr; }
    int32_t clear();
    int32_t copy();  // accumulators -> device RGBA8 (src/kernel/framebuffer.cl:31-54)
    int32_t read_rgba8(uint8_t* rgba) const;
    int32_t read_accum(uint32_t* rgb) const;
    void timing(bool b) { m_timing = b; }
    double get_clear_exec_time() const { return m_clear_ms; }
    double get_copy_exec_time() const { return m_copy_ms; }
    rtx_fb* handle() const { return m_h; }
    size_t width() const { return m_size[0]; }
    size_t height() const { return m_size[1]; }

private:
    rtx_fb* m_h = nullptr;
    size_t m_size[2] = {0, 0};
    bool m_timing = false;
    double m_clear_ms = 0, m_copy_ms = 0;
};

class BVHBuilder {
public:
    int32_t initialize();
    int32_t build_lbvh(Scene& scene, size_t cq_i = 0);
    int32_t refit_lbvh(Scene& scene, size_t cq_i = 0);
    // With logging on and a log set, build_lbvh and refit_lbvh write their
    // stage times to the log (src/rt/bvh-builder.cpp:266-270, 351-354,
    // 456-459, 791-795); the build then waits for the device.
    void set_log(Log* log) { m_log = log; }
    void timing(bool b) { m_timing = b; }
    void logging(bool l) { m_logging = l; }
    double get_exec_time() const { return m_time_ms; }
    void update_configuration(const RendererConfig& conf);

private:
    int32_t log_stages(const char* title);
    Log* m_log = nullptr;
    bool m_timing = true, m_logging = false;
    double m_time_ms = 0;
    int m_depth = 32, m_min_leaf = 1;
};

// A copy of a built scene's accelerator (the bvh_mem DeviceMemory the
// reference's threaded renderer copies from the scene and traces,
// src/rt/scene.cpp:99-128, src/rt/tracer.cpp:90-135): single-root scenes.
class BVHMemory {
public:
    BVHMemory() = default;
    ~BVHMemory();
    BVHMemory(const BVHMemory&) = delete;
    BVHMemory& operator=(const BVHMemory&) = delete;
    // Copies the scene's current accelerator (a device copy on the context's
    // stream).  0 = ok, -1 = error.
    int32_t copy_from(Scene& scene);
    bool valid() const { return m_h != nullptr && m_of != nullptr; }
    rtx_bvh* handle() const { return m_h; }

private:
    rtx_bvh* m_h = nullptr;
    const Scene* m_of = nullptr;
};

class PrimaryRayGenerator {
public:
    int32_t initialize();
    int32_t set_spp(size_t spp, pixel_sample_cl const* pixel_samples);
    size_t get_spp() const { return m_samples.size() / 3; }
    int32_t generate(const Camera& cam, RayBundle& bundle, size_t size[2], const size_t ray_count,
                     const size_t offset);
    void timing(bool b) { m_timing = b; }
    double get_exec_time() const { return m_time_ms; }
    void update_configuration(const RendererConfig& conf);

private:
    std::vector<float> m_samples{0.f, 0.f, 1.f};
    int m_quad = 32;
    bool m_timing = false, m_initialized = false;
    double m_time_ms = 0;
};

class Tracer {
public:
    int32_t initialize();
    int32_t trace(Scene& scene, int32_t ray_count, RayBundle& rays, HitBundle& hits, bool secondary = false);
    // Against an accelerator copy of this scene (include/rt/tracer.hpp:21, 26).
    int32_t trace(Scene& scene, BVHMemory& bvh_mem, int32_t ray_count, RayBundle& rays, HitBundle& hits,
                  bool secondary = false);
    int32_t shadow_trace(Scene& scene, int32_t ray_count, RayBundle& rays, HitBundle& hits, bool secondary = false);
    int32_t shadow_trace(Scene& scene, BVHMemory& bvh_mem, int32_t ray_count, RayBundle& rays, HitBundle& hits,
                         bool secondary = false);
    void timing(bool b) { m_timing = b; }
    double get_trace_exec_time() const { return m_trace_ms; }
    double get_shadow_exec_time() const { return m_shadow_ms; }
    void update_configuration(const RendererConfig& conf) { (void)conf; }

private:
    bool m_timing = false, m_initialized = false;
    double m_trace_ms = 0, m_shadow_ms = 0;
};

class SecondaryRayGenerator {
public:
    int32_t initialize();
    int32_t generate(Scene& scene, RayBundle& ray_in, size_t rays_in, HitBundle& hits, RayBundle& ray_out,
                     size_t* rays_out);
    void set_max_rays(size_t max) { m_max_rays = max; }
    void timing(bool b) { m_timing = b; }
    double get_exec_time() const { return m_time_ms; }
    void update_configuration(const RendererConfig& conf) { (void)conf; }

private:
    size_t m_max_rays = 0;
    bool m_timing = false, m_initialized = false;
    double m_time_ms = 0;
};

class RayShader {
public:
    int32_t initialize();
    int32_t shade(RayBundle& rays, HitBundle& hb, Scene& scene, Cubemap& cm, FrameBuffer& fb, size_t size,
                  bool primary = false);
    void timing(bool b) { m_timing = b; }
    double get_exec_time() const { return m_time_ms; }
    void update_configuration(const RendererConfig& conf) { (void)conf; }

private:
    bool m_timing = false, m_initialized = false;
    double m_time_ms = 0;
};

// Renderer (src/rt/renderer.cpp): one frame = set_up_frame ->
// render_to_framebuffer (tiles of rays: primary, trace, shadow, shade, then
// max_bounces rounds of secondary generation with the count read back) ->
// copy_framebuffer -> conclude_frame, driven through the stage classes.
class Renderer {
public:
    Renderer();
    uint32_t initialize();
    uint32_t configure_from_defaults();
    uint32_t resize_output(size_t width, size_t height);
    uint32_t set_up_frame(Scene& scene);
    uint32_t update_accelerator(Scene& scene);
    uint32_t update_configuration();
    uint32_t render_to_framebuffer(Scene& scene);
    uint32_t copy_framebuffer();
    uint32_t conclude_frame(Scene& scene);
    uint32_t get_framebuffer_w() const { return (uint32_t)fb_w; }
    uint32_t get_framebuffer_h() const { return (uint32_t)fb_h; }
    int32_t set_samples_per_pixel(size_t spp, pixel_sample_cl const* pixel_samples);
    size_t get_samples_per_pixel() const { return prim_ray_gen.get_spp(); }
    void set_static_bvh(bool b) { static_bvh = b; }
    bool is_static_bvh() const { return static_bvh; }
    void set_max_bounces(uint32_t b) { max_bounces = b; }
    uint32_t get_max_bounces() const { return max_bounces; }
    void clear_stats() {
        stats.clear_times();
        stats.clear_mean_times();
    }
    const FrameStats& get_frame_stats() const { return stats; }
    FrameBuffer& framebuffer() { return m_framebuffer; }
    RendererConfig config;

private:
    bool initialized = false;
    RayBundle ray_bundle_1, ray_bundle_2;
    HitBundle hit_bundle;
    FrameBuffer m_framebuffer;
    BVHBuilder bvh_builder;
    PrimaryRayGenerator prim_ray_gen;
    SecondaryRayGenerator sec_ray_gen;
    Tracer tracer;
    RayShader ray_shader;
    FrameStats stats;
    size_t fb_w = 512, fb_h = 512;
    size_t tile_size = 0;
    uint32_t max_bounces = 9;
    bool static_bvh = false;
    std::chrono::steady_clock::time_point frame_start;
};

#endif  // RT_STAGES_HPP
