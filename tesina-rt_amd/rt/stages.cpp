// C++ stage classes (include/rt/stages.hpp) over the C ABI.  Each method
// cites the reference method it stands in for; results go through the same
// gfx950 kernels as the C ABI, so these classes add no device code.
#include "rt/stages.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <map>
#include <sstream>

namespace {
using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); }
rtx_ctx* dev() { return DeviceInterface::instance()->ctx(); }
void report(const char* what) { std::cerr << what << ": " << rtx_last_error() << "\n"; }

// Per-kernel totals of the context's profile (launches, ms); waits for the
// device.
using ProfileTotals = std::map<std::string, std::pair<unsigned long long, double>>;
ProfileTotals profile_totals(rtx_ctx* ctx) {
    ProfileTotals t;
    const int64_t n = rtx_profile_report(ctx, nullptr, 0);
    if (n <= 1) return t;
    std::string buf((size_t)n, '\0');
    rtx_profile_report(ctx, &buf[0], (uint64_t)n);
    std::istringstream in(buf.c_str());
    std::string name;
    unsigned long long c = 0;
    double ms = 0;
    while (in >> name >> c >> ms) t[name] = {c, ms};
    return t;
}
}  // namespace

// ------------------------------------------------------ DeviceInterface --
DeviceInterface* DeviceInterface::instance() {
    static DeviceInterface di;
    return &di;
}

int32_t DeviceInterface::initialize(int device) {
    if (m_ctx) return 0;
    if (rtx_init(device, &m_ctx) != RTX_OK) {
        report("DeviceInterface::initialize");
        m_ctx = nullptr;
        return -1;
    }
    return 0;
}

void DeviceInterface::finish_commands() {
    if (m_ctx) rtx_sync(m_ctx, nullptr);
}

int32_t DeviceInterface::shutdown() {
    if (!m_ctx) return 0;
    int32_t e = rtx_shutdown(m_ctx);
    m_ctx = nullptr;
    return e ? -1 : 0;
}

// --------------------------------------------------------------- Camera --
Camera::Camera() : pos{0, 0, 0}, dir{0, 0, 1}, right{1, 0, 0}, up{0, 1, 0}, fov(0.785398f), aspect(1.f),
                   right_handed(false) {}

Camera::Camera(const vec3& position, const vec3& direction, const vec3& u, float f, float a, bool rh) : Camera() {
    fov = f;
    aspect = a;
    set(position, direction, u, f, a, rh);
}

// Camera::set (src/rt/camera.cpp:31-59).
void Camera::set(const vec3& position, const vec3& direction, const vec3& u, float f, float a, bool rh) {
    if (f > 0.f) fov = f;
    if (a > 0.f) aspect = a;
    right_handed = rh;
    rtx_camera c{};
    const float p[3] = {position.x, position.y, position.z};
    const float d[3] = {direction.x, direction.y, direction.z};
    const float uu[3] = {u.x, u.y, u.z};
    rtx_camera_set(&c, p, d, uu, f, aspect, rh ? 1 : 0);
    pos = vec3{c.pos[0], c.pos[1], c.pos[2]};
    dir = vec3{c.dir[0], c.dir[1], c.dir[2]};
    right = vec3{c.right[0], c.right[1], c.right[2]};
    up = vec3{c.up[0], c.up[1], c.up[2]};
}

rtx_camera Camera::basis() const {
    rtx_camera c{};
    const vec3* v[4] = {&pos, &dir, &right, &up};
    float* o[4] = {c.pos, c.dir, c.right, c.up};
    for (int k = 0; k < 4; ++k) {
        o[k][0] = v[k]->x;
        o[k][1] = v[k]->y;
        o[k][2] = v[k]->z;
    }
    return c;
}

// ------------------------------------------------------- RendererConfig --
// Defaults of src/rt/renderer-config.cpp:4-16.
RendererConfig::RendererConfig()
    : use_lbvh(1), bvh_refit_only(0), bvh_depth(32), bvh_min_leaf_size(1), sec_ray_use_atomics(0),
      sec_ray_use_disc(1), prim_ray_quad_size(32), prim_ray_use_zcurve(0), tile_to_cores_ratio(128) {}

// ---------------------------------------------------------------- Scene --
Scene::Scene() { m_lights.count = 0; }
Scene::~Scene() { destroy(); }

int32_t Scene::initialize() {
    if (m_h) return 0;
    if (!DeviceInterface::instance()->good() && DeviceInterface::instance()->initialize()) return -1;
    if (rtx_scene_create(dev(), &m_h) != RTX_OK) {
        report("Scene::initialize");
        return -1;
    }
    return 0;
}

int32_t Scene::destroy() {
    if (m_h) rtx_scene_destroy(m_h);
    m_h = nullptr;
    if (m_inst) rtx_instances_destroy(m_inst);
    m_inst = nullptr;
    m_img_target = nullptr;
    m_cube_on = nullptr;
    for (rtx_scene* m : m_meshes) rtx_scene_destroy(m);
    m_meshes.clear();
    m_host.clear();
    m_obj_mesh.clear();
    m_obj_mat.clear();
    m_tr.clear();
    m_tr_inv.clear();
    m_bvhs_built = false;
    return 0;
}

// ------------------------------------------------------- images --
int32_t TextureAtlas::destroy() {
    m_rgba.clear();
    m_dims.clear();
    m_dirty = true;
    return 0;
}

texture_id TextureAtlas::load_texture_rgba(const uint8_t* rgba, uint32_t width, uint32_t height) {
    if (!rgba || !width || !height || texture_count() >= RTX_MAX_TEXTURES) return invalid_tex_id;
    m_rgba.insert(m_rgba.end(), rgba, rgba + 4ull * width * height);
    m_dims.push_back((int32_t)width);
    m_dims.push_back((int32_t)height);
    m_dirty = true;
    return (texture_id)texture_count() - 1;
}

int32_t Cubemap::initialize_rgba(const uint8_t* const faces[6], const uint32_t dims[12]) {
    std::vector<uint8_t> data;
    for (int f = 0; f < 6; ++f) {
        if (!faces[f] || !dims[2 * f] || !dims[2 * f + 1]) return -1;
        data.insert(data.end(), faces[f], faces[f] + 4ull * dims[2 * f] * dims[2 * f + 1]);
    }
    m_rgba.swap(data);
    m_dims.assign(dims, dims + 12);
    ++m_version;
    enabled = true;
    return 0;
}

int32_t Cubemap::destroy() {
    m_rgba.clear();
    m_dims.clear();
    ++m_version;
    enabled = false;
    return 0;
}

rtx_scene* Scene::shading_handle() {
    if (m_built || !m_bvhs_built || !m_inst) return m_h;
    rtx_scene* sh = nullptr;
    if (rtx_instances_shading_scene(m_inst, &sh) != RTX_OK) {
        report("Scene::shading_handle");
        return nullptr;
    }
    return sh;
}

int32_t Scene::sync_images(const Cubemap& cm) {
    rtx_scene* target = shading_handle();
    if (!target) return -1;
    const bool moved = target != m_img_target;  // a new target starts without images
    m_img_target = nullptr;
    if (texture_atlas.m_dirty || moved) {
        const int32_t n = (int32_t)texture_atlas.texture_count();
        if (rtx_scene_set_textures(target, n ? texture_atlas.m_rgba.data() : nullptr,
                                   n ? texture_atlas.m_dims.data() : nullptr, n) != RTX_OK) {
            report("Scene::sync_images");
            return -1;
        }
        texture_atlas.m_dirty = false;
    }
    const Cubemap* want = (cm.enabled && cm.m_dims.size() == 12) ? &cm : nullptr;
    if (moved || want != m_cube_on || (want && want->m_version != m_cube_version)) {
        if (rtx_scene_set_cubemap(target, want ? want->m_rgba.data() : nullptr,
                                  want ? want->m_dims.data() : nullptr) != RTX_OK) {
            report("Scene::sync_images");
            return -1;
        }
        m_cube_on = want;
        m_cube_version = want ? want->m_version : 0;
    }
    m_img_target = target;
    return 0;
}

// ------------------------------------------------------------- objects --
int32_t Scene::add_mesh(const float* pos, const float* nrm, const float* uv, uint32_t V, const uint32_t* idx,
                        uint32_t T) {
    if (!DeviceInterface::instance()->good() && DeviceInterface::instance()->initialize()) return -1;
    rtx_scene* m = nullptr;
    if (rtx_scene_create(dev(), &m) != RTX_OK) {
        report("Scene::add_mesh");
        return -1;
    }
    std::vector<int32_t> mm(T, 0);
    if (rtx_scene_set_vertices(m, pos, nrm, uv, V) != RTX_OK ||
        rtx_scene_set_triangles(m, idx, mm.data(), T) != RTX_OK) {
        report("Scene::add_mesh");
        rtx_scene_destroy(m);
        return -1;
    }
    m_meshes.push_back(m);
    HostMesh h;
    h.pos.assign(pos, pos + 3 * (size_t)V);
    h.nrm = nrm ? std::vector<float>(nrm, nrm + 3 * (size_t)V) : std::vector<float>(3 * (size_t)V, 0.f);
    h.uv = uv ? std::vector<float>(uv, uv + 2 * (size_t)V) : std::vector<float>(2 * (size_t)V, 0.f);
    h.idx.assign(idx, idx + 3 * (size_t)T);
    m_host.push_back(std::move(h));
    return (int32_t)m_meshes.size() - 1;
}

int32_t Scene::load_obj_file_as_aggregate(const std::string& filename) {
    rtx_obj* o = nullptr;
    if (rtx_obj_load(filename.c_str(), 0, &o) != RTX_OK) {
        report("Scene::load_obj_file_as_aggregate");
        return -1;
    }
    uint32_t V = 0, T = 0;
    rtx_obj_info(o, &V, &T, nullptr);
    std::vector<float> pos(3 * (size_t)V), nrm(3 * (size_t)V), uv(2 * (size_t)V);
    std::vector<uint32_t> idx(3 * (size_t)T);
    rtx_obj_read(o, pos.data(), nrm.data(), uv.data(), idx.data());
    rtx_obj_destroy(o);
    return add_mesh(pos.data(), nrm.data(), uv.data(), V, idx.data(), T);
}

std::vector<int32_t> Scene::load_obj_file(const std::string& filename) {
    std::vector<int32_t> ids;
    rtx_obj* o = nullptr;
    if (rtx_obj_load(filename.c_str(), 0, &o) != RTX_OK) {
        report("Scene::load_obj_file");
        return ids;
    }
    uint32_t G = 0;
    rtx_obj_info(o, nullptr, nullptr, &G);
    for (uint32_t g = 0; g < G; ++g) {
        rtx_obj_group gi;
        uint32_t nv = 0;
        rtx_obj_group_info(o, g, &gi);
        rtx_obj_read_group(o, g, &nv, nullptr, nullptr, nullptr, nullptr);
        std::vector<float> pos(3 * (size_t)nv), nrm(3 * (size_t)nv), uv(2 * (size_t)nv);
        std::vector<uint32_t> idx(3 * (size_t)gi.triangle_count);
        rtx_obj_read_group(o, g, &nv, pos.data(), nrm.data(), uv.data(), idx.data());
        const int32_t id = add_mesh(pos.data(), nrm.data(), uv.data(), nv, idx.data(), gi.triangle_count);
        if (id < 0) break;
        ids.push_back(id);
    }
    rtx_obj_destroy(o);
    return ids;
}

int32_t Scene::set_object_material(int32_t object, const rtx_material& mat) {
    if (object < 0 || (size_t)object >= m_obj_mesh.size()) return -1;
    m_obj_mat[(size_t)object] = mat;
    return 0;
}

// M (v, w) with each row summed in index order, then homogenized by 1/w
// (GeometricProperties::transform, src/rt/geom.cpp:45-64; the normal's w is
// set to 1 before).
static void transform_h(const float* M, const float* v, float w, bool normal, float* out) {
    float r[4];
    for (int i = 0; i < 4; ++i) {
        float acc = 0.f;
        acc += M[4 * i] * v[0];
        acc += M[4 * i + 1] * v[1];
        acc += M[4 * i + 2] * v[2];
        acc += M[4 * i + 3] * w;
        r[i] = acc;
    }
    if (normal) r[3] = 1.f;
    const float inv = 1.f / r[3];
    out[0] = r[0] * inv;
    out[1] = r[1] * inv;
    out[2] = r[2] * inv;
}

int32_t Scene::create_aggregate_mesh() {
    if (!m_h || m_obj_mesh.empty()) return -1;
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> idx;
    std::vector<int32_t> matmap;
    for (size_t o = 0; o < m_obj_mesh.size(); ++o) {
        const HostMesh& h = m_host[(size_t)m_obj_mesh[o]];
        const float* M = &m_tr[16 * o];
        const uint32_t base = (uint32_t)(pos.size() / 3);
        const size_t V = h.pos.size() / 3;
        for (size_t v = 0; v < V; ++v) {
            float p[3], n[3];
            transform_h(M, &h.pos[3 * v], 1.f, false, p);
            transform_h(M, &h.nrm[3 * v], 0.f, true, n);
            pos.insert(pos.end(), p, p + 3);
            nrm.insert(nrm.end(), n, n + 3);
        }
        uv.insert(uv.end(), h.uv.begin(), h.uv.end());
        for (uint32_t i : h.idx) idx.push_back(i + base);
        matmap.insert(matmap.end(), h.idx.size() / 3, (int32_t)o);
    }
    if (set_geometry(pos.data(), nrm.data(), uv.data(), (uint32_t)(pos.size() / 3), idx.data(), matmap.data(),
                     (uint32_t)(idx.size() / 3)) ||
        set_materials(m_obj_mat))
        return -1;
    return 0;
}

// Scene::add_object: identity transform until set_object_transform.
int32_t Scene::add_object(int32_t mesh) {
    if (mesh < 0 || (size_t)mesh >= m_meshes.size()) return -1;
    static const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    m_obj_mesh.push_back(mesh);
    rtx_material black{};  // material_cl(): black, texture -1
    black.texture = -1;
    m_obj_mat.push_back(black);
    m_tr.insert(m_tr.end(), I, I + 16);
    m_tr_inv.insert(m_tr_inv.end(), I, I + 16);
    return (int32_t)m_obj_mesh.size() - 1;
}

int32_t Scene::set_object_transform(int32_t object, const float* tr16, const float* tr_inv16) {
    if (object < 0 || (size_t)object >= m_obj_mesh.size() || !tr16 || !tr_inv16) return -1;
    std::copy(tr16, tr16 + 16, m_tr.begin() + 16 * (size_t)object);
    std::copy(tr_inv16, tr_inv16 + 16, m_tr_inv.begin() + 16 * (size_t)object);
    return 0;
}

// Scene::create_bvhs (src/rt/scene.cpp:639-700): one tree per mesh that an
// object uses, then the roots.
int32_t Scene::create_bvhs(int depth, int min_leaf) {
    if (m_obj_mesh.empty()) return -1;
    std::vector<bool> used(m_meshes.size(), false);
    rtx_lbvh_params p{std::min(std::max(depth, 1), 32), std::min(std::max(min_leaf, 1), 256)};
    for (int32_t m : m_obj_mesh) {
        if (used[(size_t)m]) continue;
        used[(size_t)m] = true;
        if (rtx_lbvh_build(m_meshes[(size_t)m], &p, nullptr) != RTX_OK) {
            report("Scene::create_bvhs");
            return -1;
        }
    }
    if (m_inst) rtx_instances_destroy(m_inst);
    m_inst = nullptr;
    m_img_target = nullptr;
    std::vector<rtx_scene*> sc;
    for (int32_t m : m_obj_mesh) sc.push_back(m_meshes[(size_t)m]);
    if (rtx_instances_create(dev(), sc.data(), m_tr.data(), m_tr_inv.data(), (int32_t)sc.size(), &m_inst) !=
            RTX_OK ||
        rtx_instances_set_lights(m_inst, &m_lights) != RTX_OK ||
        rtx_instances_set_materials(m_inst, m_obj_mat.data()) != RTX_OK) {
        report("Scene::create_bvhs");
        return -1;
    }
    m_T = 0;
    for (size_t m = 0; m < m_meshes.size(); ++m)
        if (used[m]) m_T += rtx_scene_triangle_count(m_meshes[m]);
    m_bvhs_built = true;
    return 0;
}

// Scene::update_bvh_roots (src/rt/scene.cpp:596-629): the objects' current
// transforms go to the roots.
int32_t Scene::update_bvh_roots() {
    if (!m_bvhs_built) return -1;
    return rtx_instances_set_transforms(m_inst, m_tr.data(), m_tr_inv.data()) == RTX_OK ? 0 : -1;
}

int32_t Scene::set_geometry(const float* pos, const float* nrm, const float* uv, uint32_t V, const uint32_t* idx,
                            const int32_t* matmap, uint32_t T) {
    if (!m_h || rtx_scene_set_vertices(m_h, pos, nrm, uv, V) != RTX_OK) return -1;
    m_V = V;
    return set_triangles(idx, matmap, T);
}

int32_t Scene::set_vertices_ref(const void* vertex80, uint32_t V) {
    if (!m_h || rtx_scene_set_vertices_ref(m_h, vertex80, V) != RTX_OK) return -1;
    m_V = V;
    m_built = false;
    return 0;
}

int32_t Scene::set_triangles(const uint32_t* idx, const int32_t* matmap, uint32_t T) {
    if (!m_h || rtx_scene_set_triangles(m_h, idx, matmap, T) != RTX_OK) return -1;
    m_T = T;
    m_built = false;
    return 0;
}

int32_t Scene::update_vertex_positions(const float* pos) {
    if (!m_h || rtx_scene_update_positions(m_h, pos, nullptr) != RTX_OK) return -1;
    DeviceInterface::instance()->finish_commands();  // the host array may be released on return
    return 0;
}

int32_t Scene::set_materials(const std::vector<rtx_material>& mats) {
    if (!m_h || mats.empty()) return -1;
    return rtx_scene_set_materials(m_h, mats.data(), (uint32_t)mats.size()) == RTX_OK ? 0 : -1;
}

int32_t Scene::push_lights() {
    if (m_h && rtx_scene_set_lights(m_h, &m_lights) != RTX_OK) return -1;
    if (m_inst && rtx_instances_set_lights(m_inst, &m_lights) != RTX_OK) return -1;
    for (rtx_scene* m : m_meshes)
        if (rtx_scene_set_lights(m, &m_lights) != RTX_OK) return -1;
    return 0;
}

// Scene::set_dir_light normalizes the direction on the host (src/rt/scene.cpp:987-1003).
int32_t Scene::set_dir_light(const vec3& d, const vec3& color) {
    if ((!m_h && m_meshes.empty())) return -1;
    float n = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    if (!(n > 0.f)) return -1;
    rtx_light L{};
    L.type = RTX_LIGHT_DIR;
    L.dir[0] = d.x / n;
    L.dir[1] = d.y / n;
    L.dir[2] = d.z / n;
    L.color[0] = color.x;
    L.color[1] = color.y;
    L.color[2] = color.z;
    m_lights.l[0] = L;
    m_lights.count = 1;
    return push_lights();
}

int32_t Scene::set_spot_light(const vec3& p, const vec3& d, const vec3& color, float angle, float radius) {
    if ((!m_h && m_meshes.empty())) return -1;
    rtx_light L{};
    L.type = RTX_LIGHT_SPOT;
    L.pos[0] = p.x, L.pos[1] = p.y, L.pos[2] = p.z;
    L.dir[0] = d.x, L.dir[1] = d.y, L.dir[2] = d.z;
    L.color[0] = color.x, L.color[1] = color.y, L.color[2] = color.z;
    L.angle = angle;
    L.radius = radius;
    m_lights.l[0] = L;
    m_lights.count = 1;
    return push_lights();
}

int32_t Scene::set_ambient_light(const vec3& c) {
    if ((!m_h && m_meshes.empty())) return -1;
    m_lights.ambient[0] = c.x;
    m_lights.ambient[1] = c.y;
    m_lights.ambient[2] = c.z;
    return push_lights();
}

int32_t Scene::add_point_light(const vec3& p, const vec3& color) {
    if ((!m_h && m_meshes.empty()) || m_lights.count >= RTX_MAX_LIGHTS) return -1;
    rtx_light& L = m_lights.l[m_lights.count++];
    L = rtx_light{};
    L.type = RTX_LIGHT_POINT;
    L.pos[0] = p.x, L.pos[1] = p.y, L.pos[2] = p.z;
    L.color[0] = color.x, L.color[1] = color.y, L.color[2] = color.z;
    return push_lights();
}

int32_t Scene::set_lights(const rtx_lights& lights) {
    if ((!m_h && m_meshes.empty())) return -1;
    m_lights = lights;
    return push_lights();
}

// -------------------------------------------------------------- bundles --
RayBundle::~RayBundle() {
    if (m_h) rtx_rays_destroy(m_h);
}

// RayBundle::initialize (src/rt/ray.cpp:48-66): zero rays is an error.
int32_t RayBundle::initialize(size_t rays) {
    if (!rays || !DeviceInterface::instance()->good()) return -1;
    if (m_h) rtx_rays_destroy(m_h);
    m_h = nullptr;
    if (rtx_rays_create(dev(), rays, &m_h) != RTX_OK) {
        report("RayBundle::initialize");
        m_h = nullptr;
        return -1;
    }
    m_count = rays;
    return 0;
}

int32_t RayBundle::resize(size_t rays) { return initialize(rays); }

HitBundle::~HitBundle() {
    if (m_h) rtx_hits_destroy(m_h);
}

int32_t HitBundle::initialize(size_t sz) {
    if (!sz || !DeviceInterface::instance()->good()) return -1;
    if (m_h) rtx_hits_destroy(m_h);
    m_h = nullptr;
    if (rtx_hits_create(dev(), sz, &m_h) != RTX_OK) {
        report("HitBundle::initialize");
        m_h = nullptr;
        return -1;
    }
    m_count = sz;
    return 0;
}

int32_t HitBundle::resize(size_t sz) { return initialize(sz); }

// ----------------------------------------------------------- FrameBuffer --
FrameBuffer::~FrameBuffer() {
    if (m_h) rtx_fb_destroy(m_h);
}

int32_t FrameBuffer::initialize(size_t sz[2]) {
    if (!DeviceInterface::instance()->good() || !sz[0] || !sz[1]) return -1;
    if (m_h) rtx_fb_destroy(m_h);
    m_h = nullptr;
    if (rtx_fb_create(dev(), (int32_t)sz[0], (int32_t)sz[1], &m_h) != RTX_OK) {
        report("FrameBuffer::initialize");
        m_h = nullptr;
        return -1;
    }
    m_size[0] = sz[0];
    m_size[1] = sz[1];
    return 0;
}

int32_t FrameBuffer::resize(size_t sz[2]) { return initialize(sz); }

// FrameBuffer::clear (src/rt/framebuffer.cpp:118-137).
int32_t FrameBuffer::clear() {
    if (!m_h) return -1;
    auto t = clk::now();
    if (rtx_fb_clear(m_h, nullptr) != RTX_OK) return -1;
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_clear_ms = ms_since(t);
    }
    return 0;
}

// FrameBuffer::copy (src/rt/framebuffer.cpp:141-165), into the device RGBA8 image.
int32_t FrameBuffer::copy() {
    if (!m_h) return -1;
    auto t = clk::now();
    if (rtx_fb_resolve(m_h, nullptr) != RTX_OK) return -1;
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_copy_ms = ms_since(t);
    }
    return 0;
}

int32_t FrameBuffer::read_rgba8(uint8_t* rgba) const {
    return m_h && rtx_fb_read_rgba8(m_h, rgba) == RTX_OK ? 0 : -1;
}

int32_t FrameBuffer::read_accum(uint32_t* rgb) const {
    return m_h && rtx_fb_read_accum(m_h, rgb) == RTX_OK ? 0 : -1;
}

// ------------------------------------------------------------ BVHBuilder --
int32_t BVHBuilder::initialize() {
    if (!DeviceInterface::instance()->good()) return -1;
    m_depth = 32;
    m_min_leaf = 1;
    return 0;
}

// BVHBuilder::update_configuration clamps (src/rt/bvh-builder.cpp:194-199);
// this build's codes are 32-bit, so depth is clamped to 32.
void BVHBuilder::update_configuration(const RendererConfig& conf) {
    m_depth = std::min(std::max(conf.bvh_depth, 1), 32);
    m_min_leaf = std::min(std::max(conf.bvh_min_leaf_size, 1), 256);
}

// BVHBuilder::build_lbvh (src/rt/bvh-builder.cpp:201-741): permutes the
// scene's index buffer and material map in place and records the node
// counts on the scene.
int32_t BVHBuilder::build_lbvh(Scene& scene, size_t cq_i) {
    (void)cq_i;
    if (!scene.valid()) return -1;
    auto t = clk::now();
    rtx_lbvh_params p{m_depth, m_min_leaf};
    if (log_stages(nullptr)) return -1;
    if (rtx_lbvh_build(scene.handle(), &p, nullptr) != RTX_OK) {
        report("BVHBuilder::build_lbvh");
        log_stages("-----BVH construction failed; kernels run: ");
        return -1;
    }
    if (log_stages("-----BVH construction times: ")) return -1;
    if (rtx_lbvh_info(scene.handle(), &scene.bvh_node_count, scene.bvh_level_node_count) != RTX_OK) return -1;
    if (m_timing) m_time_ms = ms_since(t);
    scene.m_built = true;
    return 0;
}

// With logging on and a log set: called with no title before the build
// (turns the context's kernel timing on and takes its totals) and with the
// title after it (waits for the device and writes each build kernel's time,
// the reference's stage lines first, then the total; puts the timing back).
int32_t BVHBuilder::log_stages(const char* title) {
    static thread_local ProfileTotals before;
    static thread_local int was_on = 0;
    if (!m_logging || !m_log) return 0;
    rtx_ctx* ctx = DeviceInterface::instance()->ctx();
    if (!title) {
        was_on = rtx_profile_enabled(ctx);
        if (rtx_profile_enable(ctx, 1) != RTX_OK) return -1;
        before = profile_totals(ctx);
        return 0;
    }
    const ProfileTotals after = profile_totals(ctx);
    if (!was_on) rtx_profile_enable(ctx, 0);
    // the reference's stage names where a kernel is that stage
    static const std::pair<const char*, const char*> named[] = {
        {"lbvh_prim_bounds", "BBox builder time"}, {"lbvh_morton", "Morton encoder time"},
        {"lbvh_sort_pass", "Morton sorter"}, {"lbvh_refit_level", "Node bbox compute time"}};
    Log& log = *m_log;
    log << title << "\n";
    double total = 0;
    std::map<std::string, double> ms;
    for (const auto& kv : after) {
        if (kv.first.compare(0, 5, "lbvh_") != 0) continue;
        auto b = before.find(kv.first);
        const double d = kv.second.second - (b == before.end() ? 0.0 : b->second.second);
        if (kv.second.first == (b == before.end() ? 0ull : b->second.first)) continue;  // not launched
        ms[kv.first] = d;
        total += d;
    }
    for (const auto& n : named) {
        auto it = ms.find(n.first);
        if (it == ms.end()) continue;
        log << "\t" << n.second << ": " << it->second << " ms\n";
        ms.erase(it);
    }
    for (const auto& kv : ms) log << "\t" << kv.first << ": " << kv.second << " ms\n";
    log << "\tTotal: " << total << " ms\n";
    return 0;
}

// BVHBuilder::refit_lbvh (src/rt/bvh-builder.cpp:743-850).
int32_t BVHBuilder::refit_lbvh(Scene& scene, size_t cq_i) {
    (void)cq_i;
    if (!scene.valid() || !scene.ready()) return -1;
    auto t = clk::now();
    if (log_stages(nullptr)) return -1;
    if (rtx_lbvh_refit(scene.handle(), nullptr) != RTX_OK) {
        report("BVHBuilder::refit_lbvh");
        log_stages("-----BVH refit failed; kernels run: ");
        return -1;
    }
    if (log_stages("-----BVH refit times: ")) return -1;
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_time_ms = ms_since(t);
    }
    return 0;
}

// --------------------------------------------------- PrimaryRayGenerator --
int32_t PrimaryRayGenerator::initialize() {
    if (m_initialized || !DeviceInterface::instance()->good()) return -1;
    m_samples = {0.f, 0.f, 1.f};
    m_quad = 32;
    m_initialized = true;
    return 0;
}

int32_t PrimaryRayGenerator::set_spp(size_t spp, pixel_sample_cl const* ps) {
    if (!spp || !ps) return -1;
    m_samples.resize(3 * spp);
    for (size_t i = 0; i < spp; ++i) {
        m_samples[3 * i] = ps[i].ox;
        m_samples[3 * i + 1] = ps[i].oy;
        m_samples[3 * i + 2] = ps[i].contribution;
    }
    return 0;
}

void PrimaryRayGenerator::update_configuration(const RendererConfig& conf) {
    m_quad = conf.prim_ray_quad_size > 0 ? conf.prim_ray_quad_size : 32;
}

// PrimaryRayGenerator::generate (src/rt/primary-ray-generator.cpp:58-165).
int32_t PrimaryRayGenerator::generate(const Camera& cam, RayBundle& bundle, size_t size[2], const size_t ray_count,
                                      const size_t offset) {
    if (!m_initialized || !bundle.valid()) return -1;
    auto t = clk::now();
    rtx_camera c = cam.basis();
    const int spp = (int)get_spp();
    if (rtx_raygen_primary(bundle.handle(), &c, (int32_t)size[0], (int32_t)size[1], m_quad, spp, m_samples.data(),
                           offset, ray_count, nullptr) != RTX_OK) {
        report("PrimaryRayGenerator::generate");
        return -1;
    }
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_time_ms = ms_since(t);
    }
    return 0;
}

// ---------------------------------------------------------------- Tracer --
int32_t Tracer::initialize() {
    if (!DeviceInterface::instance()->good()) return -1;
    m_initialized = true;
    return 0;
}

// The scene the single-root kernels trace: the aggregate mesh, or the mesh
// of a scene's only object (node 0 of the shared array, no transform).
static rtx_scene* single_root(const Scene& scene) {
    return scene.aggregate_built() ? scene.handle() : scene.mesh_handle(scene.object_mesh(0));
}

// ------------------------------------------------------------ BVHMemory --
BVHMemory::~BVHMemory() {
    if (m_h) rtx_bvh_destroy(m_h);
}

int32_t BVHMemory::copy_from(Scene& scene) {
    if (!scene.ready() || scene.root_count() > 1) return -1;  // single-root scenes
    if (!m_h && rtx_bvh_create(DeviceInterface::instance()->ctx(), &m_h) != RTX_OK) {
        report("BVHMemory::copy_from");
        return -1;
    }
    m_of = nullptr;
    if (rtx_bvh_copy(m_h, single_root(scene), nullptr) != RTX_OK) {
        report("BVHMemory::copy_from");
        return -1;
    }
    m_of = &scene;
    return 0;
}

// Tracer::trace (src/rt/tracer.cpp:140-269).
int32_t Tracer::trace(Scene& scene, int32_t ray_count, RayBundle& rays, HitBundle& hits, bool secondary) {
    (void)secondary;
    if (!m_initialized || ray_count < 0 || !scene.ready()) return -1;
    auto t = clk::now();
    const int32_t rc = scene.root_count() > 1
                           ? rtx_trace_instances(scene.instances(), rays.handle(), hits.handle(), (uint64_t)ray_count,
                                                 nullptr)
                           : rtx_trace(single_root(scene), rays.handle(), hits.handle(), (uint64_t)ray_count, nullptr);
    if (rc != RTX_OK) {
        report("Tracer::trace");
        return -1;
    }
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_trace_ms = ms_since(t);
    }
    return 0;
}

// Tracer::shadow_trace (src/rt/tracer.cpp:273-352).
int32_t Tracer::shadow_trace(Scene& scene, int32_t ray_count, RayBundle& rays, HitBundle& hits, bool secondary) {
    (void)secondary;
    if (!m_initialized || ray_count < 0 || !scene.ready()) return -1;
    auto t = clk::now();
    const int32_t rc = scene.root_count() > 1
                           ? rtx_trace_shadow_instances(scene.instances(), rays.handle(), hits.handle(),
                                                        (uint64_t)ray_count, nullptr)
                           : rtx_trace_shadow(single_root(scene), rays.handle(), hits.handle(), (uint64_t)ray_count,
                                              nullptr);
    if (rc != RTX_OK) {
        report("Tracer::shadow_trace");
        return -1;
    }
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_shadow_ms = ms_since(t);
    }
    return 0;
}

// Tracer::trace / shadow_trace against an accelerator copy
// (src/rt/tracer.cpp:90-135; the shadow overload alike).
int32_t Tracer::trace(Scene& scene, BVHMemory& bvh_mem, int32_t ray_count, RayBundle& rays, HitBundle& hits,
                      bool secondary) {
    (void)secondary;
    if (!m_initialized || ray_count < 0 || !scene.ready() || !bvh_mem.valid() || scene.root_count() > 1) return -1;
    auto t = clk::now();
    if (rtx_trace_with(single_root(scene), bvh_mem.handle(), rays.handle(), hits.handle(), (uint64_t)ray_count,
                       nullptr) != RTX_OK) {
        report("Tracer::trace");
        return -1;
    }
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_trace_ms = ms_since(t);
    }
    return 0;
}

int32_t Tracer::shadow_trace(Scene& scene, BVHMemory& bvh_mem, int32_t ray_count, RayBundle& rays, HitBundle& hits,
                             bool secondary) {
    (void)secondary;
    if (!m_initialized || ray_count < 0 || !scene.ready() || !bvh_mem.valid() || scene.root_count() > 1) return -1;
    auto t = clk::now();
    if (rtx_trace_shadow_with(single_root(scene), bvh_mem.handle(), rays.handle(), hits.handle(),
                              (uint64_t)ray_count, nullptr) != RTX_OK) {
        report("Tracer::shadow_trace");
        return -1;
    }
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_shadow_ms = ms_since(t);
    }
    return 0;
}

// ------------------------------------------------- SecondaryRayGenerator --
int32_t SecondaryRayGenerator::initialize() {
    if (!DeviceInterface::instance()->good()) return -1;
    m_initialized = true;
    return 0;
}

// SecondaryRayGenerator::generate, scan form (src/rt/secondary-ray-generator.cpp:57-145,
// 248-340): the new count is read back, as in the reference.
int32_t SecondaryRayGenerator::generate(Scene& scene, RayBundle& ray_in, size_t rays_in, HitBundle& hits,
                                        RayBundle& ray_out, size_t* rays_out) {
    if (!m_initialized || !rays_out) return -1;
    auto t = clk::now();
    uint64_t n = 0;
    rtx_scene* sh = scene.shading_handle();
    if (!sh || rtx_raygen_secondary(sh, ray_in.handle(), hits.handle(), ray_out.handle(), rays_in, &n, nullptr) !=
                   RTX_OK) {
        report("SecondaryRayGenerator::generate");
        return -1;
    }
    if (n > 2 * rays_in) {
        std::cerr << "Error: secondary ray generator reported wrong number of new rays.\n";
        return -1;
    }
    *rays_out = (size_t)n;
    if (m_timing) m_time_ms = ms_since(t);
    return 0;
}

// ------------------------------------------------------------- RayShader --
int32_t RayShader::initialize() {
    if (!DeviceInterface::instance()->good()) return -1;
    m_initialized = true;
    return 0;
}

// RayShader::shade (src/rt/ray-shader.cpp:33-154).
int32_t RayShader::shade(RayBundle& rays, HitBundle& hb, Scene& scene, Cubemap& cm, FrameBuffer& fb, size_t size,
                         bool primary) {
    if (!m_initialized || scene.sync_images(cm)) return -1;
    auto t = clk::now();
    if (rtx_shade(scene.shading_handle(), rays.handle(), hb.handle(), fb.handle(), size, primary ? 1 : 0, nullptr) !=
        RTX_OK) {
        report("RayShader::shade");
        return -1;
    }
    if (m_timing) {
        DeviceInterface::instance()->finish_commands();
        m_time_ms = ms_since(t);
    }
    return 0;
}

// -------------------------------------------------------------- Renderer --
Renderer::Renderer() {}

// Renderer::configure_from_defaults (src/rt/renderer.cpp:279-295).
uint32_t Renderer::configure_from_defaults() {
    fb_w = 512;
    fb_h = 512;
    static_bvh = false;
    max_bounces = 9;
    config.bvh_depth = 32;
    config.tile_to_cores_ratio = 128;
    config.use_lbvh = 1;
    config.bvh_refit_only = 0;
    config.sec_ray_use_disc = 0;
    config.prim_ray_quad_size = 32;
    config.prim_ray_use_zcurve = 0;
    return 0;
}

// Renderer::initialize (src/rt/renderer.cpp:350-453).
uint32_t Renderer::initialize() {
    if (initialized) return 0;
    if (!DeviceInterface::instance()->good() && DeviceInterface::instance()->initialize()) return (uint32_t)-1;
    size_t sz[2] = {fb_w, fb_h};
    if (m_framebuffer.initialize(sz) || bvh_builder.initialize() || prim_ray_gen.initialize() ||
        tracer.initialize() || sec_ray_gen.initialize() || ray_shader.initialize())
        return (uint32_t)-1;
    bvh_builder.timing(true);
    prim_ray_gen.timing(true);
    tracer.timing(true);
    sec_ray_gen.timing(true);
    ray_shader.timing(true);
    m_framebuffer.timing(true);
    tile_size = 0;
    initialized = true;
    return update_configuration();
}

uint32_t Renderer::resize_output(size_t width, size_t height) {
    fb_w = width;
    fb_h = height;
    size_t sz[2] = {fb_w, fb_h};
    if (initialized && m_framebuffer.resize(sz)) return (uint32_t)-1;
    return initialized ? update_configuration() : 0;
}

int32_t Renderer::set_samples_per_pixel(size_t spp, pixel_sample_cl const* ps) {
    return prim_ray_gen.set_spp(spp, ps);
}

// Renderer::update_configuration (src/rt/renderer.cpp:79-111): one tile of
// CUs * 1024 * tile_to_cores_ratio rays at most (the whole frame on MI355X),
// ray bundles of 3 tiles.
uint32_t Renderer::update_configuration() {
    const size_t old_tile = tile_size;
    const size_t pixels = fb_w * fb_h;
    size_t tile = (size_t)(256.0 * 1024.0 * config.tile_to_cores_ratio);
    tile_size = std::min(pixels * get_samples_per_pixel(), std::max<size_t>(tile, 1));
    if (tile_size != old_tile) {
        const size_t bundle = tile_size * 3;
        if (ray_bundle_1.resize(bundle) || ray_bundle_2.resize(bundle) || hit_bundle.resize(bundle)) {
            std::cerr << "Error resizing ray and hit bundles.\n";
            return (uint32_t)-1;
        }
        sec_ray_gen.set_max_rays(bundle);
    }
    bvh_builder.update_configuration(config);
    prim_ray_gen.update_configuration(config);
    sec_ray_gen.update_configuration(config);
    tracer.update_configuration(config);
    ray_shader.update_configuration(config);
    return 0;
}

// Renderer::set_up_frame (src/rt/renderer.cpp:12-51).
uint32_t Renderer::set_up_frame(Scene& scene) {
    frame_start = clk::now();
    if (!initialized) return (uint32_t)-1;
    stats.clear_times();
    stats.total_ray_count = 0;
    stats.total_sec_ray_count = 0;
    if (m_framebuffer.clear()) return (uint32_t)-1;
    stats.stage_times[FB_CLEAR] = m_framebuffer.get_clear_exec_time();
    return update_accelerator(scene);
}

// Renderer::update_accelerator (src/rt/renderer.cpp:53-77).
uint32_t Renderer::update_accelerator(Scene& scene) {
    if (config.use_lbvh && !(static_bvh && scene.ready())) {
        if (config.bvh_refit_only && scene.ready()) {
            if (bvh_builder.refit_lbvh(scene)) return (uint32_t)-1;
        } else if (bvh_builder.build_lbvh(scene)) {
            return (uint32_t)-1;
        }
    }
    stats.stage_times[BVH_BUILD] = bvh_builder.get_exec_time();
    return 0;
}

// Renderer::render_to_framebuffer (src/rt/renderer.cpp:113-214).
uint32_t Renderer::render_to_framebuffer(Scene& scene) {
    const size_t pixel_count = fb_w * fb_h;
    const size_t sample_count = pixel_count * prim_ray_gen.get_spp();
    size_t fb_size[2] = {fb_w, fb_h};
    for (size_t offset = 0; offset < sample_count; offset += tile_size) {
        RayBundle* ray_in = &ray_bundle_1;
        RayBundle* ray_out = &ray_bundle_2;
        const size_t cur = std::min(tile_size, sample_count - offset);
        if (prim_ray_gen.generate(scene.camera, *ray_in, fb_size, cur, offset)) return (uint32_t)-1;
        stats.stage_times[PRIM_RAY_GEN] += prim_ray_gen.get_exec_time();
        stats.total_ray_count += cur;
        if (tracer.trace(scene, (int32_t)cur, *ray_in, hit_bundle)) return (uint32_t)-1;
        stats.stage_times[PRIM_TRACE] += tracer.get_trace_exec_time();
        if (tracer.shadow_trace(scene, (int32_t)cur, *ray_in, hit_bundle)) return (uint32_t)-1;
        stats.stage_times[PRIM_SHADOW_TRACE] += tracer.get_shadow_exec_time();
        if (ray_shader.shade(*ray_in, hit_bundle, scene, scene.cubemap, m_framebuffer, cur, true)) return (uint32_t)-1;
        stats.stage_times[SHADE] += ray_shader.get_exec_time();
        size_t sec = cur;
        for (uint32_t bounce = 0; bounce < max_bounces; ++bounce) {
            size_t sec_in = sec;
            if (sec_ray_gen.generate(scene, *ray_in, sec_in, hit_bundle, *ray_out, &sec)) return (uint32_t)-1;
            stats.stage_times[SEC_RAY_GEN] += sec_ray_gen.get_exec_time();
            std::swap(ray_in, ray_out);
            if (!sec) break;
            stats.total_ray_count += sec;
            stats.total_sec_ray_count += sec;
            if (tracer.trace(scene, (int32_t)sec, *ray_in, hit_bundle, true)) return (uint32_t)-1;
            stats.stage_times[SEC_TRACE] += tracer.get_trace_exec_time();
            if (tracer.shadow_trace(scene, (int32_t)sec, *ray_in, hit_bundle, true)) return (uint32_t)-1;
            stats.stage_times[SEC_SHADOW_TRACE] += tracer.get_shadow_exec_time();
            if (ray_shader.shade(*ray_in, hit_bundle, scene, scene.cubemap, m_framebuffer, sec)) return (uint32_t)-1;
            stats.stage_times[SHADE] += ray_shader.get_exec_time();
        }
    }
    return 0;
}

// Renderer::copy_framebuffer (src/rt/renderer.cpp:216-227).
uint32_t Renderer::copy_framebuffer() {
    if (m_framebuffer.copy()) return (uint32_t)-1;
    stats.stage_times[FB_COPY] = m_framebuffer.get_copy_exec_time();
    return 0;
}

// Renderer::conclude_frame (src/rt/renderer.cpp:248-272).
uint32_t Renderer::conclude_frame(Scene& scene) {
    (void)scene;
    DeviceInterface::instance()->finish_commands();
    stats.frame_time = ms_since(frame_start);
    stats.frame_acc_time += stats.frame_time;
    for (int i = 0; i < STAGE_COUNT; ++i) stats.stage_acc_times[i] += stats.stage_times[i];
    stats.acc_frames++;
    return 0;
}
