// Traces a scene of objects through the C++ stage classes (include/rt/stages.hpp)
// as the reference drives a multi-object scene: add objects, create_bvhs,
// update_bvh_roots, then Tracer::trace and Tracer::shadow_trace, which take
// the multi-root kernels when root_count() > 1 (src/rt/tracer.cpp:97, 223).
//
//   stage_objects <scene_dir> <out_dir>
// scene_dir: meshes.txt  "M", then per mesh "V T"
//            mesh<k>.pos.f32 mesh<k>.nrm.f32 mesh<k>.uv.f32 mesh<k>.idx.u32
//            objects.txt "N", then per object "mesh tr[16] tr_inv[16]"
//            moved.txt   (optional) N lines "tr[16] tr_inv[16]": second pass
//                        after Object moves + update_bvh_roots
//            rays.f32    ori4 | dir4 interleaved per ray (8 floats), lights.bin
//            objmats.bin (optional) one rtx_material per object (Object::mat)
//            render.txt  (optional) "width height max_bounces" and the view
//                        "px py pz lx ly lz ux uy uz fov": after each pass, a
//                        frame through Renderer with use_lbvh = 0, the way the
//                        reference renders a scene of objects; images.txt
//                        (stage_io.hpp) gives it textures and a cubemap
// out_dir:   pass<p>.flags.u32 .t.f32 .id.i32 .bary.f32 .n4.f32 .p4.f32 .uv2.f32,
//            frame<p>.accum.u32 frame<p>.stats.txt "rays secondary_rays"
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "stage_io.hpp"

static bool read_mats(std::istream& in, float* tr, float* tri) {
    for (int k = 0; k < 16; ++k) in >> tr[k];
    for (int k = 0; k < 16; ++k) in >> tri[k];
    return (bool)in;
}

static int dump(HitBundle& hits, size_t n, const std::string& prefix) {
    std::vector<uint32_t> flags(n);
    std::vector<float> t(n), bary(2 * n), n4(4 * n), p4(4 * n), uv2(2 * n);
    std::vector<int32_t> id(n);
    if (rtx_hits_read(hits.handle(), flags.data(), t.data(), id.data(), bary.data(), n4.data(), p4.data(),
                      uv2.data(), n) != RTX_OK)
        return 1;
    save(prefix + ".flags.u32", flags);
    save(prefix + ".t.f32", t);
    save(prefix + ".id.i32", id);
    save(prefix + ".bary.f32", bary);
    save(prefix + ".n4.f32", n4);
    save(prefix + ".p4.f32", p4);
    save(prefix + ".uv2.f32", uv2);
    return 0;
}

struct FrameDriver {
    int w = 0, h = 0, bounces = 0;
    float view[10] = {};
    Renderer r;

    int open(const std::string& in, Scene& scene) {
        std::ifstream f(in + "/render.txt");
        f >> w >> h >> bounces;
        for (float& x : view) f >> x;
        if (!f || w <= 0 || h <= 0) return 2;
        if (const int e = load_images(in, scene)) return e;
        vec3 p{view[0], view[1], view[2]}, look{view[3], view[4], view[5]}, up{view[6], view[7], view[8]};
        scene.camera.set(p, vec3{look.x - p.x, look.y - p.y, look.z - p.z}, up, view[9], (float)w / (float)h);
        r.configure_from_defaults();
        r.config.use_lbvh = 0;  // the objects' own trees
        r.resize_output((size_t)w, (size_t)h);
        r.set_max_bounces((uint32_t)bounces);
        return r.initialize() ? 1 : 0;
    }

    int frame(Scene& scene, const std::string& prefix) {
        if (r.set_up_frame(scene) || r.render_to_framebuffer(scene) || r.copy_framebuffer() ||
            r.conclude_frame(scene))
            return 1;
        std::vector<uint32_t> acc((size_t)w * h * 3);
        if (r.framebuffer().read_accum(acc.data())) return 1;
        save(prefix + ".accum.u32", acc);
        const FrameStats& st = r.get_frame_stats();
        std::ofstream s(prefix + ".stats.txt");
        s << st.get_ray_count() << " " << st.get_secondary_ray_count() << "\n";
        return 0;
    }
};

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: stage_objects <scene_dir> <out_dir>\n";
        return 2;
    }
    const std::string in = argv[1], out = argv[2];
    auto rays = load<float>(in + "/rays.f32");
    auto lights = load<rtx_lights>(in + "/lights.bin");
    if (rays.empty() || lights.size() != 1) {
        std::cerr << "bad scene directory\n";
        return 2;
    }
    const size_t n = rays.size() / 8;
    std::vector<float> ori4(4 * n), dir4(4 * n);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) {
            ori4[4 * i + k] = rays[8 * i + k];
            dir4[4 * i + k] = rays[8 * i + 4 + k];
        }
    if (DeviceInterface::instance()->initialize(0)) return 1;
    {
        Scene scene;
        if (scene.initialize()) return 1;
        std::ifstream mf(in + "/meshes.txt");
        int M = 0;
        mf >> M;
        for (int k = 0; k < M; ++k) {
            uint32_t V = 0, T = 0;
            mf >> V >> T;
            const std::string b = in + "/mesh" + std::to_string(k);
            auto pos = load<float>(b + ".pos.f32");
            auto nrm = load<float>(b + ".nrm.f32");
            auto uv = load<float>(b + ".uv.f32");
            auto idx = load<uint32_t>(b + ".idx.u32");
            if (pos.size() != 3 * (size_t)V || idx.size() != 3 * (size_t)T) return 2;
            if (scene.add_mesh(pos.data(), nrm.data(), uv.data(), V, idx.data(), T) != k) return 1;
        }
        std::ifstream of(in + "/objects.txt");
        int N = 0;
        of >> N;
        for (int i = 0; i < N; ++i) {
            int mesh = -1;
            float tr[16], tri[16];
            of >> mesh;
            if (!read_mats(of, tr, tri)) return 2;
            const int32_t o = scene.add_object(mesh);
            if (o != i || scene.set_object_transform(o, tr, tri)) return 1;
        }
        auto omats = load<rtx_material>(in + "/objmats.bin");
        if (!omats.empty()) {
            if (omats.size() != (size_t)N) return 2;
            for (int i = 0; i < N; ++i)
                if (scene.set_object_material(i, omats[(size_t)i])) return 1;
        }
        if (scene.set_lights(lights[0]) || scene.create_bvhs(30, 1)) return 1;
        std::unique_ptr<FrameDriver> fd;
        if (std::ifstream(in + "/render.txt")) {
            fd.reset(new FrameDriver);
            if (const int e = fd->open(in, scene)) return e;
        }
        RayBundle rb;
        HitBundle hb;
        Tracer tracer;
        if (rb.initialize(n) || hb.initialize(n) || tracer.initialize()) return 1;
        if (rtx_rays_write(rb.handle(), ori4.data(), dir4.data(), nullptr, nullptr, n) != RTX_OK) return 1;
        if (tracer.trace(scene, (int32_t)n, rb, hb) || tracer.shadow_trace(scene, (int32_t)n, rb, hb)) return 1;
        if (dump(hb, n, out + "/pass0")) return 1;
        if (fd && fd->frame(scene, out + "/frame0")) return 1;
        std::ifstream mv(in + "/moved.txt");
        int passes = 1;
        if (mv) {
            for (int i = 0; i < N; ++i) {
                float tr[16], tri[16];
                if (!read_mats(mv, tr, tri) || scene.set_object_transform(i, tr, tri)) return 2;
            }
            if (scene.update_bvh_roots() || tracer.trace(scene, (int32_t)n, rb, hb) ||
                tracer.shadow_trace(scene, (int32_t)n, rb, hb) || dump(hb, n, out + "/pass1") ||
                (fd && fd->frame(scene, out + "/frame1")))
                return 1;
            passes = 2;
        }
        std::printf("stage_objects: %zu rays, %d objects, %zu roots, %d passes\n", n, N, scene.root_count(),
                    passes);
    }
    DeviceInterface::instance()->shutdown();
    return 0;
}
