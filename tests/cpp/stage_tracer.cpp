// Drives the Tracer overloads that trace an accelerator copy and the
// BVHBuilder's log (include/rt/stages.hpp), the way the reference's threaded
// renderer keeps a BVH copy while the scene is rebuilt.
//
//   stage_tracer <scene_dir> <out_dir>
// scene_dir: pos.f32 nrm.f32 uv.f32 idx.u32 matmap.i32 mats.bin lights.bin,
//            view.txt "px py pz lx ly lz ux uy uz fov",
//            params.txt "width height max_bounces bvh_depth min_leaf frames"
// out_dir:   bvh.log (the builder's log)
//            copy_{flags,t,id}.bin: primary rays traced (closest hit, then
//              shadow rays) against the copy taken after the first build,
//              once the scene has moved (+1 in x) and been rebuilt
//            moved_{flags,t,id}.bin: the same rays against the moved scene
//            first_{t,id}.bin: the same rays against the first build
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "stage_io.hpp"

struct Hits {
    std::vector<uint32_t> flags;
    std::vector<float> t;
    std::vector<int32_t> id;
};

static int read_hits(HitBundle& hb, size_t n, Hits& h) {
    h.flags.assign(n, 0);
    h.t.assign(n, 0.f);
    h.id.assign(n, 0);
    return rtx_hits_read(hb.handle(), h.flags.data(), h.t.data(), h.id.data(), nullptr, nullptr, nullptr, nullptr,
                         n) == RTX_OK
               ? 0
               : 1;
}

static void save_hits(const std::string& prefix, const Hits& h, bool flags) {
    if (flags) save(prefix + "_flags.bin", h.flags);
    save(prefix + "_t.bin", h.t);
    save(prefix + "_id.bin", h.id);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: stage_tracer <scene_dir> <out_dir>\n";
        return 2;
    }
    const std::string in = argv[1], out = argv[2];
    auto pos = load<float>(in + "/pos.f32");
    auto nrm = load<float>(in + "/nrm.f32");
    auto uv = load<float>(in + "/uv.f32");
    auto idx = load<uint32_t>(in + "/idx.u32");
    auto mm = load<int32_t>(in + "/matmap.i32");
    auto mats = load<rtx_material>(in + "/mats.bin");
    auto lights = load<rtx_lights>(in + "/lights.bin");
    float view[10];
    int w, h, bounces, depth, min_leaf, frames;
    {
        std::ifstream v(in + "/view.txt");
        for (float& x : view) v >> x;
        std::ifstream p(in + "/params.txt");
        p >> w >> h >> bounces >> depth >> min_leaf >> frames;
    }
    if (pos.empty() || idx.empty() || mats.empty() || lights.size() != 1) {
        std::cerr << "bad scene directory\n";
        return 2;
    }
    if (DeviceInterface::instance()->initialize(0)) return 1;
    {
        const uint32_t V = (uint32_t)(pos.size() / 3), T = (uint32_t)(idx.size() / 3);
        Scene scene;
        if (scene.initialize() || scene.set_geometry(pos.data(), nrm.data(), uv.data(), V, idx.data(), mm.data(), T) ||
            scene.set_materials(mats) || scene.set_lights(lights[0]))
            return 1;
        vec3 p{view[0], view[1], view[2]}, look{view[3], view[4], view[5]}, up{view[6], view[7], view[8]};
        scene.camera.set(p, vec3{look.x - p.x, look.y - p.y, look.z - p.z}, up, view[9], (float)w / (float)h);
        RendererConfig conf;
        conf.bvh_depth = depth;
        conf.bvh_min_leaf_size = min_leaf;

        Log log;
        if (log.initialize(out + "/bvh.log")) return 1;
        log.silent = true;
        BVHBuilder builder;
        if (builder.initialize()) return 1;
        builder.update_configuration(conf);
        builder.set_log(&log);
        builder.logging(true);
        if (builder.build_lbvh(scene)) return 1;

        BVHMemory bvh_mem;
        if (bvh_mem.copy_from(scene)) return 1;

        const size_t n = (size_t)w * h;
        RayBundle rays;
        HitBundle hits;
        PrimaryRayGenerator prg;
        Tracer tracer;
        if (rays.initialize(n) || hits.initialize(n) || prg.initialize() || tracer.initialize()) return 1;
        size_t size[2] = {(size_t)w, (size_t)h};
        if (prg.generate(scene.camera, rays, size, n, 0)) return 1;
        Hits first, copy, moved;
        if (tracer.trace(scene, (int32_t)n, rays, hits) || read_hits(hits, n, first)) return 1;

        // the scene moves and is rebuilt; the copy keeps the first tree
        std::vector<float> pos2(pos);
        for (size_t i = 0; i < pos2.size(); i += 3) pos2[i] += 1.0f;
        if (scene.update_vertex_positions(pos2.data()) || builder.build_lbvh(scene)) return 1;
        if (tracer.trace(scene, bvh_mem, (int32_t)n, rays, hits) ||
            tracer.shadow_trace(scene, bvh_mem, (int32_t)n, rays, hits) || read_hits(hits, n, copy))
            return 1;
        if (tracer.trace(scene, (int32_t)n, rays, hits) || tracer.shadow_trace(scene, (int32_t)n, rays, hits) ||
            read_hits(hits, n, moved))
            return 1;
        save_hits(out + "/first", first, false);
        save_hits(out + "/copy", copy, true);
        save_hits(out + "/moved", moved, true);
        std::printf("stage_tracer: %zu rays\n", n);
    }
    DeviceInterface::instance()->shutdown();
    return 0;
}
