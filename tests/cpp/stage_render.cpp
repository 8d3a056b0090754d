// Renders a scene through the C++ stage classes (include/rt/stages.hpp), the
// way the reference's apps drive Renderer (src/rt/rt.cpp:306-409):
// set_up_frame -> render_to_framebuffer -> copy_framebuffer -> conclude_frame.
//
//   stage_render <scene_dir> <out_dir>
// scene_dir: pos.f32 nrm.f32 uv.f32 idx.u32 matmap.i32 mats.bin lights.bin,
//            view.txt "px py pz lx ly lz ux uy uz fov",
//            params.txt "width height max_bounces bvh_depth min_leaf frames"
//            optional images.txt (stage_io.hpp)
// out_dir:   accum.u32 rgba.u8 stats.txt
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "stage_io.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: stage_render <scene_dir> <out_dir>\n";
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
        Scene scene;
        if (scene.initialize() ||
            scene.set_geometry(pos.data(), nrm.data(), uv.data(), (uint32_t)(pos.size() / 3), idx.data(), mm.data(),
                               (uint32_t)(idx.size() / 3)) ||
            scene.set_materials(mats) || scene.set_lights(lights[0]))
            return 1;
        if (const int e = load_images(in, scene)) return e;
        vec3 p{view[0], view[1], view[2]}, look{view[3], view[4], view[5]}, up{view[6], view[7], view[8]};
        scene.camera.set(p, vec3{look.x - p.x, look.y - p.y, look.z - p.z}, up, view[9], (float)w / (float)h);
        Renderer r;
        r.configure_from_defaults();
        r.resize_output((size_t)w, (size_t)h);
        r.set_max_bounces((uint32_t)bounces);
        r.config.bvh_depth = depth;
        r.config.bvh_min_leaf_size = min_leaf;
        if (r.initialize()) return 1;
        for (int f = 0; f < frames; ++f) {
            if (r.set_up_frame(scene) || r.render_to_framebuffer(scene) || r.copy_framebuffer() ||
                r.conclude_frame(scene))
                return 1;
        }
        std::vector<uint32_t> acc((size_t)w * h * 3);
        std::vector<uint8_t> rgba((size_t)w * h * 4);
        if (r.framebuffer().read_accum(acc.data()) || r.framebuffer().read_rgba8(rgba.data())) return 1;
        save(out + "/accum.u32", acc);
        save(out + "/rgba.u8", rgba);
        const FrameStats& st = r.get_frame_stats();
        std::ofstream s(out + "/stats.txt");
        s << st.get_ray_count() << " " << st.get_secondary_ray_count() << " " << scene.bvh_node_count << " "
          << st.get_mean_frame_time() << "\n";
        std::printf("stage_render: %zu rays (%zu secondary), %u nodes, %.3f ms/frame\n", st.get_ray_count(),
                    st.get_secondary_ray_count(), scene.bvh_node_count, st.get_mean_frame_time());
    }
    DeviceInterface::instance()->shutdown();
    return 0;
}
