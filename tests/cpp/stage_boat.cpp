// The reference's boat scene (boat_set_scene, include/rt/test-params.hpp:10-50)
// through the C++ stage classes: OBJ models loaded as aggregate meshes,
// objects with transforms and materials, create_aggregate_mesh, then the
// Renderer's frame (src/rt/rt.cpp:306-409, 515-530).
//
//   stage_boat <models_dir> <out_dir> <width> <height> <view>
// out_dir: accum.u32
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "rt/stages.hpp"

int main(int argc, char** argv) {
    if (argc < 6) {
        std::cerr << "usage: stage_boat <models_dir> <out_dir> <width> <height> <view>\n";
        return 2;
    }
    const std::string models = argv[1], out = argv[2];
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]), view = std::atoi(argv[5]);
    const vec3 cam_pos[3] = {{20.0186f, -5.49632f, 71.8718f}, {-56.9387f, -2.41959f, 29.574f},
                             {68.6859f, 5.18034f, 13.6691f}};
    const vec3 cam_dir[3] = {{0.131732f, 0.0845093f, -0.987677f}, {0.927574f, -0.22893f, -0.295292f},
                             {-0.820819f, -0.478219f, -0.31235f}};
    if (DeviceInterface::instance()->initialize(0)) return 1;
    {
        Scene scene;
        if (scene.initialize()) return 1;
        // setScale(2) then setPos(0, -8, 0): M = T(0, -8, 0) * S(2), rpy 0
        const float M[16] = {2, 0, 0, 0, 0, 2, 0, -8, 0, 0, 2, 0, 0, 0, 0, 1};
        const float Mi[16] = {0.5f, 0, 0, 0, 0, 0.5f, 0, 4, 0, 0, 0.5f, 0, 0, 0, 0, 1};
        const int32_t water = scene.load_obj_file_as_aggregate(models + "/frame_water1.obj");
        const int32_t boat = scene.load_obj_file_as_aggregate(models + "/frame_boat1.obj");
        if (water < 0 || boat < 0) return 1;
        rtx_material blue{}, red{};
        blue.diffuse[2] = 1.f;
        blue.reflectiveness = 0.98f;
        blue.refractive_index = 1.333f;
        blue.texture = -1;
        red.diffuse[0] = 1.f;
        red.shininess = 1.f;
        red.texture = -1;
        const int32_t ow = scene.add_object(water), ob = scene.add_object(boat);
        if (scene.set_object_transform(ow, M, Mi) || scene.set_object_material(ow, blue) ||
            scene.set_object_transform(ob, M, Mi) || scene.set_object_material(ob, red))
            return 1;
        if (scene.set_dir_light(vec3{0.05f, -1.f, -0.02f}, vec3{0.7f, 0.7f, 0.7f}) ||
            scene.set_ambient_light(vec3{0.2f, 0.2f, 0.2f}) || scene.create_aggregate_mesh())
            return 1;
        scene.camera.set(cam_pos[view], cam_dir[view], vec3{0.f, 1.f, 0.f}, (float)(M_PI / 4.), (float)w / (float)h);
        Renderer r;
        r.configure_from_defaults();
        r.resize_output((size_t)w, (size_t)h);
        r.set_max_bounces(3);
        r.config.bvh_depth = 30;
        r.config.bvh_min_leaf_size = 1;
        if (r.initialize()) return 1;
        if (r.set_up_frame(scene) || r.render_to_framebuffer(scene) || r.copy_framebuffer() || r.conclude_frame(scene))
            return 1;
        std::vector<uint32_t> acc((size_t)w * h * 3);
        if (r.framebuffer().read_accum(acc.data())) return 1;
        std::ofstream f(out + "/accum.u32", std::ios::binary);
        f.write(reinterpret_cast<const char*>(acc.data()), (std::streamsize)(acc.size() * 4));
        const FrameStats& st = r.get_frame_stats();
        std::printf("stage_boat: %zu triangles, %zu rays (%zu secondary)\n", scene.triangle_count(),
                    st.get_ray_count(), st.get_secondary_ray_count());
    }
    DeviceInterface::instance()->shutdown();
    return 0;
}
