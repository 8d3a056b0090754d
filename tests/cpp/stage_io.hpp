// File helpers of the stage drivers: raw arrays in and out, and the optional
// images of a scene directory.
//   images.txt: "N", then N lines "w h" (textures tex<k>.rgba), then "C"
//               (1: cubemap faces face0..face5.rgba follow as 6 lines "w h")
#pragma once
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include "rt/stages.hpp"

template <class T>
static std::vector<T> load(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return {};
    size_t n = (size_t)f.tellg();
    std::vector<T> v(n / sizeof(T));
    f.seekg(0);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    return v;
}

template <class T>
static void save(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

// Loads images.txt of `dir` into the scene's atlas and cubemap.  0 when the
// file is absent or read whole, 2 on a malformed directory, 1 on a failed call.
static int load_images(const std::string& dir, Scene& scene) {
    std::ifstream imf(dir + "/images.txt");
    if (!imf) return 0;
    int n = 0, cube = 0;
    imf >> n;
    for (int k = 0; k < n; ++k) {
        uint32_t tw = 0, th = 0;
        imf >> tw >> th;
        auto px = load<uint8_t>(dir + "/tex" + std::to_string(k) + ".rgba");
        if (px.size() != 4ull * tw * th || scene.texture_atlas.load_texture_rgba(px.data(), tw, th) != k) return 2;
    }
    imf >> cube;
    if (cube) {
        std::vector<uint8_t> faces[6];
        const uint8_t* fp[6];
        uint32_t dims[12];
        for (int f = 0; f < 6; ++f) {
            imf >> dims[2 * f] >> dims[2 * f + 1];
            faces[f] = load<uint8_t>(dir + "/face" + std::to_string(f) + ".rgba");
            if (faces[f].size() != 4ull * dims[2 * f] * dims[2 * f + 1]) return 2;
            fp[f] = faces[f].data();
        }
        if (scene.cubemap.initialize_rgba(fp, dims)) return 1;
    }
    return 0;
}
