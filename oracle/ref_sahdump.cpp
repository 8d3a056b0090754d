// Test infrastructure: builds the reference's CPU SAH BVH (BVH::construct,
// src/rt/bvh.cpp, compiled from the reference tree by Makefile.ref) over a
// triangle soup and dumps it, for the SAH builder's parity tests.
//   ref_sahdump <in.bin> <out.bin>
// in:  uint32 V, T; float pos[3V]; uint32 idx[3T]
// out: uint32 N; N x {float lo[3], hi[3]; uint32 a, b, parent, leaf};
//      uint32 order[T] (sorted position -> input triangle)
#include <cstdio>
#include <vector>

#include <rt/bvh.hpp>

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t V = 0, T = 0;
    if (std::fread(&V, 4, 1, f) != 1 || std::fread(&T, 4, 1, f) != 1) return 2;
    std::vector<float> pos(3 * (size_t)V);
    std::vector<uint32_t> idx(3 * (size_t)T);
    if (std::fread(pos.data(), 4, pos.size(), f) != pos.size() || std::fread(idx.data(), 4, idx.size(), f) != idx.size())
        return 2;
    std::fclose(f);
    Mesh mesh;
    mesh.vertices.resize(V);
    for (uint32_t v = 0; v < V; ++v)
        for (int k = 0; k < 3; ++k) mesh.vertices[v].position.s[k] = pos[3 * (size_t)v + k];
    mesh.triangles.resize(T);
    for (uint32_t t = 0; t < T; ++t)
        for (int k = 0; k < 3; ++k) mesh.triangles[t].v[k] = idx[3 * (size_t)t + k];
    BVH bvh;
    if (bvh.construct(mesh)) return 1;
    FILE* o = std::fopen(argv[2], "wb");
    const uint32_t N = (uint32_t)bvh.nodeArraySize();
    std::fwrite(&N, 4, 1, o);
    for (uint32_t n = 0; n < N; ++n) {
        const BVHNode& nd = bvh.nodeArray()[n];
        const float box[6] = {nd.m_bbox.lo.s[0], nd.m_bbox.lo.s[1], nd.m_bbox.lo.s[2],
                              nd.m_bbox.hi.s[0], nd.m_bbox.hi.s[1], nd.m_bbox.hi.s[2]};
        const uint32_t w[4] = {nd.m_l_child, nd.m_r_child, nd.m_parent, (uint32_t)(nd.m_leaf != 0)};
        std::fwrite(box, 4, 6, o);
        std::fwrite(w, 4, 4, o);
    }
    std::fwrite(bvh.triangle_order_array(), 4, T, o);
    std::fclose(o);
    return 0;
}
