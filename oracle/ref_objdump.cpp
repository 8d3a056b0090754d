// Test infrastructure: dumps what the reference's own OBJ loader (ModelOBJ,
// src/rt/obj-loader.cpp, compiled from the reference tree by Makefile.ref)
// makes of a model, for the loader's parity tests.  Writes to stdout:
//   "V T G" then the aggregate mesh (get_aggregate_mesh, :1354-1390) as
//   V lines "px py pz nx ny nz u v" (%a, exact) and T lines "i0 i1 i2";
//   then per model mesh (getModelMesh order) "first count alpha shininess
//   dr dg db <color map or ->" and its get_meshes mesh (:1292-1352):
//   "nv nt" with nv vertex lines and nt triangle lines.
#include <cstdio>

#include <rt/obj-loader.hpp>

static void vertex_line(const Vertex& v) {
    std::printf("%a %a %a %a %a %a %a %a\n", v.position.s[0], v.position.s[1], v.position.s[2], v.normal.s[0],
                v.normal.s[1], v.normal.s[2], v.texCoord.s[0], v.texCoord.s[1]);
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const bool rebuild = argc > 2 && argv[2][0] == '1';
    ModelOBJ obj;
    if (!obj.import(argv[1], rebuild)) return 1;
    Mesh agg;
    obj.get_aggregate_mesh(&agg);
    std::vector<Mesh> meshes;
    obj.get_meshes(meshes);
    std::printf("%zu %zu %d\n", agg.vertexCount(), agg.triangleCount(), obj.getNumberOfModelMeshes());
    for (size_t i = 0; i < agg.vertexCount(); ++i) vertex_line(agg.vertices[i]);
    for (size_t t = 0; t < agg.triangleCount(); ++t)
        std::printf("%u %u %u\n", agg.triangles[t].v[0], agg.triangles[t].v[1], agg.triangles[t].v[2]);
    for (int g = 0; g < obj.getNumberOfModelMeshes(); ++g) {
        const ModelOBJ::ModelMesh& mm = obj.getModelMesh(g);
        const ModelOBJ::Material* mat = mm.pMaterial;
        std::printf("%d %d %a %a %a %a %a %s\n", mm.startIndex / 3, mm.triangleCount, mat->alpha, mat->shininess,
                    mat->diffuse[0], mat->diffuse[1], mat->diffuse[2],
                    mat->colorMapFilename.empty() ? "-" : mat->colorMapFilename.c_str());
        const Mesh& m = meshes[g];
        std::printf("%zu %zu\n", m.vertexCount(), m.triangleCount());
        for (size_t i = 0; i < m.vertexCount(); ++i) vertex_line(m.vertices[i]);
        for (size_t t = 0; t < m.triangleCount(); ++t)
            std::printf("%u %u %u\n", m.triangles[t].v[0], m.triangles[t].v[1], m.triangles[t].v[2]);
    }
    return 0;
}
