// Wavefront OBJ/MTL ingest with the semantics of the reference's loader
// (ModelOBJ, src/rt/obj-loader.cpp:163-227, 306-545, 546-622, 798-1290; its
// meshes via get_aggregate_mesh / get_meshes, :1292-1390), so that a model
// loads into the same vertex and index arrays:
//   - faces of 3+ corners in the forms v, v/vt, v//vn, v/vt/vn (the first
//     corner decides the form of the face), fanned around the first corner;
//     a negative index i resolves to i + count - 1, count being the elements
//     read so far (the reference's rule, one below the OBJ convention);
//   - a corner becomes a vertex {position, uv, normal}; corners of the same
//     position index with identical bytes share one vertex, numbered in order
//     of first use (addVertex's cache, keyed by the position index);
//   - without normals in the file (or when asked) vertex normals are the sum
//     of the unnormalized face normals of their triangles, times 1/sqrt(len2);
//   - triangles keep the file order; runs of one material form groups, ordered
//     by material alpha, opaque first, runs of equal alpha in file order;
//   - MTL keywords dispatch on their first letter as the reference's do:
//     n* newmtl, N* shininess/1000 (Ns, and Ni too), Ka/Kd/Ks, Tr (alpha =
//     1 - Tr), d* alpha, i* illum (1: no specular), m* containing map_Kd or
//     map_bump; a later mtllib replaces the materials from index 0 while the
//     names read before stay known; without materials one default material;
//     usemtl of an unknown name selects material 0;
//   - a name argument (mtllib, usemtl, newmtl, map_*) is the second word
//     after the keyword when there are two, as sscanf("%s %s") into one
//     buffer leaves it, else the first.
// Tangents and bitangents (built by the reference only for bump-mapped
// materials) are not produced: the ray-tracing path reads positions,
// normals and uvs.  Parsing is line based; a face does not continue onto the
// next line.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rtx.h"

void rtx_set_error(const std::string& msg);

struct ObjMaterial {
    float ambient[4] = {0.2f, 0.2f, 0.2f, 1.f};
    float diffuse[4] = {0.8f, 0.8f, 0.8f, 1.f};
    float specular[4] = {0.f, 0.f, 0.f, 1.f};
    float shininess = 0.f;
    float alpha = 1.f;
    std::string name = "default", color_map, bump_map;
};

struct ObjGroup {
    uint32_t first, count;  // triangles
    int material;
};

struct rtx_obj {
    std::vector<float> pos, uv, nrm;  // per vertex: 3, 2, 3
    std::vector<uint32_t> idx;        // 3 per triangle
    std::vector<int> tri_mat;
    std::vector<ObjMaterial> mats;
    std::vector<ObjGroup> groups;
};

namespace {

struct Line {
    std::vector<std::string> tok;
};

std::vector<Line> read_lines(const std::string& path, bool* ok) {
    std::vector<Line> out;
    FILE* f = std::fopen(path.c_str(), "rb");
    *ok = f != nullptr;
    if (!f) return out;
    std::string text;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    std::fclose(f);
    size_t i = 0;
    while (i < text.size()) {
        size_t e = text.find('\n', i);
        if (e == std::string::npos) e = text.size();
        Line ln;
        size_t k = i;
        while (k < e) {
            while (k < e && std::isspace((unsigned char)text[k])) ++k;
            size_t s = k;
            while (k < e && !std::isspace((unsigned char)text[k])) ++k;
            if (k > s) ln.tok.emplace_back(text, s, k - s);
        }
        if (!ln.tok.empty()) out.push_back(std::move(ln));
        i = e + 1;
    }
    return out;
}

bool to_float(const std::string& s, float* v) {
    char* end = nullptr;
    errno = 0;
    *v = std::strtof(s.c_str(), &end);
    return end != s.c_str();
}

// One face corner: position, texture and normal indices as written (0 = absent).
struct Corner {
    long v = 0, vt = 0, vn = 0;
};

// Parses a corner in the face's form; false when the token does not fit it.
enum Form { F_V, F_VT, F_VN, F_VTVN };
bool parse_corner(const std::string& t, Form form, Corner* c) {
    const char* p = t.c_str();
    char* e = nullptr;
    c->v = std::strtol(p, &e, 10);
    if (e == p) return false;
    if (form == F_V) return true;
    if (*e != '/') return false;
    p = e + 1;
    if (form == F_VN) {
        if (*p != '/') return false;
        c->vn = std::strtol(p + 1, &e, 10);
        return e != p + 1;
    }
    c->vt = std::strtol(p, &e, 10);
    if (e == p) return false;
    if (form == F_VT) return true;
    if (*e != '/') return false;
    p = e + 1;
    c->vn = std::strtol(p, &e, 10);
    return e != p;
}

Form face_form(const std::string& t) {
    if (t.find("//") != std::string::npos) return F_VN;
    Corner c;
    if (parse_corner(t, F_VTVN, &c)) return F_VTVN;
    if (parse_corner(t, F_VT, &c)) return F_VT;
    return F_V;
}

// Index as written -> 0-based (negative: i + count - 1, the reference's rule).
long resolve(long i, long count) { return i < 0 ? i + count - 1 : i - 1; }

// The name argument of a keyword line.
std::string name_arg(const Line& l) { return l.tok.size() > 2 ? l.tok[2] : (l.tok.size() > 1 ? l.tok[1] : ""); }

bool load_mtl(const std::string& path, rtx_obj* m, std::map<std::string, int>* names) {
    bool ok = false;
    std::vector<Line> lines = read_lines(path, &ok);
    if (!ok) return false;
    m->mats.clear();
    ObjMaterial* cur = nullptr;
    auto f3 = [](const Line& l, float* d) {
        for (int k = 0; k < 3 && k + 1 < (int)l.tok.size(); ++k) to_float(l.tok[k + 1], &d[k]);
        d[3] = 1.f;
    };
    for (const Line& l : lines) {
        const std::string& k = l.tok[0];
        if (k[0] == 'n') {
            m->mats.emplace_back();
            cur = &m->mats.back();
            cur->name = name_arg(l);
            (*names)[cur->name] = (int)m->mats.size() - 1;
            continue;
        }
        if (!cur) continue;
        if (k[0] == 'N') {
            if (l.tok.size() > 1) to_float(l.tok[1], &cur->shininess);
            cur->shininess /= 1000.0f;  // [0, 1000] -> [0, 1]
        } else if (k[0] == 'K' && k.size() > 1 && k[1] == 'a') {
            f3(l, cur->ambient);
        } else if (k[0] == 'K' && k.size() > 1 && k[1] == 'd') {
            f3(l, cur->diffuse);
        } else if (k[0] == 'K' && k.size() > 1 && k[1] == 's') {
            f3(l, cur->specular);
        } else if (k[0] == 'T' && k.size() > 1 && k[1] == 'r') {
            if (l.tok.size() > 1) to_float(l.tok[1], &cur->alpha);
            cur->alpha = 1.0f - cur->alpha;
        } else if (k[0] == 'd') {
            if (l.tok.size() > 1) to_float(l.tok[1], &cur->alpha);
        } else if (k[0] == 'i') {
            if (l.tok.size() > 1 && std::atoi(l.tok[1].c_str()) == 1) {
                cur->specular[0] = cur->specular[1] = cur->specular[2] = 0.f;
                cur->specular[3] = 1.f;
            }
        } else if (k[0] == 'm' && k.find("map_Kd") != std::string::npos) {
            cur->color_map = name_arg(l);
        } else if (k[0] == 'm' && k.find("map_bump") != std::string::npos) {
            cur->bump_map = name_arg(l);
        }
    }
    return true;
}

struct VKey {
    long v;
    uint32_t bits[5];  // uv, normal
    bool operator==(const VKey& o) const { return v == o.v && std::memcmp(bits, o.bits, sizeof bits) == 0; }
};
struct VKeyHash {
    size_t operator()(const VKey& k) const {
        uint64_t h = (uint64_t)k.v * 0x9E3779B97F4A7C15ull;
        for (uint32_t b : k.bits) h = (h ^ b) * 0x100000001B3ull;
        return (size_t)h;
    }
};

// Summed face normals, then times 1/sqrt(len2) (generateNormals).
void vertex_normals(rtx_obj* m) {
    const size_t V = m->pos.size() / 3, T = m->idx.size() / 3;
    std::fill(m->nrm.begin(), m->nrm.end(), 0.f);
    for (size_t t = 0; t < T; ++t) {
        const uint32_t a = m->idx[3 * t], b = m->idx[3 * t + 1], c = m->idx[3 * t + 2];
        float e1[3], e2[3], n[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = m->pos[3 * b + k] - m->pos[3 * a + k];
            e2[k] = m->pos[3 * c + k] - m->pos[3 * a + k];
        }
        n[0] = (e1[1] * e2[2]) - (e1[2] * e2[1]);
        n[1] = (e1[2] * e2[0]) - (e1[0] * e2[2]);
        n[2] = (e1[0] * e2[1]) - (e1[1] * e2[0]);
        for (uint32_t v : {a, b, c})
            for (int k = 0; k < 3; ++k) m->nrm[3 * v + k] += n[k];
    }
    for (size_t v = 0; v < V; ++v) {
        float* n = &m->nrm[3 * v];
        const float inv = 1.0f / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        n[0] *= inv;
        n[1] *= inv;
        n[2] *= inv;
    }
}

}  // namespace

int32_t rtx_obj_load(const char* path, int32_t rebuild_normals, rtx_obj** out) {
    if (!path || !out) {
        rtx_set_error("rtx_obj_load: null argument");
        return RTX_EINVAL;
    }
    bool ok = false;
    std::vector<Line> lines = read_lines(path, &ok);
    if (!ok) {
        rtx_set_error(std::string("rtx_obj_load: cannot open ") + path);
        return RTX_EINVAL;
    }
    rtx_obj* m = new rtx_obj();
    std::string dir(path);
    const size_t slash = dir.find_last_of("\\/");
    dir = slash == std::string::npos ? std::string() : dir.substr(0, slash + 1);
    std::map<std::string, int> names;
    // materials first (the reference reads every mtllib in a first pass)
    for (const Line& l : lines)
        if (l.tok[0][0] == 'm') load_mtl(dir + name_arg(l), m, &names);
    if (m->mats.empty()) {
        m->mats.emplace_back();
        names["default"] = 0;
    }
    std::vector<float> vp, vt, vn;
    std::unordered_map<VKey, uint32_t, VKeyHash> cache;
    int active = 0;
    auto fail = [&](const std::string& msg) {
        rtx_set_error("rtx_obj_load: " + msg);
        delete m;
        return RTX_EINVAL;
    };
    for (const Line& l : lines) {
        const std::string& k = l.tok[0];
        if (k == "v" || k == "vn" || k == "vt") {
            std::vector<float>& dst = k == "v" ? vp : (k == "vn" ? vn : vt);
            const int nc = k == "vt" ? 2 : 3;
            for (int c = 0; c < nc; ++c) {
                float x = 0.f;
                if (c + 1 < (int)l.tok.size()) to_float(l.tok[c + 1], &x);
                dst.push_back(x);
            }
        } else if (k[0] == 'u') {
            auto it = names.find(name_arg(l));
            active = (it == names.end() || it->second >= (int)m->mats.size()) ? 0 : it->second;
        } else if (k[0] == 'f') {
            if (l.tok.size() < 4) return fail("face with fewer than 3 corners");
            const Form form = face_form(l.tok[1]);
            const long nv = (long)vp.size() / 3, nt = (long)vt.size() / 2, nn = (long)vn.size() / 3;
            std::vector<uint32_t> ids;
            for (size_t c = 1; c < l.tok.size(); ++c) {
                Corner cr;
                if (!parse_corner(l.tok[c], form, &cr)) break;  // the face ends at a corner of another form
                const long v = resolve(cr.v, nv);
                if (v < 0 || v >= nv) return fail("position index out of range");
                float vert[8] = {vp[3 * v], vp[3 * v + 1], vp[3 * v + 2], 0.f, 0.f, 0.f, 0.f, 0.f};
                if (form == F_VT || form == F_VTVN) {
                    const long t = resolve(cr.vt, nt);
                    if (t < 0 || t >= nt) return fail("texture index out of range");
                    vert[3] = vt[2 * t];
                    vert[4] = vt[2 * t + 1];
                }
                if (form == F_VN || form == F_VTVN) {
                    const long n = resolve(cr.vn, nn);
                    if (n < 0 || n >= nn) return fail("normal index out of range");
                    vert[5] = vn[3 * n];
                    vert[6] = vn[3 * n + 1];
                    vert[7] = vn[3 * n + 2];
                }
                VKey key;
                key.v = v;
                std::memcpy(key.bits, vert + 3, sizeof key.bits);
                auto it = cache.find(key);
                uint32_t id;
                if (it == cache.end()) {
                    id = (uint32_t)(m->pos.size() / 3);
                    cache.emplace(key, id);
                    m->pos.insert(m->pos.end(), vert, vert + 3);
                    m->uv.insert(m->uv.end(), vert + 3, vert + 5);
                    m->nrm.insert(m->nrm.end(), vert + 5, vert + 8);
                } else {
                    id = it->second;
                }
                ids.push_back(id);
            }
            if (ids.size() < 3) return fail("face with fewer than 3 corners");
            for (size_t c = 2; c < ids.size(); ++c) {  // fan around the first corner
                m->idx.insert(m->idx.end(), {ids[0], ids[c - 1], ids[c]});
                m->tri_mat.push_back(active);
            }
        }
    }
    if (rebuild_normals || vn.empty()) vertex_normals(m);
    // groups: runs of one material, opaque first (stable among equal alpha)
    for (uint32_t t = 0; t < (uint32_t)m->tri_mat.size(); ++t) {
        if (t == 0 || m->tri_mat[t] != m->tri_mat[t - 1])
            m->groups.push_back(ObjGroup{t, 0, m->tri_mat[t]});
        ++m->groups.back().count;
    }
    std::stable_sort(m->groups.begin(), m->groups.end(), [&](const ObjGroup& a, const ObjGroup& b) {
        return m->mats[a.material].alpha > m->mats[b.material].alpha;
    });
    *out = m;
    return RTX_OK;
}

int32_t rtx_obj_destroy(rtx_obj* m) {
    delete m;
    return RTX_OK;
}

int32_t rtx_obj_info(rtx_obj* m, uint32_t* vertices, uint32_t* triangles, uint32_t* groups) {
    if (!m) {
        rtx_set_error("rtx_obj_info: null model");
        return RTX_EINVAL;
    }
    if (vertices) *vertices = (uint32_t)(m->pos.size() / 3);
    if (triangles) *triangles = (uint32_t)(m->idx.size() / 3);
    if (groups) *groups = (uint32_t)m->groups.size();
    return RTX_OK;
}

int32_t rtx_obj_read(rtx_obj* m, float* pos3, float* nrm3, float* uv2, uint32_t* idx3) {
    if (!m) {
        rtx_set_error("rtx_obj_read: null model");
        return RTX_EINVAL;
    }
    if (pos3) std::copy(m->pos.begin(), m->pos.end(), pos3);
    if (nrm3) std::copy(m->nrm.begin(), m->nrm.end(), nrm3);
    if (uv2) std::copy(m->uv.begin(), m->uv.end(), uv2);
    if (idx3) std::copy(m->idx.begin(), m->idx.end(), idx3);
    return RTX_OK;
}

int32_t rtx_obj_group_info(rtx_obj* m, uint32_t g, rtx_obj_group* out) {
    if (!m || !out || g >= m->groups.size()) {
        rtx_set_error("rtx_obj_group_info: null argument or group out of range");
        return RTX_EINVAL;
    }
    const ObjGroup& G = m->groups[g];
    const ObjMaterial& M = m->mats[G.material];
    std::memset(out, 0, sizeof *out);
    out->first_triangle = G.first;
    out->triangle_count = G.count;
    out->material = G.material;
    std::copy(M.ambient, M.ambient + 4, out->ambient);
    std::copy(M.diffuse, M.diffuse + 4, out->diffuse);
    std::copy(M.specular, M.specular + 4, out->specular);
    out->shininess = M.shininess;
    out->alpha = M.alpha;
    std::snprintf(out->name, sizeof out->name, "%s", M.name.c_str());
    std::snprintf(out->color_map, sizeof out->color_map, "%s", M.color_map.c_str());
    std::snprintf(out->bump_map, sizeof out->bump_map, "%s", M.bump_map.c_str());
    return RTX_OK;
}

// get_meshes: group g as a mesh of its own, vertices renumbered in order of
// first use within the group.  (The reference's get_meshes also leaves the
// group's triangle count of zero triangles before the real ones, from a
// resize followed by push_back; those are not produced here.)
int32_t rtx_obj_read_group(rtx_obj* m, uint32_t g, uint32_t* vertices, float* pos3, float* nrm3, float* uv2,
                           uint32_t* idx3) {
    if (!m || g >= m->groups.size()) {
        rtx_set_error("rtx_obj_read_group: null model or group out of range");
        return RTX_EINVAL;
    }
    const ObjGroup& G = m->groups[g];
    std::unordered_map<uint32_t, uint32_t> remap;
    std::vector<uint32_t> order;
    for (uint32_t t = G.first; t < G.first + G.count; ++t)
        for (int c = 0; c < 3; ++c) {
            const uint32_t v = m->idx[3 * (size_t)t + c];
            auto it = remap.find(v);
            uint32_t id;
            if (it == remap.end()) {
                id = (uint32_t)order.size();
                remap.emplace(v, id);
                order.push_back(v);
            } else {
                id = it->second;
            }
            if (idx3) idx3[3 * (size_t)(t - G.first) + c] = id;
        }
    if (vertices) *vertices = (uint32_t)order.size();
    for (size_t i = 0; i < order.size(); ++i) {
        const uint32_t v = order[i];
        if (pos3) std::copy(&m->pos[3 * (size_t)v], &m->pos[3 * (size_t)v] + 3, pos3 + 3 * i);
        if (nrm3) std::copy(&m->nrm[3 * (size_t)v], &m->nrm[3 * (size_t)v] + 3, nrm3 + 3 * i);
        if (uv2) std::copy(&m->uv[2 * (size_t)v], &m->uv[2 * (size_t)v] + 2, uv2 + 2 * i);
    }
    return RTX_OK;
}
