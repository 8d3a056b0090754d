// rt-seq: sequential CPU restatement of Tesina-RT's ray-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product is
// checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or the timed CPU
// baseline; the product never links it.
//
// Every function follows one reference kernel or host routine literally (one
// scalar loop per work-item, same floating-point operation order), so that
// its results define the parity contract:
//   * Morton codes, sorted permutation, BVH topology and node AABBs: bit-exact.
//   * hits (id, t, u, v, n, hit point), shadow flags, secondary rays, shading
//     and the RGBA8 resolve: the same IEEE-754 single-precision op sequence.
// Built with  g++ -O2 -ffp-contract=off  (no FMA contraction, correctly rounded
// divide and sqrt, IEEE fmin/fmax), which is the arithmetic the GPU kernels
// are compiled for as well.
//
// Places where the reference's result depends on its vendor OpenCL compiler
// (normalize, pow, min/max on NaN, write_imagef rounding) are given one
// defined meaning here; the test file that covers each says so.
//
// Pinning (tests/test_oracle_*.py): closed-form Morton vectors, std::sort /
// brute-force split checks of the hierarchy, brute-force closest/any hit over
// all triangles, analytic plane/sphere hits, closed-form shading values.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

// ---------------------------------------------------------------- layouts --
// Reference BVHNode, 48 bytes (include/rt/bvh.hpp:15-74,
// src/kernel/bvh-builder.cl:31-46): BBox{hi,lo} as float4 pairs, then
// l_child/start_index, r_child/end_index, parent, split_axis, leaf.
struct orc_node {
    float hi[4];
    float lo[4];
    uint32_t a;       // l_child or start_index
    uint32_t b;       // r_child or end_index
    uint32_t parent;
    int8_t split_axis;
    int8_t leaf;
    uint8_t pad[2];
};

// Light record of this build (reference lights_cl holds one DIR or SPOT light,
// include/rt/light.hpp:86-92; POINT and several lights are the configs'
// extension, identical in oracle and GPU).
enum { ORC_SPOT = 0, ORC_DIR = 1, ORC_POINT = 2 };
struct orc_light {
    int32_t type;
    float pos[3];
    float dir[3];
    float color[3];
    float angle;
    float radius;
};
struct orc_lights {
    float ambient[3];
    int32_t count;
    orc_light l[8];
};

// material_cl, 32 bytes (include/rt/material.hpp:57-73).
struct orc_material {
    float diffuse[4];
    float shininess;
    float reflectiveness;
    float refractive_index;
    int32_t texture;
};

struct orc_camera {
    float pos[3], dir[3], right[3], up[3];
};

}  // extern "C"

namespace {

struct f3 {
    float x, y, z;
};
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
inline f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
inline f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
// dot/cross as OpenCL defines them, evaluated left to right, no contraction.
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize(v) = v * (1 / sqrt(dot(v, v))), correctly rounded sqrt and divide.
inline f3 normalize(f3 v) {
    float inv = 1.0f / std::sqrt(dot(v, v));
    return v * inv;
}
inline f3 fmin3(f3 a, f3 b) { return f3{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline f3 fmax3(f3 a, f3 b) { return f3{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }
inline f3 recip(f3 v) { return f3{1.0f / v.x, 1.0f / v.y, 1.0f / v.z}; }

struct BBox {
    f3 hi, lo;
};

// Primitive AABB: src/kernel/bvh-builder.cl:100-130.
inline BBox prim_bbox(const float* pos, const uint32_t* idx, size_t tri) {
    f3 v0 = ld3(pos + 4 * (size_t)idx[3 * tri + 0]);
    f3 v1 = ld3(pos + 4 * (size_t)idx[3 * tri + 1]);
    f3 v2 = ld3(pos + 4 * (size_t)idx[3 * tri + 2]);
    BBox b;
    b.hi = b.lo = v0;
    b.hi = fmax3(b.hi, v1);
    b.hi = fmax3(b.hi, v2);
    b.lo = fmin3(b.lo, v1);
    b.lo = fmin3(b.lo, v2);
    return b;
}

inline void merge(BBox& a, const BBox& b) {
    a.lo = fmin3(a.lo, b.lo);
    a.hi = fmax3(a.hi, b.hi);
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ a1 + a2 -----
// Global AABB = union of the primitive AABBs (src/kernel/bvh-builder.cl:318-391;
// the tree reduction there is a min/max, so the sequential union is exact).
void orc_global_bbox(const float* pos, const uint32_t* idx, uint32_t T, float out_lo[3], float out_hi[3]) {
    BBox g;
    g.hi = mk(-1e37f, -1e37f, -1e37f);
    g.lo = mk(1e37f, 1e37f, 1e37f);
    for (uint32_t i = 0; i < T; ++i) merge(g, prim_bbox(pos, idx, i));
    out_lo[0] = g.lo.x; out_lo[1] = g.lo.y; out_lo[2] = g.lo.z;
    out_hi[0] = g.hi.x; out_hi[1] = g.hi.y; out_hi[2] = g.hi.z;
}

// ------------------------------------------------------------------ a3 -----
// morton_encode: src/kernel/bvh-builder.cl:181-252.  Returns the 64-bit code
// (code[1] << 32 | code[0]).
uint64_t orc_morton_one(const float lo[3], const float hi[3], const float glo[3], const float ghi[3], int depth) {
    f3 center = (ld3(hi) + ld3(lo)) * 0.5f;
    f3 g_lo = ld3(glo), g_hi = ld3(ghi);
    f3 inv = recip(g_hi - g_lo);
    f3 slice = (center - g_lo) * inv;
    slice = fmax3(mk(0.f, 0.f, 0.f), fmin3(mk(1.f, 1.f, 1.f), slice));
    // pow(2, ceil(depth/3.0)) is an exact power of two.
    float scale = std::ldexp(1.0f, (depth + 2) / 3);
    slice = slice * scale;
    uint32_t m[3] = {(uint32_t)slice.x, (uint32_t)slice.y, (uint32_t)slice.z};
    int offset = depth % 3 ? 3 - (depth % 3) : 0;
    int axis = (offset + 1) % 3;
    uint64_t code = 0;
    for (int i = 0; i < depth; ++i) {
        int src_bit = (i + offset) / 3;
        code |= (uint64_t)((m[axis] >> src_bit) & 1u) << i;
        axis = (axis + 1) % 3;
    }
    return code;
}

void orc_morton(const float* pos, const uint32_t* idx, uint32_t T, int depth, uint64_t* codes) {
    float glo[3], ghi[3];
    orc_global_bbox(pos, idx, T, glo, ghi);
    for (uint32_t i = 0; i < T; ++i) {
        BBox b = prim_bbox(pos, idx, i);
        float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
        codes[i] = orc_morton_one(lo, hi, glo, ghi, depth);
    }
}

// ------------------------------------------------------------------ a4 -----
// Canonical order of the sort: stable by (code, input index).  The reference
// bitonic network (src/kernel/bvh-builder-sort.cl:9-232) orders equal codes
// differently; equal-code runs never split (find_middle), so topology and
// AABBs do not depend on the tie order.
void orc_sort(const uint64_t* codes, uint32_t T, uint32_t* perm) {
    for (uint32_t i = 0; i < T; ++i) perm[i] = i;
    std::stable_sort(perm, perm + T, [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
}

// Emulation of the reference's indirect bitonic network, one compare-exchange
// stage at a time (morton_sort_g2, src/kernel/bvh-builder-sort.cl:26-58, with
// the host schedule of src/rt/bvh-builder.cpp:361-454).  Only for tie-order
// diagnostics; g4/g8/g16 launches group the same g2 stages.
void orc_bitonic_emulate(const uint64_t* codes, uint32_t T, uint32_t* tri) {
    int size = (int)T;
    int n2 = 1;
    while (n2 < size) n2 <<= 1;
    for (int i = 0; i < size; ++i) tri[i] = (uint32_t)i;
    for (int len = 1; len < n2; len <<= 1) {
        int dir = len << 1;
        int inv = ((size - 1) & dir) == 0 && dir != n2;
        int last_step = dir == n2;
        for (int inc = len; inc > 0; inc >>= 1) {
            // a stage's compare-exchanges touch disjoint pairs (i, i+inc) with
            // bit inc of i clear, so sequential order is exact
            for (int i = 0; i + inc < size; ++i) {
                if ((i & inc) != 0) continue;
                uint64_t x0 = codes[tri[i]], x1 = codes[tri[i + inc]];
                bool order_asc = !(dir & i) && !last_step;
                bool inc_is_smaller = x1 < x0;
                bool swap = (inc_is_smaller == order_asc) == (bool)inv;
                if (swap) std::swap(tri[i], tri[i + inc]);
            }
        }
    }
}

// ------------------------------------------------------------- a5 + a6 -----
// find_middle: src/kernel/bvh-builder.cl:484-552, followed literally.
// Returns the split index (0 = leaf); *level is updated as node->leaf is.
static uint32_t find_middle(uint32_t start, uint32_t end, int* level_io, const uint64_t* codes, int max_level,
                            int min_leaf) {
    auto codebit = [&](uint32_t x, int bit) -> bool { return (codes[x] >> bit) & 1u; };
    uint32_t first = start;
    uint32_t last = end - 1;
    int level = *level_io;
    int bit = max_level - 1 - level;
    if (last - first - 1 <= (uint32_t)min_leaf) {
        *level_io = 1;
        return 0;
    }
    if (bit < 0) {
        *level_io = 1;
        return 0;
    }
    while (codebit(first, bit) || !codebit(last, bit)) {
        level++;
        bit = max_level - 1 - level;
        if (bit < 0) {
            *level_io = 1;
            return 0;
        }
    }
    *level_io = level;
    while (first + 1 < last) {
        uint32_t middle = (first + last) / 2;
        if (!codebit(middle, bit))
            first = middle;
        else
            last = middle;
    }
    return last;
}

// Full LBVH build: src/rt/bvh-builder.cpp:201-741 with the kernels of
// src/kernel/bvh-builder.cl.  In/out: idx (3T) and matmap (T) are permuted in
// place into sorted order (src/rt/bvh-builder.cpp:481-511).  Outputs: sorted
// codes, the permutation (sorted position -> input position), nodes in
// canonical level order (each level block ordered by start), node count,
// per-iteration output counts OC[0..63], and the global AABB.
// Returns 0 on success, -1 on bad arguments.
int orc_lbvh_build(const float* pos, uint32_t* idx, int32_t* matmap, uint32_t T, int depth, int min_leaf,
                   uint64_t* codes_sorted, uint32_t* perm, orc_node* nodes, uint32_t node_cap, uint32_t* node_count,
                   uint32_t* level_counts, float* gbox6) {
    if (T == 0 || !pos || !idx) return -1;
    depth = std::min(std::max(depth, 1), 64);        // src/rt/bvh-builder.cpp:197
    min_leaf = std::min(std::max(min_leaf, 1), 256);  // src/rt/bvh-builder.cpp:198
    std::vector<uint64_t> codes(T);
    orc_morton(pos, idx, T, depth, codes.data());
    if (gbox6) orc_global_bbox(pos, idx, T, gbox6, gbox6 + 3);
    std::vector<uint32_t> p(T);
    orc_sort(codes.data(), T, p.data());
    // prim AABBs in input order (bboxes[] of the reference)
    std::vector<BBox> pb(T);
    for (uint32_t i = 0; i < T; ++i) pb[i] = prim_bbox(pos, idx, i);
    // rearrange_indices / rearrange_map / rearrange_morton_codes
    std::vector<uint32_t> idx_in(idx, idx + 3 * (size_t)T);
    for (uint32_t i = 0; i < T; ++i)
        for (int k = 0; k < 3; ++k) idx[3 * (size_t)i + k] = idx_in[3 * (size_t)p[i] + k];
    if (matmap) {
        std::vector<int32_t> mm(matmap, matmap + T);
        for (uint32_t i = 0; i < T; ++i) matmap[i] = mm[p[i]];
    }
    std::vector<uint64_t> sc(T);
    for (uint32_t i = 0; i < T; ++i) sc[i] = codes[p[i]];
    if (codes_sorted) std::memcpy(codes_sorted, sc.data(), sizeof(uint64_t) * T);
    if (perm) std::memcpy(perm, p.data(), sizeof(uint32_t) * T);

    // process_task x depth (src/kernel/bvh-builder.cl:582-671), sequential
    // in-order processing = canonical level order.
    std::vector<orc_node> nd;
    nd.reserve(2 * (size_t)T);
    orc_node root;
    std::memset(&root, 0, sizeof root);
    root.a = 0;
    root.b = T;
    root.parent = 0;
    root.leaf = 0;
    nd.push_back(root);
    std::vector<int> lvl(1, 0);  // node->leaf used as the level during the build
    uint32_t OC[64] = {0};
    size_t begin = 0, end = 1;
    for (int it = 0; it < depth; ++it) {
        for (size_t n = begin; n < end; ++n) {
            int level = lvl[n];
            uint32_t mid = find_middle(nd[n].a, nd[n].b, &level, sc.data(), depth, min_leaf);
            if (!mid) {
                lvl[n] = level;  // == 1
                continue;
            }
            // set_children: src/kernel/bvh-builder.cl:554-580
            uint32_t c = (uint32_t)nd.size();
            orc_node l, r;
            std::memset(&l, 0, sizeof l);
            std::memset(&r, 0, sizeof r);
            l.a = nd[n].a;
            l.b = mid;
            r.a = mid;
            r.b = nd[n].b;
            l.parent = r.parent = (uint32_t)n;
            nd[n].a = c;
            nd[n].b = c + 1;
            nd.push_back(l);
            nd.push_back(r);
            lvl.push_back(level + 1);
            lvl.push_back(level + 1);
            lvl[n] = 0;
            OC[it] += 2;
        }
        begin = end;
        end = nd.size();
    }
    for (size_t n = 0; n < nd.size(); ++n) nd[n].leaf = (int8_t)std::min(lvl[n], 127);
    uint32_t N = (uint32_t)nd.size();
    if (node_count) *node_count = N;
    if (level_counts)
        for (int i = 0; i < 64; ++i) level_counts[i] = OC[i];
    if (N > node_cap || !nodes) return nodes ? -1 : 0;

    // build_leaf_bbox (src/kernel/bvh-builder.cl:262-288), then
    // build_node_bbox by depth, deepest first (src/rt/bvh-builder.cpp:683-695).
    std::vector<BBox> nb(N);
    for (uint32_t n = 0; n < N; ++n) {
        if (!nd[n].leaf) continue;
        BBox b = pb[p[nd[n].a]];
        for (uint32_t i = nd[n].a + 1; i < nd[n].b; ++i) merge(b, pb[p[i]]);
        nb[n] = b;
    }
    // level blocks: [0,1), then OC-sized blocks
    std::vector<size_t> lb(1, 0);
    size_t acc = 1;
    lb.push_back(acc);
    for (int it = 0; it < depth; ++it) {
        acc += OC[it];
        lb.push_back(acc);
    }
    for (int d = depth - 1; d >= 0; --d) {
        for (size_t n = lb[d]; n < lb[d + 1]; ++n) {
            if (nd[n].leaf) continue;
            BBox b = nb[nd[n].a];
            merge(b, nb[nd[n].b]);
            nb[n] = b;
        }
    }
    for (uint32_t n = 0; n < N; ++n) {
        orc_node o = nd[n];
        o.hi[0] = nb[n].hi.x; o.hi[1] = nb[n].hi.y; o.hi[2] = nb[n].hi.z; o.hi[3] = 0.f;
        o.lo[0] = nb[n].lo.x; o.lo[1] = nb[n].lo.y; o.lo[2] = nb[n].lo.z; o.lo[3] = 0.f;
        nodes[n] = o;
    }
    return 0;
}

// refit_lbvh (src/rt/bvh-builder.cpp:743-850): prim AABBs recomputed from the
// already-permuted index buffer with an identity map (Q14), topology kept.
int orc_lbvh_refit(const float* pos, const uint32_t* idx, uint32_t T, orc_node* nodes, uint32_t N, int depth,
                   const uint32_t* level_counts) {
    if (!nodes || !N) return -1;
    depth = std::min(std::max(depth, 1), 64);
    std::vector<BBox> nb(N);
    for (uint32_t n = 0; n < N; ++n) {
        if (!nodes[n].leaf) continue;
        BBox b = prim_bbox(pos, idx, nodes[n].a);
        for (uint32_t i = nodes[n].a + 1; i < nodes[n].b; ++i) merge(b, prim_bbox(pos, idx, i));
        nb[n] = b;
    }
    (void)T;
    std::vector<size_t> lb(1, 0);
    size_t acc = 1;
    lb.push_back(acc);
    for (int it = 0; it < depth; ++it) {
        acc += level_counts[it];
        lb.push_back(acc);
    }
    for (int d = depth - 1; d >= 0; --d)
        for (size_t n = lb[d]; n < lb[d + 1] && n < N; ++n) {
            if (nodes[n].leaf) continue;
            BBox b = nb[nodes[n].a];
            merge(b, nb[nodes[n].b]);
            nb[n] = b;
        }
    for (uint32_t n = 0; n < N; ++n) {
        nodes[n].hi[0] = nb[n].hi.x; nodes[n].hi[1] = nb[n].hi.y; nodes[n].hi[2] = nb[n].hi.z;
        nodes[n].lo[0] = nb[n].lo.x; nodes[n].lo[1] = nb[n].lo.y; nodes[n].lo[2] = nb[n].lo.z;
    }
    return 0;
}

// ------------------------------------------------------------------ a9 -----
// Camera::set (src/rt/camera.cpp:31-59), left-handed unless right_handed.
void orc_camera_set(const float pos[3], const float dir[3], const float up_in[3], float fov, float aspect,
                    int right_handed, orc_camera* cam) {
    f3 d = normalize(ld3(dir));
    f3 U = ld3(up_in);
    f3 right = right_handed ? cross(d, U) : cross(U, d);
    f3 up = right_handed ? cross(right, d) : cross(d, right);
    float plane_h = (float)std::tan((double)(fov * 0.5f));
    float rn = std::sqrt(dot(right, right));
    float un = std::sqrt(dot(up, up));
    right = right * ((plane_h * aspect) / rn);
    up = up * (plane_h / un);
    for (int k = 0; k < 3; ++k) cam->pos[k] = pos[k];
    cam->dir[0] = d.x; cam->dir[1] = d.y; cam->dir[2] = d.z;
    cam->right[0] = right.x; cam->right[1] = right.y; cam->right[2] = right.z;
    cam->up[0] = up.x; cam->up[1] = up.y; cam->up[2] = up.z;
}

// Sample -> pixel mapping of generate_primary_rays
// (src/kernel/primary-ray-generator.cl:44-84).
static inline void sample_xy(int i, int width, int height, int quad, int* xo, int* yo) {
    int block_side = quad, block_area = quad * quad;
    int x_blocks = width / block_side, y_blocks = height / block_side;
    int block_rays = x_blocks * y_blocks * block_area;
    int right_rays = (width - x_blocks * block_side) * height;
    int x, y;
    if (i < block_rays) {
        int block = i / block_area;
        int y_block = block / x_blocks, x_block = block % x_blocks;
        int block_pos = i % block_area;
        x = x_block * block_side + block_pos % block_side;
        y = y_block * block_side + block_pos / block_side;
    } else if (i < block_rays + right_rays) {
        int start_x = x_blocks * block_side;
        int slab = width - start_x;
        int ri = i - block_rays;
        x = start_x + ri % slab;
        y = ri / slab;
    } else {
        int start_y = y_blocks * block_side;
        int slab = height - start_y;
        int ti = i - (block_rays + right_rays);
        x = ti / slab;
        y = start_y + ti % slab;
    }
    *xo = x;
    *yo = y;
}

// generate_primary_rays (src/kernel/primary-ray-generator.cl:28-103).
// Rays are written SoA: ori4 = (ori, tMin), dir4 = (dir, tMax).
// psamples: spp triples (ox, oy, contribution).
void orc_raygen_primary(const orc_camera* cam, int width, int height, int quad, const float* psamples,
                        uint64_t offset, uint64_t count, float* ori4, float* dir4, int32_t* pixel, float* contrib) {
    int pixels = width * height;
    f3 dir = ld3(cam->dir), right = ld3(cam->right), up = ld3(cam->up);
    for (uint64_t k = 0; k < count; ++k) {
        int gid = (int)(offset + k);
        int i = gid % pixels;
        const float* ps = psamples + 3 * (gid / pixels);
        int x, y;
        sample_xy(i, width, height, quad, &x, &y);
        float xn = (0.5f + (float)x + ps[0]) / (float)width;
        float yn = (0.5f + (float)y + ps[1]) / (float)height;
        f3 d = normalize(dir + right * (xn * 2.f - 1.f) + up * (yn * 2.f - 1.f));
        ori4[4 * k + 0] = cam->pos[0]; ori4[4 * k + 1] = cam->pos[1]; ori4[4 * k + 2] = cam->pos[2];
        ori4[4 * k + 3] = 0.f;
        dir4[4 * k + 0] = d.x; dir4[4 * k + 1] = d.y; dir4[4 * k + 2] = d.z;
        dir4[4 * k + 3] = 1e36f;
        pixel[k] = width * y + x;
        contrib[k] = ps[2];
    }
}

}  // extern "C"

// ---------------------------------------------------------------- a10/a11 --
namespace {

struct Ray {
    f3 ori, dir, inv;
    float tmin, tmax;
};

// bbox_hit: src/kernel/trace-bvh.cl:247-273 (shadow-trace-bvh.cl:160-186 uses
// min/max, defined here as the same IEEE minNum/maxNum).
inline bool bbox_hit(const orc_node& n, const Ray& r, float tmax) {
    float t0 = r.tmin, t1 = tmax;
    f3 lo = mk(n.lo[0], n.lo[1], n.lo[2]), hi = mk(n.hi[0], n.hi[1], n.hi[2]);
    f3 tl = (lo - r.ori) * r.inv;
    f3 th = (hi - r.ori) * r.inv;
    f3 tM = fmax3(tl, th);
    f3 tm = fmin3(tl, th);
    if (std::fabs(r.inv.x) > 1e-6f) { t0 = std::fmax(t0, tm.x); t1 = std::fmin(t1, tM.x); }
    if (std::fabs(r.inv.y) > 1e-6f) { t0 = std::fmax(t0, tm.y); t1 = std::fmin(t1, tM.y); }
    if (std::fabs(r.inv.z) > 1e-6f) { t0 = std::fmax(t0, tm.z); t1 = std::fmin(t1, tM.z); }
    return t0 <= t1;
}

// Möller–Trumbore as in leaf_hit (src/kernel/trace-bvh.cl:282-327).
// Returns false on reject; fills t, u, v.
inline bool tri_hit(const float* pos, const uint32_t* idx, uint32_t tri, const Ray& r, float* t, float* u, float* v) {
    f3 v0 = ld3(pos + 4 * (size_t)idx[3 * (size_t)tri + 0]);
    f3 v1 = ld3(pos + 4 * (size_t)idx[3 * (size_t)tri + 1]);
    f3 v2 = ld3(pos + 4 * (size_t)idx[3 * (size_t)tri + 2]);
    f3 e1 = v1 - v0, e2 = v2 - v0;
    f3 h = cross(r.dir, e2);
    float a = dot(e1, h);
    if (a > -1e-26f && a < 1e-26f) return false;
    float f = 1.f / a;
    f3 s = r.ori - v0;
    float uu = f * dot(s, h);
    if (uu < 0.f || uu > 1.f) return false;
    f3 q = cross(s, e1);
    float vv = f * dot(r.dir, q);
    if (vv < 0.f || uu + vv > 1.f) return false;
    *t = f * dot(e2, q);
    *u = uu;
    *v = vv;
    return true;
}

struct Hit {
    int id;
    float t, u, v;
};

// The stackless walks of src/kernel/trace-bvh-no-stack.cl:408-520 and
// src/kernel/shadow-trace-bvh-no-stack.cl:213-311, selected for every trace
// of the oracle by orc_set_traversal(1): the tree is walked by parent
// pointers (no stack, so no MAX_LEVELS), the near child first as in the
// stack walk.  Closest hit: per leaf the hits with t <= the leaf's running
// best (starting at the ray's tMax), merged into the ray's best only when
// strictly nearer (merge_hit_info, :210-219), so across leaves the first of
// equal hits is kept.  Going up from the root ends the walk.
static int g_stackless = 0;

static Hit trace_ray_ns(const Ray& r0, const float* pos, const uint32_t* idx, const orc_node* nodes,
                        uint32_t root) {
    Ray r = r0;
    Hit best{-1, r.tmax, 0.f, 0.f};
    bool up = false;
    uint32_t last = root, curr = root, first = 0, second = 0;
    for (;;) {
        const orc_node& n = nodes[curr];
        if (!n.leaf) {
            const orc_node& L = nodes[n.a];
            const orc_node& R = nodes[n.b];
            f3 delta = mk(R.lo[0], R.lo[1], R.lo[2]) - mk(L.lo[0], L.lo[1], L.lo[2]);
            bool order = dot(r.dir, delta) > 0.f;
            first = order ? n.a : n.b;
            second = order ? n.b : n.a;
        }
        const uint32_t parent = curr == root ? root : n.parent;
        if (up) {
            if (last == root) break;
            if (last == first) {
                last = curr;
                curr = second;
                up = false;
            } else {
                last = curr;
                curr = parent;
            }
        } else if (bbox_hit(n, r, r.tmax)) {
            if (n.leaf) {
                Hit info{-1, r.tmax, 0.f, 0.f};
                for (uint32_t i = n.a; i < n.b; ++i) {
                    float t, u, v;
                    if (!tri_hit(pos, idx, i, r, &t, &u, &v)) continue;
                    if (t <= info.t && t >= r.tmin) info = Hit{(int)i, t, u, v};
                }
                if (best.id < 0 || (info.id >= 0 && info.t < best.t)) best = info;
                if (best.id >= 0) r.tmax = best.t;
                last = curr;
                curr = parent;
                up = true;
            } else {
                last = curr;
                curr = first;
            }
        } else {
            last = curr;
            curr = parent;
            up = true;
        }
    }
    return best;
}

static bool trace_shadow_ns(const Ray& r, const float* pos, const uint32_t* idx, const orc_node* nodes,
                            uint32_t root) {
    bool up = false;
    uint32_t last = root, curr = root, first = 0, second = 0;
    for (;;) {
        const orc_node& n = nodes[curr];
        if (!n.leaf) {
            const orc_node& L = nodes[n.a];
            const orc_node& R = nodes[n.b];
            f3 delta = mk(R.lo[0], R.lo[1], R.lo[2]) - mk(L.lo[0], L.lo[1], L.lo[2]);
            bool order = dot(r.dir, delta) > 0.f;
            first = order ? n.a : n.b;
            second = order ? n.b : n.a;
        }
        const uint32_t parent = curr == root ? root : n.parent;
        if (up) {
            if (last == root) return false;
            if (last == first) {
                last = curr;
                curr = second;
                up = false;
            } else {
                last = curr;
                curr = parent;
            }
        } else if (bbox_hit(n, r, r.tmax)) {
            if (n.leaf) {
                for (uint32_t i = n.a; i < n.b; ++i) {
                    float t, u, v;
                    if (tri_hit(pos, idx, i, r, &t, &u, &v) && t < r.tmax && t > r.tmin) return true;
                }
                last = curr;
                curr = parent;
                up = true;
            } else {
                last = curr;
                curr = first;
            }
        } else {
            last = curr;
            curr = parent;
            up = true;
        }
    }
}

// trace_ray: src/kernel/trace-bvh.cl:330-417.  Stack of 32, near child first
// by dot(dir, r.lo - l.lo) > 0 (the adir branches all test that one scalar).
Hit trace_ray(const Ray& r, const float* pos, const uint32_t* idx, const orc_node* nodes, uint64_t* nvisit,
              uint32_t root = 0) {
    // nvisit (optional): [0] nodes visited, [1] triangles tested; root: bvh_root
    if (g_stackless) return trace_ray_ns(r, pos, idx, nodes, root);
    Hit best{-1, r.tmax, 0.f, 0.f};
    float tmax = r.tmax;
    uint32_t stack[32];
    uint32_t level = 0, curr = root;
    for (;;) {
        const orc_node& n = nodes[curr];
        if (nvisit) ++*nvisit;
        bool pop;
        if (!bbox_hit(n, r, tmax)) {
            pop = true;
        } else if (n.leaf) {
            if (nvisit) nvisit[1] += n.b - n.a;
            for (uint32_t i = n.a; i < n.b; ++i) {
                float t, u, v;
                if (!tri_hit(pos, idx, i, r, &t, &u, &v)) continue;
                if (t <= best.t && t >= r.tmin) best = Hit{(int)i, t, u, v};
            }
            if (best.id >= 0) tmax = best.t;
            pop = true;
        } else {
            const orc_node& L = nodes[n.a];
            const orc_node& R = nodes[n.b];
            f3 delta = mk(R.lo[0], R.lo[1], R.lo[2]) - mk(L.lo[0], L.lo[1], L.lo[2]);
            bool order = dot(r.dir, delta) > 0.f;
            uint32_t first = order ? n.a : n.b, second = order ? n.b : n.a;
            // trees deeper than the stack: the reference writes past its
            // 32 entries there (undefined); an entry at 32 or above is never
            // read, since the ray ends at its next pop, so it is not stored
            if (level < 32) stack[level] = second;
            level++;
            curr = first;
            pop = false;
        }
        if (pop) {
            if (level > 0 && level < 32) {
                level--;
                curr = stack[level];
            } else {
                return best;
            }
        }
    }
}

// trace_shadow_ray: src/kernel/shadow-trace-bvh.cl:238-315 with leaf_hit_any
// (:188-234): any hit with tMin < t < tMax.
bool trace_shadow(const Ray& r, const float* pos, const uint32_t* idx, const orc_node* nodes, uint64_t* nvisit,
                  uint32_t root = 0) {
    if (g_stackless) return trace_shadow_ns(r, pos, idx, nodes, root);
    uint32_t stack[32];
    uint32_t level = 0, curr = root;
    for (;;) {
        const orc_node& n = nodes[curr];
        bool pop;
        if (nvisit) ++nvisit[0];
        if (!bbox_hit(n, r, r.tmax)) {
            pop = true;
        } else if (n.leaf) {
            for (uint32_t i = n.a; i < n.b; ++i) {
                float t, u, v;
                if (nvisit) ++nvisit[1];
                if (!tri_hit(pos, idx, i, r, &t, &u, &v)) continue;
                if (t < r.tmax && t > r.tmin) return true;
            }
            pop = true;
        } else {
            const orc_node& L = nodes[n.a];
            const orc_node& R = nodes[n.b];
            f3 delta = mk(R.lo[0], R.lo[1], R.lo[2]) - mk(L.lo[0], L.lo[1], L.lo[2]);
            bool order = dot(r.dir, delta) > 0.f;
            uint32_t first = order ? n.a : n.b, second = order ? n.b : n.a;
            // trees deeper than the stack: the reference writes past its
            // 32 entries there (undefined); an entry at 32 or above is never
            // read, since the ray ends at its next pop, so it is not stored
            if (level < 32) stack[level] = second;
            level++;
            curr = first;
            pop = false;
        }
        if (pop) {
            if (level > 0 && level < 32) {
                level--;
                curr = stack[level];
            } else {
                return false;
            }
        }
    }
}

inline Ray load_ray(const float* ori4, const float* dir4, uint64_t k) {
    Ray r;
    r.ori = ld3(ori4 + 4 * k);
    r.tmin = ori4[4 * k + 3];
    r.dir = ld3(dir4 + 4 * k);
    r.tmax = dir4[4 * k + 3];
    r.inv = recip(r.dir);
    return r;
}

}  // namespace

extern "C" {

// Hit record bits (flags): bit0 hit, bit1 inverse_n, bits 8..15 shadowed by
// light l.  SampleTraceInfo of src/kernel/trace-bvh.cl:62-74 stored SoA.
enum { ORC_HIT = 1u, ORC_INV_N = 2u };

// trace_single + complete_trace_info (src/kernel/trace-bvh.cl:190-245, 465-481).
// nrm: raw vertex normals (V*4), uv: V*2.  Outputs per ray: flags, t, id,
// u, v (barycentrics), n4, p4 (hit point), uv2.  visits (optional) counts
// node visits for effective-bandwidth reporting.
void orc_trace(const float* pos, const float* nrm, const float* uvs, const uint32_t* idx, const orc_node* nodes,
               const float* ori4, const float* dir4, uint64_t count, uint32_t* flags, float* t, int32_t* id,
               float* bary2, float* n4, float* p4, float* uv2, uint64_t* visits) {
    for (uint64_t k = 0; k < count; ++k) {
        Ray r = load_ray(ori4, dir4, k);
        Hit h = trace_ray(r, pos, idx, nodes, visits);
        if (h.id < 0) {
            flags[k] = 0;
            t[k] = 0.f;
            id[k] = -1;
            if (bary2) { bary2[2 * k] = 0.f; bary2[2 * k + 1] = 0.f; }
            for (int c = 0; c < 4; ++c) { n4[4 * k + c] = 0.f; p4[4 * k + c] = 0.f; }
            uv2[2 * k] = uv2[2 * k + 1] = 0.f;
            continue;
        }
        f3 hp = r.ori + r.dir * h.t;
        uint32_t i0 = idx[3 * (size_t)h.id], i1 = idx[3 * (size_t)h.id + 1], i2 = idx[3 * (size_t)h.id + 2];
        float u = h.u, v = h.v, w = 1.f - (u + v);
        f3 n0 = normalize(ld3(nrm + 4 * (size_t)i0));
        f3 n1 = normalize(ld3(nrm + 4 * (size_t)i1));
        f3 n2 = normalize(ld3(nrm + 4 * (size_t)i2));
        f3 n = normalize(w * n0 + v * n2 + u * n1);
        uint32_t fl = ORC_HIT;
        if (dot(n, r.dir) > 0.f) {
            fl |= ORC_INV_N;
            n = -n;
        }
        float tu0 = uvs[2 * (size_t)i0], tv0 = uvs[2 * (size_t)i0 + 1];
        float tu1 = uvs[2 * (size_t)i1], tv1 = uvs[2 * (size_t)i1 + 1];
        float tu2 = uvs[2 * (size_t)i2], tv2 = uvs[2 * (size_t)i2 + 1];
        flags[k] = fl;
        t[k] = h.t;
        id[k] = h.id;
        if (bary2) { bary2[2 * k] = u; bary2[2 * k + 1] = v; }
        n4[4 * k] = n.x; n4[4 * k + 1] = n.y; n4[4 * k + 2] = n.z; n4[4 * k + 3] = 0.f;
        p4[4 * k] = hp.x; p4[4 * k + 1] = hp.y; p4[4 * k + 2] = hp.z; p4[4 * k + 3] = 0.f;
        uv2[2 * k] = w * tu0 + v * tu2 + u * tu1;
        uv2[2 * k + 1] = w * tv0 + v * tv2 + u * tv1;
    }
}

// Brute-force closest hit over every triangle (no BVH): the textbook check
// that pins the traversal.  Same Möller–Trumbore and tie rule (later index
// wins an exact t tie).
void orc_trace_brute(const float* pos, const uint32_t* idx, uint32_t T, const float* ori4, const float* dir4,
                     uint64_t count, int32_t* id, float* t) {
    for (uint64_t k = 0; k < count; ++k) {
        Ray r = load_ray(ori4, dir4, k);
        Hit best{-1, r.tmax, 0.f, 0.f};
        for (uint32_t i = 0; i < T; ++i) {
            float tt, u, v;
            if (!tri_hit(pos, idx, i, r, &tt, &u, &v)) continue;
            if (tt <= best.t && tt >= r.tmin) best = Hit{(int)i, tt, u, v};
        }
        id[k] = best.id;
        t[k] = best.id >= 0 ? best.t : 0.f;
    }
}

// Shadow ray of light l from a hit point (shadow_trace_single,
// src/kernel/shadow-trace-bvh.cl:375-418; POINT is the configs' extension:
// toward pos, tMax = distance).  Returns 1 = shadowed, 0 = lit, and the ray
// (for tests) when out pointers are given.
static int shadow_ray_for(const orc_light& L, f3 p, Ray* out) {
    Ray r;
    r.ori = p;
    r.tmin = 0.01f;
    r.tmax = 1e37f;
    if (L.type == ORC_DIR) {
        r.dir = -ld3(L.dir);
    } else if (L.type == ORC_SPOT) {
        r.dir = normalize(ld3(L.pos) - p);
        if (dot(-r.dir, ld3(L.dir)) < L.angle) return 1;
    } else if (L.type == ORC_POINT) {
        f3 toL = ld3(L.pos) - p;
        r.tmax = std::sqrt(dot(toL, toL));
        r.dir = normalize(toL);
    } else {
        return 1;
    }
    r.inv = recip(r.dir);
    *out = r;
    return -1;
}

void orc_trace_shadow(const float* pos, const uint32_t* idx, const orc_node* nodes, const orc_lights* lights,
                      uint64_t count, uint32_t* flags, const float* p4, uint64_t* visits) {
    for (uint64_t k = 0; k < count; ++k) {
        uint32_t fl = flags[k] & 0xffu;
        if (fl & ORC_HIT) {
            f3 p = ld3(p4 + 4 * k);
            for (int l = 0; l < lights->count; ++l) {
                Ray r;
                int s = shadow_ray_for(lights->l[l], p, &r);
                if (s < 0) s = trace_shadow(r, pos, idx, nodes, visits) ? 1 : 0;
                if (s) fl |= 1u << (8 + l);
            }
        }
        flags[k] = fl;
    }
}

// Brute-force any hit (pins the shadow traversal).
void orc_trace_shadow_brute(const float* pos, const uint32_t* idx, uint32_t T, const orc_lights* lights,
                            uint64_t count, uint32_t* flags, const float* p4) {
    for (uint64_t k = 0; k < count; ++k) {
        uint32_t fl = flags[k] & 0xffu;
        if (fl & ORC_HIT) {
            f3 p = ld3(p4 + 4 * k);
            for (int l = 0; l < lights->count; ++l) {
                Ray r;
                int s = shadow_ray_for(lights->l[l], p, &r);
                if (s < 0) {
                    s = 0;
                    for (uint32_t i = 0; i < T && !s; ++i) {
                        float t, u, v;
                        if (tri_hit(pos, idx, i, r, &t, &u, &v) && t < r.tmax && t > r.tmin) s = 1;
                    }
                }
                if (s) fl |= 1u << (8 + l);
            }
        }
        flags[k] = fl;
    }
}

// ------------------------------------------------------------------ a12 ----
// mark_secondary_rays + exclusive scan (total at [n]) + generate_secondary_rays
// (src/kernel/secondary-ray-generator.cl:43-181, src/gpu/scan.cpp:17-157).
// A parent whose children would reach the capacity (new_id + 2 >= cap) emits
// none (:61-63); the returned count is the number of rays actually written
// (the written rays form a prefix), where the reference would report the scan
// total and trace unwritten slots.
// Returns a mask: bit0 = refracted child, bit1 = reflected child.
static int sec_children(const orc_material& m, f3 dir, f3 n_in, uint32_t fl, float* R_out, f3* refr_dir,
                        f3* refl_dir) {
    f3 d = -normalize(dir);
    f3 n = normalize(n_in);
    float R = 1.f;
    int c = 0;
    if (m.refractive_index > 0.f) {
        float etai, etat;
        if (fl & ORC_INV_N) {
            etat = 1.f;
            etai = m.refractive_index;
        } else {
            etai = 1.f;
            etat = m.refractive_index;
        }
        float cosi = dot(d, n);
        float rel = etai / etat;
        float w = rel * cosi;
        float k = 1.f + (w - rel) * (w + rel);
        if (k > 0.f) {
            float R0 = (etat - 1.f) / (etat + 1.f);
            R0 *= R0;
            float x = 1.f - cosi;
            float x2 = x * x;
            float cos5 = x2 * x2 * x;  // pow(x, 5) as repeated products
            R = R0 + (1.f - R0) * cos5;
            k = std::sqrt(k);
            if (refr_dir) *refr_dir = normalize((w - k) * n - etai * d);
            c |= 1;
        }
    }
    if (m.reflectiveness > 0.f) {
        f3 dd = -d;
        if (refl_dir) *refl_dir = normalize(dd - 2.f * n * dot(dd, n));
        c |= 2;
    }
    *R_out = R;
    return c;
}

uint64_t orc_raygen_secondary(const orc_material* mats, const int32_t* matmap, const float* ori4_in,
                              const float* dir4_in, const int32_t* pixel_in, const float* contrib_in,
                              const uint32_t* flags, const int32_t* id, const float* n4, const float* p4,
                              uint64_t n_in, uint64_t cap, float* ori4, float* dir4, int32_t* pixel, float* contrib,
                              uint64_t* total_out) {
    std::vector<uint64_t> cnt(n_in + 1, 0);
    for (uint64_t k = 0; k < n_in; ++k) {
        if (!(flags[k] & ORC_HIT)) continue;
        float R;
        int c = sec_children(mats[matmap[id[k]]], ld3(dir4_in + 4 * k), ld3(n4 + 4 * k), flags[k], &R, nullptr,
                             nullptr);
        cnt[k] = (uint64_t)((c & 1) + (c >> 1));
    }
    uint64_t acc = 0;
    for (uint64_t k = 0; k <= n_in; ++k) {
        uint64_t c = cnt[k];
        cnt[k] = acc;
        acc += c;
    }
    uint64_t total = cnt[n_in];
    if (total_out) *total_out = total;
    uint64_t written = total;
    for (uint64_t k = 0; k < n_in; ++k) {
        if (!(flags[k] & ORC_HIT)) continue;
        uint64_t nid = cnt[k];
        if (cnt[k + 1] == nid) continue;
        if (nid + 2 >= cap) {
            written = std::min(written, nid);
            continue;
        }
        const orc_material& m = mats[matmap[id[k]]];
        float R;
        f3 rd = mk(0, 0, 0), fd = mk(0, 0, 0);
        f3 dirv = ld3(dir4_in + 4 * k);
        int c = sec_children(m, dirv, ld3(n4 + 4 * k), flags[k], &R, &rd, &fd);
        f3 hp = ld3(p4 + 4 * k);
        auto emit = [&](f3 d, float contribution) {
            ori4[4 * nid] = hp.x; ori4[4 * nid + 1] = hp.y; ori4[4 * nid + 2] = hp.z; ori4[4 * nid + 3] = 0.0001f;
            dir4[4 * nid] = d.x; dir4[4 * nid + 1] = d.y; dir4[4 * nid + 2] = d.z; dir4[4 * nid + 3] = 1e37f;
            pixel[nid] = pixel_in[k];
            contrib[nid] = contribution;
            ++nid;
        };
        // refraction first, then reflection (generate_secondary_rays).
        if (c & 1) emit(rd, contrib_in[k] * m.reflectiveness * (1.f - R));
        if (c & 2) emit(fd, contrib_in[k] * m.reflectiveness * R);
    }
    return written;
}

// ------------------------------------------------------------ a13 / a14 ----
// Textures of the shading stage: RGBA8 texels of every image in one array;
// entry e of `tab` = {first texel, width, height}.  Entries [0, count) are
// the material textures (TextureAtlas, src/rt/texture-atlas.cpp:25-60; texel
// (x, y) at (x + y*width), as create_tex_gl_from_file lays it out,
// src/cl-gl/opengl-init.cpp:143-200); with `cube`, entries count..count+5
// are the cubemap faces +x, -x, +y, -y, +z, -z (Cubemap::initialize,
// src/rt/cubemap.cpp:4-57).
struct orc_texset {
    const uint8_t* rgba;
    const int64_t* tab;
    int32_t count;
    int32_t cube;
};

// UNORM_INT8 texel to float: c / 255.
static f3 texel(const orc_texset* tx, int e, int x, int y) {
    const int64_t off = tx->tab[3 * e], w = tx->tab[3 * e + 1];
    const uint8_t* p = tx->rgba + 4 * (off + x + (int64_t)y * w);
    return mk(p[0] / 255.f, p[1] / 255.f, p[2] / 255.f);
}

// read_imagef with CLK_NORMALIZED_COORDS_TRUE | CLK_FILTER_LINEAR and
// CLK_ADDRESS_REPEAT (material textures) or CLK_ADDRESS_CLAMP_TO_EDGE (the
// cubemap), ray-shader.cl:145-153: the OpenCL 1.2 formulas (section 8.2),
//   repeat: u = (s - floor(s)) * w; i0 = floor(u - 0.5), i1 = i0 + 1, each
//           wrapped into [0, w); clamp: u = s * w, i0, i1 clamped to [0, w);
//   a = frac(u - 0.5) (likewise v, j0, j1, b);
//   T = (1-a)(1-b) T00 + a(1-b) T10 + (1-a)b T01 + ab T11,
// evaluated in float32 in that order.  OpenCL leaves the weights' precision
// to the implementation; this is the definition both the oracle and the
// kernel follow.  A non-finite coordinate samples as 0.
static f3 sample_linear(const orc_texset* tx, int e, float s, float t, bool repeat) {
    const int w = (int)tx->tab[3 * e + 1], h = (int)tx->tab[3 * e + 2];
    if (!std::isfinite(s)) s = 0.f;
    if (!std::isfinite(t)) t = 0.f;
    float u, v;
    if (repeat) {
        u = (s - std::floor(s)) * (float)w;
        v = (t - std::floor(t)) * (float)h;
    } else {
        u = s * (float)w;
        v = t * (float)h;
    }
    const float fu = std::floor(u - 0.5f), fv = std::floor(v - 0.5f);
    const float a = (u - 0.5f) - fu, b = (v - 0.5f) - fv;
    int i0, i1, j0, j1;
    if (repeat) {
        i0 = (int)fu;
        j0 = (int)fv;
        i1 = i0 + 1;
        j1 = j0 + 1;
        if (i0 < 0) i0 += w;
        if (j0 < 0) j0 += h;
        if (i1 > w - 1) i1 -= w;
        if (j1 > h - 1) j1 -= h;
    } else {
        const int iu = (int)std::fmin(std::fmax(fu, -1.f), (float)w);
        const int iv = (int)std::fmin(std::fmax(fv, -1.f), (float)h);
        i0 = std::min(std::max(iu, 0), w - 1);
        i1 = std::min(std::max(iu + 1, 0), w - 1);
        j0 = std::min(std::max(iv, 0), h - 1);
        j1 = std::min(std::max(iv + 1, 0), h - 1);
    }
    const float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
    return ((w00 * texel(tx, e, i0, j0) + w10 * texel(tx, e, i1, j0)) + w01 * texel(tx, e, i0, j1)) +
           w11 * texel(tx, e, i1, j1);
}

static bool in_range(float f1, float f2) { return f1 <= 1.f && f1 >= -1.f && f2 <= 1.f && f2 >= -1.f; }

// The miss branch of shade_sample (ray-shader.cl:275-307): the first face
// whose projection of the direction lies in [-1, 1]^2, in the order
// +x, -x, +y, -y, +z, -z; no face (a NaN direction) gives black.
static f3 cubemap_rgb(const orc_texset* tx, f3 d) {
    const f3 inv = recip(d);
    const f3 xp = d * inv.x, yp = d * inv.y, zp = d * inv.z;
    const int c = tx->count;
    if (d.x > 0.f && in_range(xp.y, xp.z)) return sample_linear(tx, c + 0, (-xp.z + 1.f) * 0.5f, (xp.y + 1.f) * 0.5f, false);
    if (d.x < 0.f && in_range(xp.y, xp.z)) return sample_linear(tx, c + 1, (-xp.z + 1.f) * 0.5f, (-xp.y + 1.f) * 0.5f, false);
    if (d.y > 0.f && in_range(yp.x, yp.z)) return sample_linear(tx, c + 2, (yp.x + 1.f) * 0.5f, (-yp.z + 1.f) * 0.5f, false);
    if (d.y < 0.f && in_range(yp.x, yp.z)) return sample_linear(tx, c + 3, (-yp.x + 1.f) * 0.5f, (-yp.z + 1.f) * 0.5f, false);
    if (d.z > 0.f && in_range(zp.x, zp.y)) return sample_linear(tx, c + 4, (zp.x + 1.f) * 0.5f, (zp.y + 1.f) * 0.5f, false);
    if (d.z < 0.f && in_range(zp.x, zp.y)) return sample_linear(tx, c + 5, (zp.x + 1.f) * 0.5f, (-zp.y + 1.f) * 0.5f, false);
    return mk(0.f, 0.f, 0.f);
}

// shade_sample (src/kernel/ray-shader.cl:116-333): ambient*diffuse + per
// unshadowed light (diffuse*cos^2 + Phong^8), times (1 - reflectiveness),
// contribution-scaled, clamped, floor(x * 1e6) added to the pixel.  The
// diffuse colour is the material's, or its texture sampled at the hit's uv
// (an index without a texture samples black).  accum: uint32 r,g,b per pixel.
static void shade_one(const orc_material& m, const orc_lights* lights, uint32_t fl, f3 d, f3 n, f3 p,
                      float contribution, float out[3], const orc_texset* tx, float tu, float tv) {
    f3 diffuse = mk(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
    if (m.texture >= 0)
        diffuse = (tx && m.texture < tx->count) ? sample_linear(tx, m.texture, tu, tv, true) : mk(0.f, 0.f, 0.f);
    f3 val = ld3(lights->ambient) * diffuse;
    for (int l = 0; l < lights->count; ++l) {
        const orc_light& L = lights->l[l];
        f3 Ld;
        float LD;
        if (L.type == ORC_DIR) {
            Ld = ld3(L.dir);
            LD = 1.f;
        } else if (L.type == ORC_SPOT) {
            Ld = normalize(p - ld3(L.pos));
            LD = dot(Ld, ld3(L.dir)) - L.angle;
            LD /= 1.f - L.angle;
            LD = clampf(LD, 0.f, 1.f);
        } else {
            Ld = normalize(p - ld3(L.pos));
            LD = 1.f;
        }
        f3 rgb = ld3(L.color) * LD;
        if (fl & (1u << (8 + l))) continue;
        f3 r = d - 2.f * n * dot(d, n);
        float cosL = clampf(dot(n, -Ld), 0.f, 1.f);
        cosL = cosL * cosL;
        float spec = clampf(dot(r, -Ld), 0.f, 1.f);
        float s2 = spec * spec, s4 = s2 * s2;
        spec = s4 * s4;
        spec *= cosL * m.shininess;
        val = val + rgb * diffuse * cosL;
        val = val + rgb * spec;
    }
    val = val * (1.f - m.reflectiveness);
    out[0] = val.x * contribution;
    out[1] = val.y * contribution;
    out[2] = val.z * contribution;
}

// tx may be null (no textures, no cubemap); uv2 is read only for textured
// materials.  Without a cubemap a miss adds nothing (the `return` of the
// miss branch).
void orc_shade_tex(const orc_material* mats, const int32_t* matmap, const orc_lights* lights, const float* dir4,
                   const int32_t* pixel, const float* contrib, const uint32_t* flags, const int32_t* id,
                   const float* n4, const float* p4, const float* uv2, uint64_t count, const orc_texset* tx,
                   uint32_t* accum) {
    for (uint64_t k = 0; k < count; ++k) {
        float c[3];
        if (!(flags[k] & ORC_HIT)) {
            if (!tx || !tx->cube) continue;
            const f3 rgb = cubemap_rgb(tx, ld3(dir4 + 4 * k));
            c[0] = contrib[k] * rgb.x;
            c[1] = contrib[k] * rgb.y;
            c[2] = contrib[k] * rgb.z;
        } else {
            shade_one(mats[matmap[id[k]]], lights, flags[k], ld3(dir4 + 4 * k), ld3(n4 + 4 * k), ld3(p4 + 4 * k),
                      contrib[k], c, tx, uv2 ? uv2[2 * k] : 0.f, uv2 ? uv2[2 * k + 1] : 0.f);
        }
        for (int ch = 0; ch < 3; ++ch) {
            float f = clampf(c[ch], 0.f, 1.f);
            accum[3 * (size_t)pixel[k] + ch] += (uint32_t)std::floor(f * 1e6f);
        }
    }
}

void orc_shade(const orc_material* mats, const int32_t* matmap, const orc_lights* lights, const float* dir4,
               const int32_t* pixel, const float* contrib, const uint32_t* flags, const int32_t* id, const float* n4,
               const float* p4, uint64_t count, uint32_t* accum) {
    orc_shade_tex(mats, matmap, lights, dir4, pixel, contrib, flags, id, n4, p4, nullptr, count, nullptr, accum);
}

// One texture sample as the shading stage takes it (for the sampler's pins).
void orc_sample_texture(const orc_texset* tx, int32_t entry, const float* st, uint64_t n, int32_t repeat,
                        float* rgb) {
    for (uint64_t k = 0; k < n; ++k) {
        const f3 c = sample_linear(tx, entry, st[2 * k], st[2 * k + 1], repeat != 0);
        rgb[3 * k] = c.x;
        rgb[3 * k + 1] = c.y;
        rgb[3 * k + 2] = c.z;
    }
}

void orc_sample_cubemap(const orc_texset* tx, const float* dir4, uint64_t n, float* rgb) {
    for (uint64_t k = 0; k < n; ++k) {
        const f3 c = cubemap_rgb(tx, ld3(dir4 + 4 * k));
        rgb[3 * k] = c.x;
        rgb[3 * k + 1] = c.y;
        rgb[3 * k + 2] = c.z;
    }
}

// copy kernel (src/kernel/framebuffer.cl:31-54): float(acc) * (1/1e6),
// saturate, x255, round to nearest even -> RGBA8, alpha 255.
void orc_resolve(const uint32_t* accum, uint64_t pixels, uint8_t* rgba) {
    const float inv = 1.f / 1e6f;
    for (uint64_t i = 0; i < pixels; ++i) {
        for (int c = 0; c < 3; ++c) {
            float f = (float)accum[3 * i + c] * inv;
            f = clampf(f, 0.f, 1.f) * 255.f;
            rgba[4 * i + c] = (uint8_t)std::nearbyint(f);
        }
        rgba[4 * i + 3] = 255;
    }
}

// --------------------------------------------------------------- a15 -------
// Renderer::render_to_framebuffer (src/rt/renderer.cpp:113-214) for one tile
// covering the whole frame: primary -> trace -> shadow -> shade, then
// max_bounces rounds of secondary -> trace -> shadow -> shade.  The scene
// must already hold a built BVH (orc_lbvh_build).  Returns primary+secondary
// ray count (the reference's count); *shadow_rays counts shadow rays traced.
struct orc_frame_stats {
    uint64_t primary, secondary, shadow, bounces_run;
    uint64_t per_bounce[16];
};

uint64_t orc_render_tex(const float* pos, const float* nrm, const float* uvs, const uint32_t* idx,
                        const int32_t* matmap, const orc_node* nodes, const orc_material* mats,
                        const orc_lights* lights, const orc_camera* cam, int width, int height, int quad,
                        int max_bounces, uint64_t cap, const orc_texset* tx, uint32_t* accum, orc_frame_stats* st) {
    uint64_t P = (uint64_t)width * height;
    if (cap < P) cap = P;
    std::vector<float> o4a(4 * cap), d4a(4 * cap), o4b(4 * cap), d4b(4 * cap), ca(cap), cb(cap);
    std::vector<int32_t> pa(cap), pb(cap), id(cap);
    std::vector<uint32_t> fl(cap);
    std::vector<float> t(cap), n4(4 * cap), p4(4 * cap), uv2(2 * cap);
    const float ps[3] = {0.f, 0.f, 1.f};
    orc_raygen_primary(cam, width, height, quad, ps, 0, P, o4a.data(), d4a.data(), pa.data(), ca.data());
    std::memset(accum, 0, sizeof(uint32_t) * 3 * P);
    orc_frame_stats s;
    std::memset(&s, 0, sizeof s);
    s.primary = P;
    uint64_t n = P;
    float *o_in = o4a.data(), *d_in = d4a.data(), *c_in = ca.data();
    int32_t* p_in = pa.data();
    float *o_out = o4b.data(), *d_out = d4b.data(), *c_out = cb.data();
    int32_t* p_out = pb.data();
    auto count_shadow = [&](uint64_t nn) {
        for (uint64_t k = 0; k < nn; ++k)
            if (fl[k] & ORC_HIT) s.shadow += (uint64_t)lights->count;
    };
    orc_trace(pos, nrm, uvs, idx, nodes, o_in, d_in, n, fl.data(), t.data(), id.data(), nullptr, n4.data(), p4.data(),
              uv2.data(), nullptr);
    orc_trace_shadow(pos, idx, nodes, lights, n, fl.data(), p4.data(), nullptr);
    count_shadow(n);
    orc_shade_tex(mats, matmap, lights, d_in, p_in, c_in, fl.data(), id.data(), n4.data(), p4.data(), uv2.data(), n,
                  tx, accum);
    for (int b = 0; b < max_bounces; ++b) {
        uint64_t m = orc_raygen_secondary(mats, matmap, o_in, d_in, p_in, c_in, fl.data(), id.data(), n4.data(),
                                          p4.data(), n, cap, o_out, d_out, p_out, c_out, nullptr);
        std::swap(o_in, o_out);
        std::swap(d_in, d_out);
        std::swap(p_in, p_out);
        std::swap(c_in, c_out);
        n = m;
        if (!n) break;
        s.secondary += n;
        if (b < 16) s.per_bounce[b] = n;
        s.bounces_run++;
        orc_trace(pos, nrm, uvs, idx, nodes, o_in, d_in, n, fl.data(), t.data(), id.data(), nullptr, n4.data(),
                  p4.data(), uv2.data(), nullptr);
        orc_trace_shadow(pos, idx, nodes, lights, n, fl.data(), p4.data(), nullptr);
        count_shadow(n);
        orc_shade_tex(mats, matmap, lights, d_in, p_in, c_in, fl.data(), id.data(), n4.data(), p4.data(), uv2.data(), n,
                  tx, accum);
    }
    if (st) *st = s;
    return s.primary + s.secondary;
}

uint64_t orc_render(const float* pos, const float* nrm, const float* uvs, const uint32_t* idx, const int32_t* matmap,
                    const orc_node* nodes, const orc_material* mats, const orc_lights* lights, const orc_camera* cam,
                    int width, int height, int quad, int max_bounces, uint64_t cap, uint32_t* accum,
                    orc_frame_stats* st) {
    return orc_render_tex(pos, nrm, uvs, idx, matmap, nodes, mats, lights, cam, width, height, quad, max_bounces, cap,
                          nullptr, accum, st);
}

// ------------------------------------------------------------ FP64 shadow --
// Hit distance of ray k against its reported triangle, recomputed in double
// (Möller–Trumbore in double).  Writes t64; returns max |t32 - t64| / |t64|.
double orc_hit_t_f64(const float* pos, const uint32_t* idx, const float* ori4, const float* dir4, const int32_t* id,
                     const float* t32, uint64_t count, double* t64) {
    double worst = 0.0;
    for (uint64_t k = 0; k < count; ++k) {
        if (id[k] < 0) {
            if (t64) t64[k] = 0.0;
            continue;
        }
        double P[3], D[3], V0[3], V1[3], V2[3];
        for (int c = 0; c < 3; ++c) {
            P[c] = ori4[4 * k + c];
            D[c] = dir4[4 * k + c];
            V0[c] = pos[4 * (size_t)idx[3 * (size_t)id[k]] + c];
            V1[c] = pos[4 * (size_t)idx[3 * (size_t)id[k] + 1] + c];
            V2[c] = pos[4 * (size_t)idx[3 * (size_t)id[k] + 2] + c];
        }
        double e1[3], e2[3], s[3];
        for (int c = 0; c < 3; ++c) {
            e1[c] = V1[c] - V0[c];
            e2[c] = V2[c] - V0[c];
            s[c] = P[c] - V0[c];
        }
        double h[3] = {D[1] * e2[2] - D[2] * e2[1], D[2] * e2[0] - D[0] * e2[2], D[0] * e2[1] - D[1] * e2[0]};
        double a = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
        double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
        double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / a;
        if (t64) t64[k] = tt;
        double err = std::fabs((double)t32[k] - tt) / std::max(std::fabs(tt), 1e-30);
        worst = std::max(worst, err);
    }
    return worst;
}

// ------------------------------------------------------- multi-root -----
// BVHRoot (include/rt/multi-bvh.hpp:9-14): a root node of the shared node
// array and the object's transform tr (object to world) and its inverse,
// row-major 4x4 (sqmat4: four float4 rows).
struct orc_root {
    int32_t node;
    float tr[16];
    float tr_inv[16];
};

// OpenCL dot(float4, float4), defined as the sequential sum.
static inline float dot4(const float* a, const float* b) {
    return ((a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]) + a[3] * b[3];
}
// multiply_vector / multiply_point: the rows of M dotted with (v, 0) or
// (v, 1), then a division by w when |w| exceeds `wmin`: 1e-26 in
// src/kernel/trace-bvh.cl:83-113, 0.00001 in src/kernel/shadow-trace-bvh.cl:107-143.
static f3 mul_h(const float* M, f3 v, float vw, float wmin) {
    const float x[4] = {v.x, v.y, v.z, vw};
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = dot4(M + 4 * i, x);
    if (std::fabs(r[3]) > wmin) {
        r[0] = r[0] / r[3];
        r[1] = r[1] / r[3];
        r[2] = r[2] / r[3];
    }
    return mk(r[0], r[1], r[2]);
}
// transform_ray (src/kernel/trace-bvh.cl:115-121): direction and origin into
// the object's frame, tMin / tMax unchanged (an affine map keeps the ray
// parameter of every point).
static Ray transform_ray(const Ray& r, const float* M, float wmin) {
    Ray o = r;
    o.dir = mul_h(M, r.dir, 0.f, wmin);
    o.ori = mul_h(M, r.ori, 1.f, wmin);
    o.inv = recip(o.dir);
    return o;
}

// trace_multi + transform_hit_info + complete_transformed_hit_info
// (src/kernel/trace-bvh.cl:123-183, 419-463): every root traced from the
// ray's own tMax; a root's hit point is taken back to the world and its t is
// the world distance from the origin; the first root with the smallest t
// wins.  The normal is the interpolated object normal times tr (not its
// inverse transpose), flipped when it faces the ray.  Hit ids index the
// shared index buffer.
void orc_trace_multi(const float* pos, const float* nrm, const float* uvs, const uint32_t* idx,
                     const orc_node* nodes, const orc_root* roots, int nroots, const float* ori4,
                     const float* dir4, uint64_t count, uint32_t* flags, float* t, int32_t* id, float* bary2,
                     float* n4, float* p4, float* uv2) {
    for (uint64_t k = 0; k < count; ++k) {
        const Ray r = load_ray(ori4, dir4, k);
        Hit best{-1, 0.f, 0.f, 0.f};
        int best_root = -1;
        for (int i = 0; i < nroots; ++i) {
            const Ray tr = transform_ray(r, roots[i].tr_inv, 1e-26f);
            Hit h = trace_ray(tr, pos, idx, nodes, nullptr, (uint32_t)roots[i].node);
            if (h.id < 0) continue;
            const f3 hp = mul_h(roots[i].tr, tr.ori + tr.dir * h.t, 1.f, 1e-26f);
            const f3 dv = hp - r.ori;
            h.t = std::sqrt(dot(dv, dv));
            if (best.id < 0 || h.t < best.t) {
                best = h;
                best_root = i;
            }
        }
        if (best.id < 0) {
            flags[k] = 0;
            t[k] = 0.f;
            id[k] = -1;
            if (bary2) { bary2[2 * k] = 0.f; bary2[2 * k + 1] = 0.f; }
            for (int c = 0; c < 4; ++c) { n4[4 * k + c] = 0.f; p4[4 * k + c] = 0.f; }
            uv2[2 * k] = uv2[2 * k + 1] = 0.f;
            continue;
        }
        const f3 hp = r.ori + r.dir * best.t;
        uint32_t i0 = idx[3 * (size_t)best.id], i1 = idx[3 * (size_t)best.id + 1], i2 = idx[3 * (size_t)best.id + 2];
        float u = best.u, v = best.v, w = 1.f - (u + v);
        f3 n0 = normalize(ld3(nrm + 4 * (size_t)i0));
        f3 n1 = normalize(ld3(nrm + 4 * (size_t)i1));
        f3 n2 = normalize(ld3(nrm + 4 * (size_t)i2));
        f3 n = mul_h(roots[best_root].tr, normalize(w * n0 + v * n2 + u * n1), 0.f, 1e-26f);
        uint32_t fl = ORC_HIT;
        if (dot(n, r.dir) > 0.f) {
            fl |= ORC_INV_N;
            n = -n;
        }
        float tu0 = uvs[2 * (size_t)i0], tv0 = uvs[2 * (size_t)i0 + 1];
        float tu1 = uvs[2 * (size_t)i1], tv1 = uvs[2 * (size_t)i1 + 1];
        float tu2 = uvs[2 * (size_t)i2], tv2 = uvs[2 * (size_t)i2 + 1];
        flags[k] = fl;
        t[k] = best.t;
        id[k] = best.id;
        if (bary2) { bary2[2 * k] = u; bary2[2 * k + 1] = v; }
        n4[4 * k] = n.x; n4[4 * k + 1] = n.y; n4[4 * k + 2] = n.z; n4[4 * k + 3] = 0.f;
        p4[4 * k] = hp.x; p4[4 * k + 1] = hp.y; p4[4 * k + 2] = hp.z; p4[4 * k + 3] = 0.f;
        uv2[2 * k] = w * tu0 + v * tu2 + u * tu1;
        uv2[2 * k + 1] = w * tv0 + v * tv2 + u * tv1;
    }
}

// shadow_trace_multi (src/kernel/shadow-trace-bvh.cl:317-373): from the hit
// point toward each light, tMin 0.01, tMax 1e37; shadowed when any root has
// a hit.  A spot light is tested as that kernel does: a point with
// dot(toward the light, spot dir) < angle is lit, not traced.  Point lights
// (the lights extension) stop at the light, as in the single-root kernel.
void orc_trace_shadow_multi(const float* pos, const uint32_t* idx, const orc_node* nodes,
                            const orc_root* roots, int nroots, const orc_lights* lights, uint64_t count,
                            uint32_t* flags, const float* p4) {
    for (uint64_t k = 0; k < count; ++k) {
        uint32_t fl = flags[k] & 0xffu;
        if (fl & ORC_HIT) {
            const f3 p = ld3(p4 + 4 * k);
            for (int l = 0; l < lights->count; ++l) {
                const orc_light& L = lights->l[l];
                Ray r;
                r.ori = p;
                r.tmin = 0.01f;
                r.tmax = 1e37f;
                if (L.type == ORC_DIR) {
                    r.dir = -ld3(L.dir);
                } else if (L.type == ORC_SPOT) {
                    r.dir = normalize(ld3(L.pos) - p);
                    if (dot(r.dir, ld3(L.dir)) < L.angle) continue;
                } else if (L.type == ORC_POINT) {
                    const f3 toL = ld3(L.pos) - p;
                    r.tmax = std::sqrt(dot(toL, toL));
                    r.dir = normalize(toL);
                } else {
                    continue;
                }
                r.inv = recip(r.dir);
                bool hit = false;
                for (int i = 0; i < nroots && !hit; ++i)
                    hit = trace_shadow(transform_ray(r, roots[i].tr_inv, 0.00001f), pos, idx, nodes, nullptr,
                                       (uint32_t)roots[i].node);
                if (hit) fl |= 1u << (8 + l);
            }
        }
        flags[k] = fl;
    }
}

// Renderer::render_to_framebuffer over a scene of objects (root_count() > 1,
// src/rt/renderer.cpp:113-214 with the multi-root kernels of
// src/rt/tracer.cpp:97, 223, 301, 363): orc_render's sequence with
// orc_trace_multi / orc_trace_shadow_multi; secondary rays and shading read
// the shared material map (one material per mesh, Scene::create_bvhs).
uint64_t orc_render_multi(const float* pos, const float* nrm, const float* uvs, const uint32_t* idx,
                          const orc_node* nodes, const orc_root* roots, int nroots, const int32_t* matmap,
                          const orc_material* mats, const orc_lights* lights, const orc_camera* cam, int width,
                          int height, int quad, int max_bounces, uint64_t cap, const orc_texset* tx,
                          uint32_t* accum, orc_frame_stats* st) {
    uint64_t P = (uint64_t)width * height;
    if (cap < P) cap = P;
    std::vector<float> o4a(4 * cap), d4a(4 * cap), o4b(4 * cap), d4b(4 * cap), ca(cap), cb(cap);
    std::vector<int32_t> pa(cap), pb(cap), id(cap);
    std::vector<uint32_t> fl(cap);
    std::vector<float> t(cap), bary(2 * cap), n4(4 * cap), p4(4 * cap), uv2(2 * cap);
    const float ps[3] = {0.f, 0.f, 1.f};
    orc_raygen_primary(cam, width, height, quad, ps, 0, P, o4a.data(), d4a.data(), pa.data(), ca.data());
    std::memset(accum, 0, sizeof(uint32_t) * 3 * P);
    orc_frame_stats s;
    std::memset(&s, 0, sizeof s);
    s.primary = P;
    uint64_t n = P;
    float *o_in = o4a.data(), *d_in = d4a.data(), *c_in = ca.data();
    int32_t* p_in = pa.data();
    float *o_out = o4b.data(), *d_out = d4b.data(), *c_out = cb.data();
    int32_t* p_out = pb.data();
    for (int b = -1; b < max_bounces; ++b) {
        if (b >= 0) {
            uint64_t m = orc_raygen_secondary(mats, matmap, o_in, d_in, p_in, c_in, fl.data(), id.data(), n4.data(),
                                              p4.data(), n, cap, o_out, d_out, p_out, c_out, nullptr);
            std::swap(o_in, o_out);
            std::swap(d_in, d_out);
            std::swap(p_in, p_out);
            std::swap(c_in, c_out);
            n = m;
            if (!n) break;
            s.secondary += n;
            if (b < 16) s.per_bounce[b] = n;
            s.bounces_run++;
        }
        orc_trace_multi(pos, nrm, uvs, idx, nodes, roots, nroots, o_in, d_in, n, fl.data(), t.data(), id.data(),
                        bary.data(), n4.data(), p4.data(), uv2.data());
        orc_trace_shadow_multi(pos, idx, nodes, roots, nroots, lights, n, fl.data(), p4.data());
        for (uint64_t k = 0; k < n; ++k)
            if (fl[k] & ORC_HIT) s.shadow += (uint64_t)lights->count;
        orc_shade_tex(mats, matmap, lights, d_in, p_in, c_in, fl.data(), id.data(), n4.data(), p4.data(), uv2.data(),
                      n, tx, accum);
    }
    if (st) *st = s;
    return s.primary + s.secondary;
}

// 1: every trace of the oracle walks the tree without a stack (the no-stack
// kernels); 0: the stack walk of trace-bvh.cl / shadow-trace-bvh.cl.
void orc_set_traversal(int stackless) { g_stackless = stackless; }

int orc_abi_version(void) { return 1; }

}  // extern "C"
