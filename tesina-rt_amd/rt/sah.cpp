// CPU SAH BVH, the reference's alternative accelerator (BVH::construct /
// construct_and_map and BVHNode::sort, src/rt/bvh.cpp:7-175, 198-314, with
// BBox of src/rt/bbox.cpp), restated so that it builds the same tree:
//   - a node of <= 2 triangles is a leaf;
//   - else the axis is the bbox's largest (x if dx > dy and dx > dz, else y
//     if dy > dz, else z); the node's triangles are sorted by centroid on it
//     (std::sort, as the reference); 32 SAH buckets over the node bbox's
//     extent, slot = min(uint(32 * (c - min) / len), 31); the cost of a split
//     after bucket i is 0.125 + (nl * SA(left) + nr * SA(right)) / SA(node),
//     the first lowest cost wins; the split is the first triangle whose
//     centroid reaches min + (i + 1) * (len / 32); a split that leaves one side
//     empty retries once on the next axis, then the node is a leaf;
//   - children are allocated as a pair at the end of the node array (left,
//     right = left + 1), the left subtree built first; parents are recorded;
//   - bucket boxes start from the reference's empty box: lo = FLT_MAX and hi =
//     FLT_MIN (the smallest positive float, numeric_limits<float>::min());
//     min / max are std::min / std::max; centroid = (hi + lo) * 0.5; surface
//     area = (dx (dy + dz) + dy dz) * 2 with d = max(hi - lo, 0).
// The float-to-slot conversion of a NaN (a node of zero extent) is taken as
// slot 0, what the reference's x86-64 build computes.
#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/rtx.h"

void rtx_set_error(const std::string& msg);

namespace {

struct Box {
    float lo[3], hi[3];
    static Box empty() {
        Box b;
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = std::numeric_limits<float>::max();
            b.hi[k] = std::numeric_limits<float>::min();
        }
        return b;
    }
    void merge(const Box& o) {
        for (int k = 0; k < 3; ++k) {
            hi[k] = std::max(hi[k], o.hi[k]);
            lo[k] = std::min(lo[k], o.lo[k]);
        }
    }
    float centroid(int k) const { return (hi[k] + lo[k]) * 0.5f; }
    float area() const {
        float d[3];
        for (int k = 0; k < 3; ++k) d[k] = std::max(hi[k] - lo[k], 0.f);
        return (d[0] * (d[1] + d[2]) + d[1] * d[2]) * 2.f;
    }
    int largest_axis() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx > dy && dx > dz) return 0;
        if (dy > dz) return 1;
        return 2;
    }
};

struct Node {
    Box box;
    uint32_t a = 0, b = 0;  // children (internal) or [start, end) (leaf)
    uint32_t parent = 0;
    bool leaf = false;
};

constexpr uint32_t MIN_PRIMS = 2;
constexpr uint32_t BUCKETS = 32;

struct Builder {
    const std::vector<Box>& tb;
    std::vector<uint32_t>& order;
    std::vector<Node>& nodes;
    uint32_t max_depth = 0;

    Box node_box(uint32_t s, uint32_t e) const {
        Box b = tb[order[s]];
        for (uint32_t i = s; i < e; ++i) b.merge(tb[order[i]]);
        return b;
    }

    uint32_t split_sah(const Box& nb, int axis, uint32_t s, uint32_t e) const {
        Box bbox[BUCKETS];
        uint32_t prims[BUCKETS];
        for (uint32_t k = 0; k < BUCKETS; ++k) {
            bbox[k] = Box::empty();
            prims[k] = 0;
        }
        const float len = nb.hi[axis] - nb.lo[axis], mn = nb.lo[axis], sa = nb.area();
        for (uint32_t i = s; i < e; ++i) {
            const Box& t = tb[order[i]];
            const float f = (float)BUCKETS * ((t.centroid(axis) - mn) / len);
            uint32_t slot = (f >= 0.f && f < 4294967296.f) ? (uint32_t)f : 0u;
            slot = std::min(slot, BUCKETS - 1);
            prims[slot]++;
            bbox[slot].merge(t);
        }
        float best = 0.f;
        uint32_t best_i = 0;
        for (uint32_t i = 0; i + 1 < BUCKETS; ++i) {
            Box bl = bbox[0], br = bbox[BUCKETS - 1];
            uint32_t nl = prims[0], nr = prims[BUCKETS - 1];
            for (uint32_t j = 1; j <= i; ++j) {
                bl.merge(bbox[j]);
                nl += prims[j];
            }
            for (uint32_t j = i + 1; j < BUCKETS; ++j) {
                br.merge(bbox[j]);
                nr += prims[j];
            }
            const float cost = .125f + ((float)nl * bl.area() + (float)nr * br.area()) / sa;
            if (i == 0 || cost < best) {
                best = cost;
                best_i = i;
            }
        }
        float split = mn;
        split += (float)(best_i + 1) * (len / (float)BUCKETS);
        uint32_t i = s;
        for (; i < e; ++i)
            if (tb[order[i]].centroid(axis) >= split) break;
        return i;
    }

    // BVHNode::sort for node `self` (its range already set)
    void build(uint32_t self, uint32_t depth) {
        max_depth = std::max(max_depth, depth);
        const uint32_t s = nodes[self].a, e = nodes[self].b;
        nodes[self].box = node_box(s, e);
        if (e - s <= MIN_PRIMS) {
            nodes[self].leaf = true;
            return;
        }
        int axis = nodes[self].box.largest_axis();
        uint32_t at = s;
        for (int tries = 0; tries < 2; ++tries) {
            std::sort(order.begin() + s, order.begin() + e,
                      [&](uint32_t x, uint32_t y) { return tb[x].centroid(axis) < tb[y].centroid(axis); });
            at = split_sah(nodes[self].box, axis, s, e);
            if (at == s || at == e)
                axis = (axis + 1) % 3;
            else
                break;
        }
        if (at == s || at == e) {
            nodes[self].leaf = true;
            return;
        }
        const uint32_t l = (uint32_t)nodes.size();
        nodes.resize(nodes.size() + 2);
        nodes[l].a = s;
        nodes[l].b = at;
        nodes[l].parent = self;
        nodes[l + 1].a = at;
        nodes[l + 1].b = e;
        nodes[l + 1].parent = self;
        build(l, depth + 1);
        build(l + 1, depth + 1);
        nodes[self].leaf = false;
        nodes[self].a = l;
        nodes[self].b = l + 1;
    }
};

}  // namespace

// Builds the tree of triangles idx3 (T) over positions pos3; fills the nodes
// (reference layout), the triangle order (sorted position -> input triangle)
// and the depth of the deepest node.  T >= 1.
int32_t rtx_sah_construct(const float* pos3, uint32_t V, const uint32_t* idx3, uint32_t T,
                          std::vector<rtx_node_ref>& out, std::vector<uint32_t>& order, uint32_t* max_depth) {
    std::vector<Box> tb(T);
    for (uint32_t t = 0; t < T; ++t) {
        const uint32_t i0 = idx3[3 * (size_t)t], i1 = idx3[3 * (size_t)t + 1], i2 = idx3[3 * (size_t)t + 2];
        if (i0 >= V || i1 >= V || i2 >= V) {
            rtx_set_error("rtx_sah: vertex index out of range");
            return RTX_EINVAL;
        }
        for (int k = 0; k < 3; ++k) {
            const float a = pos3[3 * (size_t)i0 + k], b = pos3[3 * (size_t)i1 + k], c = pos3[3 * (size_t)i2 + k];
            tb[t].hi[k] = std::max(a, std::max(b, c));
            tb[t].lo[k] = std::min(a, std::min(b, c));
        }
    }
    order.resize(T);
    for (uint32_t t = 0; t < T; ++t) order[t] = t;
    std::vector<Node> nodes(1);
    nodes.reserve(2 * (size_t)T);
    nodes[0].a = 0;
    nodes[0].b = T;
    nodes[0].parent = 0;
    Builder bld{tb, order, nodes};
    bld.build(0, 0);
    *max_depth = bld.max_depth;
    out.resize(nodes.size());
    for (size_t n = 0; n < nodes.size(); ++n) {
        rtx_node_ref& o = out[n];
        std::memset(&o, 0, sizeof o);
        for (int k = 0; k < 3; ++k) {
            o.lo[k] = nodes[n].box.lo[k];
            o.hi[k] = nodes[n].box.hi[k];
        }
        o.a = nodes[n].a;
        o.b = nodes[n].b;
        o.parent = nodes[n].parent;
        o.leaf = nodes[n].leaf ? 1 : 0;
    }
    return RTX_OK;
}

int32_t rtx_sah_build_host(const float* pos3, uint32_t V, const uint32_t* idx3, uint32_t T, rtx_node_ref* nodes,
                           uint32_t cap, uint32_t* count, uint32_t* order) {
    if (!pos3 || !idx3 || !T || !count) {
        rtx_set_error("rtx_sah_build_host: null argument or no triangles");
        return RTX_EINVAL;
    }
    std::vector<rtx_node_ref> nd;
    std::vector<uint32_t> ord;
    uint32_t depth = 0;
    int32_t e = rtx_sah_construct(pos3, V, idx3, T, nd, ord, &depth);
    if (e) return e;
    *count = (uint32_t)nd.size();
    if (nodes) {
        if (cap < nd.size()) {
            rtx_set_error("rtx_sah_build_host: capacity below node count");
            return RTX_EINVAL;
        }
        std::copy(nd.begin(), nd.end(), nodes);
    }
    if (order) std::copy(ord.begin(), ord.end(), order);
    return RTX_OK;
}
