// Ray bundles, primary ray generation, closest-hit and shadow traversal.
//
// Reference: PrimaryRayGenerator (src/rt/primary-ray-generator.cpp:58-165,
// src/kernel/primary-ray-generator.cl:28-103), Tracer::trace / shadow_trace
// (src/rt/tracer.cpp:140-352; src/kernel/trace-bvh.cl:190-481,
// src/kernel/shadow-trace-bvh.cl:160-418).
//
// Traversal keeps the reference's visiting order exactly (test a node's box
// when it is visited, near child first by dot(dir, right.lo - left.lo) > 0,
// pop only while 0 < level < 32, tMax shrinks to the best hit) so that hit
// ids, distances and tie winners are those of the oracle.  The MI355X layout
// is what differs: 32-byte nodes plus a 16-byte delta fetched only at
// internal nodes, triangles stored contiguously in leaf order, the stack in
// LDS ([level][thread], conflict-free), SoA rays and hits.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <type_traits>

#include "devmath.hpp"
#include "rtx_impl.hpp"
#include "stack.hpp"
#include "wavefront.hpp"

namespace {


__device__ __forceinline__ uint64_t ray_count(const uint32_t* dcount, uint64_t hcount) {
    return hcount == RTX_DEVICE_COUNT ? (uint64_t)*dcount : hcount;
}

// Work split over the L2 domains.  Work-groups are dispatched round-robin
// over the XCDs (group g on XCD g mod X), each with its own L2, so XCD x
// takes the x-th contiguous chunk of the n items: neighbouring rays (same
// screen region, same BVH nodes) then share one L2 instead of all eight.
// The grid is a multiple of X; a group walks its chunk with stride
// (grid / X) * TB.
struct XcdSpan {
    uint64_t begin, end, stride;
};
__device__ __forceinline__ XcdSpan xcd_span(uint64_t n, uint32_t X, uint32_t TB) {
    const uint32_t x = blockIdx.x % X, slot = blockIdx.x / X, per = gridDim.x / X;
    XcdSpan sp;
    const uint64_t lo = n * x / X, hi = n * (x + 1) / X;
    sp.begin = lo + (uint64_t)slot * TB;
    sp.end = hi;
    sp.stride = (uint64_t)per * TB;
    return sp;
}

// ------------------------------------------------------------- raygen ----
// Pixel of sample i (generate_primary_rays, src/kernel/primary-ray-generator.cl:
// 28-103): quad x quad blocks row by row, then the strip right of them row by
// row, then the strip above them column by column.
__device__ __forceinline__ void quad_pixel(int i, int width, int height, int quad, int& x, int& y) {
    int bs = quad, ba = quad * quad;
    int xb = width / bs, yb = height / bs;
    int block_rays = xb * yb * ba;
    int right_rays = (width - xb * bs) * height;
    if (i < block_rays) {
        int block = i / ba;
        int y_block = block / xb, x_block = block % xb;
        int bp = i % ba;
        x = x_block * bs + bp % bs;
        y = y_block * bs + bp / bs;
    } else if (i < block_rays + right_rays) {
        int sx = xb * bs, slab_w = width - sx, ri = i - block_rays;
        x = sx + ri % slab_w;
        y = ri / slab_w;
    } else {
        int sy = yb * bs, slab_h = height - sy, ti = i - (block_rays + right_rays);
        x = ti / slab_h;
        y = sy + ti % slab_h;
    }
}

// Pinhole direction through pixel (x, y) offset by the sample (s0, s1).
__device__ __forceinline__ v3 primary_dir(const rtx_camera& cam, int x, int y, int width, int height, float s0,
                                          float s1) {
    float xn = (0.5f + (float)x + s0) / (float)width;
    float yn = (0.5f + (float)y + s1) / (float)height;
    v3 dir = mk3(cam.dir[0], cam.dir[1], cam.dir[2]);
    v3 right = mk3(cam.right[0], cam.right[1], cam.right[2]);
    v3 up = mk3(cam.up[0], cam.up[1], cam.up[2]);
    return normalize(dir + right * (xn * 2.f - 1.f) + up * (yn * 2.f - 1.f));
}

__global__ __launch_bounds__(256) void k_raygen_primary(float4* __restrict__ ori4, float4* __restrict__ dir4,
                                                        int32_t* __restrict__ pixel, float* __restrict__ contrib,
                                                        uint32_t* dcount, rtx_camera cam, int width, int height,
                                                        int quad, const float* __restrict__ ps, uint64_t offset,
                                                        uint64_t count) {
    uint64_t k = blockIdx.x * 256ull + threadIdx.x;
    if (k == 0) *dcount = (uint32_t)count;
    if (k >= count) return;
    int gid = (int)(offset + k);
    int pixels = width * height;
    int i = gid % pixels;
    const float* smp = ps + 3 * (gid / pixels);
    int x, y;
    quad_pixel(i, width, height, quad, x, y);
    const v3 d = primary_dir(cam, x, y, width, height, smp[0], smp[1]);
    ori4[k] = make_float4(cam.pos[0], cam.pos[1], cam.pos[2], 0.f);
    dir4[k] = make_float4(d.x, d.y, d.z, 1e36f);
    pixel[k] = width * y + x;
    contrib[k] = smp[2];
}

// ---------------------------------------------------------- traversal --
// Per-ray constants of the slab test.  The reference skips an axis when
// |invDir| <= 1e-6 (src/kernel/trace-bvh.cl:262-270); that is fixed per ray.
struct RayC {
    v3 o, d, inv;
    float tmin;
    bool ax, ay, az;
};

__device__ __forceinline__ RayC make_ray(v3 o, v3 d, float tmin) {
    RayC r;
    r.o = o;
    r.d = d;
    r.inv = recip(d);
    r.tmin = tmin;
    r.ax = fabsf(r.inv.x) > 1e-6f;
    r.ay = fabsf(r.inv.y) > 1e-6f;
    r.az = fabsf(r.inv.z) > 1e-6f;
    return r;
}

// bbox_hit (src/kernel/trace-bvh.cl:247-273).  The exit distance is
// fmin(tMax, m) with m the box's own exit, so a box that fails with some tMax
// fails with every smaller one: testing both children when the parent is
// visited (with the current tMax) and again when a pushed child is popped
// (with the then-current tMax) gives exactly the reference's decisions.
__device__ __forceinline__ bool slab(float4 lo4, float4 hi4, const RayC& r, float tmax) {
    v3 tl = (xyz(lo4) - r.o) * r.inv;
    v3 th = (xyz(hi4) - r.o) * r.inv;
    v3 tM = vmax(tl, th);
    v3 tm = vmin(tl, th);
    float t0 = r.tmin, t1 = tmax;
    if (r.ax) { t0 = fmaxf(t0, tm.x); t1 = fminf(t1, tM.x); }
    if (r.ay) { t0 = fmaxf(t0, tm.y); t1 = fminf(t1, tM.y); }
    if (r.az) { t0 = fmaxf(t0, tm.z); t1 = fminf(t1, tM.z); }
    return t0 <= t1;
}


struct Closest {
    float t, u, v;
    int32_t id;
};

// Möller–Trumbore of leaf_hit (src/kernel/trace-bvh.cl:282-327), on vertices
// already in registers.
__device__ __forceinline__ bool tri_hit_v(float4 a, float4 b, float4 c, v3 o, v3 d, float* tt, float* uu,
                                          float* vv) {
    v3 v0 = xyz(a), v1 = xyz(b), v2 = xyz(c);
    v3 e1 = v1 - v0, e2 = v2 - v0;
    v3 h = cross(d, e2);
    float aa = dot(e1, h);
    if (aa > -1e-26f && aa < 1e-26f) return false;
    float f = 1.f / aa;
    v3 s = o - v0;
    float u = f * dot(s, h);
    if (u < 0.f || u > 1.f) return false;
    v3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.f || u + v > 1.f) return false;
    *tt = f * dot(e2, q);
    *uu = u;
    *vv = v;
    return true;
}

// Leaf triangles are tested in order, four per round: the vertices of all
// four are requested before any is tested (indices past the leaf are clamped
// to its first triangle, a valid address whose result is discarded), so a
// round costs one memory round trip.
constexpr uint32_t LEAF_ROUND = 4;

__device__ __forceinline__ void leaf_closest(const float4* __restrict__ tri, uint32_t a, uint32_t cnt, const RayC& r,
                                             Closest& best) {
    const uint32_t end = a + cnt;
    for (uint32_t t = a; t < end; t += LEAF_ROUND) {
        float4 v[3 * LEAF_ROUND];
#pragma unroll
        for (uint32_t j = 0; j < LEAF_ROUND; ++j) {
            const size_t tj = 3 * (size_t)(t + j < end ? t + j : a);
            v[3 * j] = tri[tj];
            v[3 * j + 1] = tri[tj + 1];
            v[3 * j + 2] = tri[tj + 2];
        }
#pragma unroll
        for (uint32_t j = 0; j < LEAF_ROUND; ++j) {
            float tt, u, w;
            if (t + j < end && tri_hit_v(v[3 * j], v[3 * j + 1], v[3 * j + 2], r.o, r.d, &tt, &u, &w) &&
                tt <= best.t && tt >= r.tmin) {
                best.t = tt;  // later triangle wins an exact tie
                best.id = (int32_t)(t + j);
                best.u = u;
                best.v = w;
            }
        }
    }
}

__device__ __forceinline__ bool leaf_any(const float4* __restrict__ tri, uint32_t a, uint32_t cnt, const RayC& r,
                                         float tmax) {
    const uint32_t end = a + cnt;
    for (uint32_t t = a; t < end; t += LEAF_ROUND) {
        float4 v[3 * LEAF_ROUND];
#pragma unroll
        for (uint32_t j = 0; j < LEAF_ROUND; ++j) {
            const size_t tj = 3 * (size_t)(t + j < end ? t + j : a);
            v[3 * j] = tri[tj];
            v[3 * j + 1] = tri[tj + 1];
            v[3 * j + 2] = tri[tj + 2];
        }
        bool any = false;
#pragma unroll
        for (uint32_t j = 0; j < LEAF_ROUND; ++j) {
            float tt, u, w;
            any = any || (t + j < end && tri_hit_v(v[3 * j], v[3 * j + 1], v[3 * j + 2], r.o, r.d, &tt, &u, &w) &&
                          tt < tmax && tt > r.tmin);
        }
        if (any) return true;
    }
    return false;
}

// Closest hit: the reference's depth-first order (near child first by
// dot(dir, right.lo - left.lo) > 0, src/kernel/trace-bvh.cl:392-409), visiting
// two sibling boxes per 64-byte load (siblings are adjacent in the node
// array).  Leaves are reached in the same sequence with the same tMax as in
// trace_ray (src/kernel/trace-bvh.cl:330-417), so hits and tie winners match.
template <class Stack>
__device__ __forceinline__ void traverse_closest(const float4* __restrict__ node, const float4* __restrict__ tri,
                                                 const RayC& r, Closest& best, Stack& stk) {
    {
        float4 r0 = node[0], r1 = node[1];
        if (!slab(r0, r1, r, best.t)) return;
        uint32_t cnt = __float_as_uint(r1.w);
        if (cnt) {
            leaf_closest(tri, __float_as_uint(r0.w), cnt, r, best);
            return;
        }
    }
    uint32_t cur = __float_as_uint(node[0].w);
    uint32_t level = 0;
    stk.reset();
    for (;;) {
        const float4 L0 = node[2 * (size_t)cur], L1 = node[2 * (size_t)cur + 1];
        const float4 R0 = node[2 * (size_t)cur + 2], R1 = node[2 * (size_t)cur + 3];
        const bool hl = slab(L0, L1, r, best.t);
        const bool hr = slab(R0, R1, r, best.t);
        const bool nl = dot(r.d, xyz(R0) - xyz(L0)) > 0.f;
        if (nl ? hr : hl) stk.put(level++, nl ? cur + 1 : cur);
        if (nl ? hl : hr) {
            const uint32_t na = __float_as_uint(nl ? L0.w : R0.w), nc = __float_as_uint(nl ? L1.w : R1.w);
            if (!nc) {
                cur = na;
                continue;
            }
            leaf_closest(tri, na, nc, r, best);
        }
        bool next = false;
        while (level > 0) {
            const uint32_t id = stk.get(--level);
            const float4 P0 = node[2 * (size_t)id], P1 = node[2 * (size_t)id + 1];
            if (!slab(P0, P1, r, best.t)) continue;  // the reference's test at visit time
            const uint32_t pa = __float_as_uint(P0.w), pc = __float_as_uint(P1.w);
            if (pc) {
                leaf_closest(tri, pa, pc, r, best);
                continue;
            }
            cur = pa;
            next = true;
            break;
        }
        if (!next) return;
    }
}

// Any hit with tMin < t < tMax (trace_shadow_ray + leaf_hit_any,
// src/kernel/shadow-trace-bvh.cl:188-315).  tMax is fixed, so the set of
// leaves whose boxes pass does not depend on the visiting order.
template <class Stack>
__device__ __forceinline__ bool traverse_any(const float4* __restrict__ node, const float4* __restrict__ tri,
                                             const RayC& r, float tmax, Stack& stk) {
    {
        float4 r0 = node[0], r1 = node[1];
        if (!slab(r0, r1, r, tmax)) return false;
        uint32_t cnt = __float_as_uint(r1.w);
        if (cnt) return leaf_any(tri, __float_as_uint(r0.w), cnt, r, tmax);
    }
    uint32_t cur = __float_as_uint(node[0].w);
    uint32_t level = 0;
    stk.reset();
    for (;;) {
        const float4 L0 = node[2 * (size_t)cur], L1 = node[2 * (size_t)cur + 1];
        const float4 R0 = node[2 * (size_t)cur + 2], R1 = node[2 * (size_t)cur + 3];
        const bool hl = slab(L0, L1, r, tmax);
        const bool hr = slab(R0, R1, r, tmax);
        const bool nl = dot(r.d, xyz(R0) - xyz(L0)) > 0.f;
        if (nl ? hr : hl) stk.put(level++, nl ? cur + 1 : cur);
        if (nl ? hl : hr) {
            const uint32_t na = __float_as_uint(nl ? L0.w : R0.w), nc = __float_as_uint(nl ? L1.w : R1.w);
            if (!nc) {
                cur = na;
                continue;
            }
            if (leaf_any(tri, na, nc, r, tmax)) return true;
        }
        bool next = false;
        while (level > 0) {
            // tMax is fixed: the box passed when pushed, it passes now
            const uint32_t id = stk.get(--level);
            const float4 P0 = node[2 * (size_t)id], P1 = node[2 * (size_t)id + 1];
            const uint32_t pa = __float_as_uint(P0.w), pc = __float_as_uint(P1.w);
            if (pc) {
                if (leaf_any(tri, pa, pc, r, tmax)) return true;
                continue;
            }
            cur = pa;
            next = true;
            break;
        }
        if (!next) return false;
    }
}

// Wave-cooperative leaf test.  Leaves of at least `coop` triangles (long
// runs of equal Morton codes, e.g. the pole fans of UV spheres, reach
// hundreds) are tested by every live lane of the wave together, 64
// triangles per round, instead of by their own lane alone while the rest
// of the wave waits.  The owner's ray and current tMax are broadcast; the
// round's hits are reduced to the smallest t and, among equal t, the
// highest index, which is what the sequential in-order loop with its
// `t <= best` update keeps (a later triangle wins an exact tie).
__device__ __forceinline__ float bcast(float x, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}
__device__ __forceinline__ uint32_t bcast(uint32_t x, int src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, src);
}

// Closest hit over the fat leaves flagged in this iteration (cnt > 0 on
// the owning lanes).  Must be called by all 64 lanes of the wave (`live` is
// then the full mask): the shuffles of the reduction read every lane.
// The owner lane's pointer, for callers whose lanes walk different trees.
__device__ __forceinline__ const float4* bcast(const float4* p, int src) {
    const uint64_t v = (uint64_t)p;
    const uint64_t lo = bcast((uint32_t)v, src), hi = bcast((uint32_t)(v >> 32), src);
    return reinterpret_cast<const float4*>((hi << 32) | lo);
}

// LANE_TRI: `tri` differs between lanes (each lane's own tree); the owner's
// is used for its leaf.
template <bool LANE_TRI = false>
__device__ __forceinline__ void coop_closest(const float4* __restrict__ tri, uint32_t la, uint32_t lcnt,
                                             const RayC& r, Closest& best, uint64_t live) {
    uint64_t owners = ballot(lcnt != 0u);
    if (!owners) return;
    const uint32_t nlive = (uint32_t)__builtin_popcountll(live);
    const uint32_t rank = lanes_below(live);
    const int lane = (int)lane_id();
    while (owners) {
        const int src = __builtin_ctzll(owners);
        owners &= owners - 1;
        const uint32_t a = bcast(la, src), cnt = bcast(lcnt, src);
        const float4* __restrict__ ts = LANE_TRI ? bcast(tri, src) : tri;
        const v3 o = mk3(bcast(r.o.x, src), bcast(r.o.y, src), bcast(r.o.z, src));
        const v3 d = mk3(bcast(r.d.x, src), bcast(r.d.y, src), bcast(r.d.z, src));
        const float tmin = bcast(r.tmin, src);
        for (uint32_t base = 0; base < cnt; base += nlive) {
            const float bt = bcast(best.t, src);
            const uint32_t k = base + rank;
            float tt = 0.f, u = 0.f, v = 0.f;
            bool h = false;
            if (k < cnt) {
                const size_t t3 = 3 * (size_t)(a + k);
                h = tri_hit_v(ts[t3], ts[t3 + 1], ts[t3 + 2], o, d, &tt, &u, &v) && tt <= bt && tt >= tmin;
            }
            // key: smallest t first (the +0.f folds -0 into +0, which compare
            // equal), then the highest index
            uint64_t key = ~0ull;
            if (h) key = ((uint64_t)((uint32_t)fkey(tt + 0.f) ^ 0x80000000u) << 32) | (uint32_t)~k;
            uint64_t m = key;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint64_t y = __shfl_xor(m, off, RTX_WAVE);
                m = y < m ? y : m;
            }
            const uint64_t wm = ballot(h && key == m);
            if (m != ~0ull && wm) {
                const int win = __builtin_ctzll(wm);
                const float wt = bcast(tt, win), wu = bcast(u, win), wv = bcast(v, win);
                const uint32_t wk = bcast(k, win);
                if (lane == src) {
                    best.t = wt;
                    best.u = wu;
                    best.v = wv;
                    best.id = (int32_t)(a + wk);
                }
            }
        }
    }
}

// Any hit over the flagged fat leaves; returns true on the owning lanes
// whose leaf holds a hit with tMin < t < tMax.
template <bool LANE_TRI = false>
__device__ __forceinline__ bool coop_any(const float4* __restrict__ tri, uint32_t la, uint32_t lcnt, const RayC& r,
                                         float tmax, uint64_t live) {
    uint64_t owners = ballot(lcnt != 0u);
    bool found = false;
    if (!owners) return false;
    const uint32_t nlive = (uint32_t)__builtin_popcountll(live);
    const uint32_t rank = lanes_below(live);
    const int lane = (int)lane_id();
    while (owners) {
        const int src = __builtin_ctzll(owners);
        owners &= owners - 1;
        const uint32_t a = bcast(la, src), cnt = bcast(lcnt, src);
        const float4* __restrict__ ts = LANE_TRI ? bcast(tri, src) : tri;
        const v3 o = mk3(bcast(r.o.x, src), bcast(r.o.y, src), bcast(r.o.z, src));
        const v3 d = mk3(bcast(r.d.x, src), bcast(r.d.y, src), bcast(r.d.z, src));
        const float tmin = bcast(r.tmin, src), tm = bcast(tmax, src);
        for (uint32_t base = 0; base < cnt; base += nlive) {
            const uint32_t k = base + rank;
            bool h = false;
            if (k < cnt) {
                const size_t t3 = 3 * (size_t)(a + k);
                float tt, u, v;
                h = tri_hit_v(ts[t3], ts[t3 + 1], ts[t3 + 2], o, d, &tt, &u, &v) && tt < tm && tt > tmin;
            }
            if (ballot(h)) {
                if (lane == src) found = true;
                break;
            }
        }
    }
    return found;
}

// Single-node form of the same traversal (one node per step, as
// trace_ray visits them).  SPEC issues the node's delta load together with
// the node load instead of after the box test.  Every lane stays in the
// loop until the whole wave is done, so that finished lanes help test the
// fat leaves of the others (leaves of at least `coop` triangles).  Called
// by all lanes of the wave; lanes without a ray pass start = false.
template <bool SPEC, class Stack>
__device__ __forceinline__ void traverse_closest_node(const float4* __restrict__ node,
                                                      const float4* __restrict__ delta,
                                                      const float4* __restrict__ tri, const RayC& r, Closest& best,
                                                      Stack& stk, uint32_t coop, bool start) {
    const uint64_t live = ballot(true);
    uint32_t level = 0, curr = 0;
    stk.reset();
    bool going = start;
    while (ballot(going)) {
        uint32_t la = 0, lcnt = 0;
        if (going) {
            const float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
            float4 dl4 = make_float4(0.f, 0.f, 0.f, 0.f);
            if (SPEC) {
                dl4 = delta[curr];
                // keep the delta load next to the node load: without this the
                // compiler sinks it into the internal-node branch, after the
                // box test, and each internal visit waits for two round trips
                __asm__ volatile("" : "+v"(dl4.x), "+v"(dl4.y), "+v"(dl4.z));
            }
            const uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
            bool pop = true;
            if (slab(f0, f1, r, best.t)) {
                if (cnt >= coop) {
                    la = a;
                    lcnt = cnt;
                } else if (cnt) {
                    leaf_closest(tri, a, cnt, r, best);
                } else {
                    if (!SPEC) dl4 = delta[curr];
                    const bool nl = dot(r.d, xyz(dl4)) > 0.f;
                    stk.put(level, nl ? a + 1 : a);
                    level++;
                    curr = nl ? a : a + 1;
                    pop = false;
                }
            }
            if (pop) {
                // the reference pops only below MAX_LEVELS and ends the ray
                // otherwise (src/kernel/trace-bvh.cl:361-367)
                if (level == 0 || level >= STACK)
                    going = false;
                else
                    curr = stk.get(--level);
            }
        }
        coop_closest(tri, la, lcnt, r, best, live);
    }
}

// ORDERED = false: children in index order (left first) with no near-child
// delta fetched.  An any-hit result is the existence of a hit in a leaf whose
// box and ancestors' boxes pass at the fixed tMax, which does not depend on
// the order, unless a ray's stack can reach MAX_LEVELS (trees of 32 levels);
// the host keeps the reference's order for those.
template <bool SPEC, bool ORDERED, class Stack>
__device__ __forceinline__ bool traverse_any_node(const float4* __restrict__ node, const float4* __restrict__ delta,
                                                  const float4* __restrict__ tri, const RayC& r, float tmax,
                                                  Stack& stk, uint32_t coop, bool start) {
    const uint64_t live = ballot(true);
    uint32_t level = 0, curr = 0;
    stk.reset();
    bool going = start, hit = false;
    while (ballot(going)) {
        uint32_t la = 0, lcnt = 0;
        if (going) {
            const float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
            float4 dl4 = make_float4(0.f, 0.f, 0.f, 0.f);
            if (SPEC && ORDERED) {
                dl4 = delta[curr];
                __asm__ volatile("" : "+v"(dl4.x), "+v"(dl4.y), "+v"(dl4.z));  // see traverse_closest_node
            }
            const uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
            bool pop = true;
            if (slab(f0, f1, r, tmax)) {
                if (cnt >= coop) {
                    la = a;
                    lcnt = cnt;
                } else if (cnt) {
                    if (leaf_any(tri, a, cnt, r, tmax)) {
                        hit = true;
                        going = false;
                        pop = false;
                    }
                } else {
                    if (!SPEC && ORDERED) dl4 = delta[curr];
                    const bool nl = !ORDERED || dot(r.d, xyz(dl4)) > 0.f;
                    stk.put(level, nl ? a + 1 : a);
                    level++;
                    curr = nl ? a : a + 1;
                    pop = false;
                }
            }
            if (pop) {
                // the reference pops only below MAX_LEVELS and ends the ray
                // otherwise (src/kernel/trace-bvh.cl:361-367)
                if (level == 0 || level >= STACK)
                    going = false;
                else
                    curr = stk.get(--level);
            }
        }
        if (coop_any(tri, la, lcnt, r, tmax, live)) {
            hit = true;
            going = false;
        }
    }
    return hit;
}

// Variant 4: the stackless walks of src/kernel/trace-bvh-no-stack.cl:408-520
// and shadow-trace-bvh-no-stack.cl:213-311.  Each lane walks its own ray by
// parent pointers (no stack, hence no MAX_LEVELS), the near child first:
// down into a hit box, up from a leaf or a missed box, and from the near
// child on to the far one.  Closest hit: a leaf's hits with t <= its running
// best (from the ray's tMax) replace the ray's best only when strictly nearer
// (merge_hit_info), so across leaves the first of equal hits stays.  Going
// up from the root ends the walk; `bound` (above the 3 visits per node a
// tree allows) ends a walk over a malformed parent array instead of looping.
__device__ void traverse_closest_ns(const float4* __restrict__ node, const float4* __restrict__ delta,
                                    const uint32_t* __restrict__ parent, const float4* __restrict__ tri,
                                    const RayC& r, Closest& best, uint32_t bound) {
    bool up = false;
    uint32_t last = 0, curr = 0, first = 0, second = 0;
    for (uint32_t it = 0; it < bound; ++it) {
        const float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
        const uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
        if (!cnt) {
            const bool nl = dot(r.d, xyz(delta[curr])) > 0.f;
            first = nl ? a : a + 1;
            second = nl ? a + 1 : a;
        }
        const uint32_t up_to = curr == 0 ? 0u : parent[curr];
        if (up) {
            if (last == 0) return;
            if (last == first) {
                last = curr;
                curr = second;
                up = false;
            } else {
                last = curr;
                curr = up_to;
            }
        } else if (slab(f0, f1, r, best.t)) {
            if (cnt) {
                Closest info{best.t, 0.f, 0.f, -1};
                leaf_closest(tri, a, cnt, r, info);
                if (info.id >= 0 && (best.id < 0 || info.t < best.t)) best = info;
                last = curr;
                curr = up_to;
                up = true;
            } else {
                last = curr;
                curr = first;
            }
        } else {
            last = curr;
            curr = up_to;
            up = true;
        }
    }
}

// ordered = false: children in index order with no delta fetched.  The walk
// has no stack, hence no MAX_LEVELS: its any-hit result is whether a leaf
// whose box and ancestors' boxes pass holds a hit, whatever the order.
__device__ bool traverse_any_ns(const float4* __restrict__ node, const float4* __restrict__ delta,
                                const uint32_t* __restrict__ parent, const float4* __restrict__ tri, const RayC& r,
                                float tmax, uint32_t bound, bool ordered = true) {
    bool up = false;
    uint32_t last = 0, curr = 0, first = 0, second = 0;
    for (uint32_t it = 0; it < bound; ++it) {
        const float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
        const uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
        if (!cnt) {
            const bool nl = !ordered || dot(r.d, xyz(delta[curr])) > 0.f;
            first = nl ? a : a + 1;
            second = nl ? a + 1 : a;
        }
        const uint32_t up_to = curr == 0 ? 0u : parent[curr];
        if (up) {
            if (last == 0) return false;
            if (last == first) {
                last = curr;
                curr = second;
                up = false;
            } else {
                last = curr;
                curr = up_to;
            }
        } else if (slab(f0, f1, r, tmax)) {
            if (cnt) {
                if (leaf_any(tri, a, cnt, r, tmax)) return true;
                last = curr;
                curr = up_to;
                up = true;
            } else {
                last = curr;
                curr = first;
            }
        } else {
            last = curr;
            curr = up_to;
            up = true;
        }
    }
    return false;
}

// Variant dispatch: 0 = single node, 1 = single node + early delta load,
// 2 = sibling pairs.
// Called by every lane of the wave; `start` says whether the lane has a ray.
template <int VAR, class Stack>
__device__ __forceinline__ void closest_var(const float4* __restrict__ node, const float4* __restrict__ delta,
                                            const float4* __restrict__ tri, const RayC& r, Closest& best,
                                            Stack& stk, uint32_t coop, bool start,
                                            const uint32_t* __restrict__ parent = nullptr, uint32_t bound = 0) {
    if (VAR == 4) {
        if (start) traverse_closest_ns(node, delta, parent, tri, r, best, bound);
    } else if (VAR == 2) {
        if (start) traverse_closest(node, tri, r, best, stk);
    } else {
        traverse_closest_node<VAR == 1>(node, delta, tri, r, best, stk, coop, start);
    }
}

// ordered: visit children in the reference's near-child order (else any order,
// see traverse_any_node; wave-uniform).
template <int VAR, class Stack>
__device__ __forceinline__ bool any_var(const float4* __restrict__ node, const float4* __restrict__ delta,
                                        const float4* __restrict__ tri, const RayC& r, float tmax, Stack& stk,
                                        uint32_t coop, bool start, const uint32_t* __restrict__ parent = nullptr,
                                        uint32_t bound = 0, bool ordered = true) {
    if (VAR == 4) {
        bool h = false;
        if (start) h = traverse_any_ns(node, delta, parent, tri, r, tmax, bound, ordered);
        return h;
    }
    if (VAR == 2) {
        bool h = false;
        if (start) h = traverse_any(node, tri, r, tmax, stk);
        return h;
    }
    if (!ordered) return traverse_any_node<false, false>(node, delta, tri, r, tmax, stk, coop, start);
    return traverse_any_node<VAR == 1, true>(node, delta, tri, r, tmax, stk, coop, start);
}

// -------------------------------------------------------------- trace ----
__device__ __forceinline__ void write_miss_record(uint64_t rr, uint32_t* __restrict__ flags, float* __restrict__ tout,
                                                  int32_t* __restrict__ idout, float2* __restrict__ bary,
                                                  float4* __restrict__ n4, float4* __restrict__ p4,
                                                  float2* __restrict__ uv2) {
    flags[rr] = 0u;
    tout[rr] = 0.f;
    idout[rr] = -1;
    bary[rr] = make_float2(0.f, 0.f);
    n4[rr] = make_float4(0.f, 0.f, 0.f, 0.f);
    p4[rr] = make_float4(0.f, 0.f, 0.f, 0.f);
    uv2[rr] = make_float2(0.f, 0.f);
}

// complete_trace_info (src/kernel/trace-bvh.cl:190-245) for a ray with a hit.
__device__ __forceinline__ void write_hit_record(const float4* __restrict__ tri, const float4* __restrict__ nrm,
                                                 const float2* __restrict__ uvs, const RayC& rc, const Closest& best,
                                                 uint64_t rr, uint32_t* __restrict__ flags, float* __restrict__ tout,
                                                 int32_t* __restrict__ idout, float2* __restrict__ bary,
                                                 float4* __restrict__ n4, float4* __restrict__ p4,
                                                 float2* __restrict__ uv2) {
    const v3 hp = rc.o + rc.d * best.t;
    const uint32_t i0 = __float_as_uint(tri[3 * (size_t)best.id].w);
    const uint32_t i1 = __float_as_uint(tri[3 * (size_t)best.id + 1].w);
    const uint32_t i2 = __float_as_uint(tri[3 * (size_t)best.id + 2].w);
    const float w = 1.f - (best.u + best.v);
    const v3 n0 = xyz(nrm[i0]), n1 = xyz(nrm[i1]), n2 = xyz(nrm[i2]);
    v3 nn = normalize(w * n0 + best.v * n2 + best.u * n1);
    uint32_t fl = 1u;
    if (dot(nn, rc.d) > 0.f) {
        fl |= 2u;
        nn = -nn;
    }
    const float2 t0 = uvs[i0], t1 = uvs[i1], t2 = uvs[i2];
    flags[rr] = fl;
    tout[rr] = best.t;
    idout[rr] = best.id;
    bary[rr] = make_float2(best.u, best.v);
    n4[rr] = make_float4(nn.x, nn.y, nn.z, 0.f);
    p4[rr] = make_float4(hp.x, hp.y, hp.z, 0.f);
    uv2[rr] = make_float2(w * t0.x + best.v * t2.x + best.u * t1.x, w * t0.y + best.v * t2.y + best.u * t1.y);
}

// Tracer::trace -> trace_single + complete_trace_info.  Rays with a hit are
// also appended (wave-aggregated) to the bundle's hit list, which the shadow
// stage walks densely.
template <int VAR, int TB, int SR>
__global__ __launch_bounds__(TB) void k_trace(const float4* __restrict__ node, const float4* __restrict__ delta,
                                              const float4* __restrict__ tri,
                                              const float4* __restrict__ nrm, const float2* __restrict__ uvs,
                                              const float4* __restrict__ ori4, const float4* __restrict__ dir4,
                                              const uint32_t* __restrict__ dcount, uint64_t hcount,
                                              uint32_t* __restrict__ flags, float* __restrict__ tout,
                                              int32_t* __restrict__ idout, float2* __restrict__ bary,
                                              float4* __restrict__ n4, float4* __restrict__ p4,
                                              float2* __restrict__ uv2, uint32_t* __restrict__ hit_list,
                                              uint32_t* __restrict__ hit_count, uint32_t xcds, uint32_t coop,
                                              const uint32_t* __restrict__ parent, uint32_t bound) {
    __shared__ uint32_t s_id[VAR == 4 ? 1 : SR * TB];
    uint32_t s_low[stack_private_words<SR>()];
    TraversalStack<TB, SR> stk{s_id, threadIdx.x, s_low};
    const uint64_t n = ray_count(dcount, hcount);
    const uint32_t lane = threadIdx.x & 63;
    const XcdSpan sp = xcd_span(n, xcds, TB);
    for (uint64_t base = sp.begin; base < sp.end; base += sp.stride) {
        const uint64_t rr = base + threadIdx.x;
        const bool active = rr < sp.end;
        Closest best{0.f, 0.f, 0.f, -1};
        RayC rc{};
        if (active) {
            const float4 o = ori4[rr], dd = dir4[rr];
            rc = make_ray(xyz(o), xyz(dd), o.w);
            best.t = dd.w;
        }
        closest_var<VAR>(node, delta, tri, rc, best, stk, coop, active, parent, bound);  // whole wave
        const bool hit = active && best.id >= 0;
        const uint64_t hb = ballot(hit);
        uint32_t wbase = 0;
        if (hb) {
            if (lane == (uint32_t)__builtin_ctzll(hb)) wbase = atomicAdd(hit_count, (uint32_t)__builtin_popcountll(hb));
            wbase = __shfl(wbase, __builtin_ctzll(hb), RTX_WAVE);
        }
        if (!active) continue;
        if (!hit) {
            write_miss_record(rr, flags, tout, idout, bary, n4, p4, uv2);
            continue;
        }
        hit_list[wbase + lanes_below(hb)] = (uint32_t)rr;
        write_hit_record(tri, nrm, uvs, rc, best, rr, flags, tout, idout, bary, n4, p4, uv2);
    }
}

// Hit list from the flags (when hits were written from the host).
__global__ __launch_bounds__(256) void k_hit_list(const uint32_t* __restrict__ flags, const uint32_t* dcount,
                                                  uint64_t hcount, uint32_t* __restrict__ hit_list,
                                                  uint32_t* __restrict__ hit_count) {
    const uint64_t n = ray_count(dcount, hcount);
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t base = blockIdx.x * 256ull; base < n; base += (uint64_t)gridDim.x * 256) {
        const uint64_t rr = base + threadIdx.x;
        const bool hit = rr < n && (flags[rr] & 1u);
        const uint64_t hb = ballot(hit);
        uint32_t wbase = 0;
        if (hb) {
            if (lane == (uint32_t)__builtin_ctzll(hb)) wbase = atomicAdd(hit_count, (uint32_t)__builtin_popcountll(hb));
            wbase = __shfl(wbase, __builtin_ctzll(hb), RTX_WAVE);
        }
        if (hit) hit_list[wbase + lanes_below(hb)] = (uint32_t)rr;
    }
}

__global__ void k_zero1(uint32_t* p) { *p = 0u; }
__global__ void k_zero2(uint32_t* p) {
    p[0] = 0u;
    p[1] = 0u;
}

// ------------------------------------------------------------- shadow ----
// Hit lists of up to RTX_SEG_MAX bundles traced by one launch (the shadow
// rays of every bounce of a frame are independent of each other).
#define RTX_SEG_MAX 17
struct ShadowSeg {
    const uint32_t* list;
    const uint32_t* count;
    uint32_t* flags;
    const float4* p4;
};
struct ShadowSegs {
    ShadowSeg s[RTX_SEG_MAX];
    int n;
};

// Bit 8+l of a hit's flags = "light l is blocked".  The reference assigns
// shadow_hit on every call (src/kernel/shadow-trace-bvh.cl:402, 412), so a
// second shadow pass over the same hits (a moved light) must clear the bits
// of pairs that became lit, not only set new ones.  Lights of one hit run on
// different lanes, hence the atomics.
__device__ __forceinline__ void set_shadow_bit(uint32_t* f, int l, bool shadowed) {
    if (shadowed)
        atomicOr(f, 1u << (8 + l));
    else
        atomicAnd(f, ~(1u << (8 + l)));
}

// shadow_trace_single (src/kernel/shadow-trace-bvh.cl:375-418) for every
// (hit, light) pair of the hit lists: DIR toward -dir, SPOT toward pos
// (outside the cone = shadowed), POINT toward pos with tMax = distance; any
// hit with tMin 0.01 < t < tMax sets bit 8+l of the ray's flags.
template <int VAR, int TB, int SR>
__global__ __launch_bounds__(TB) void k_shadow(const float4* __restrict__ node, const float4* __restrict__ delta,
                                               const float4* __restrict__ tri, const rtx_lights* __restrict__ lights,
                                               ShadowSegs segs, unsigned long long* shadow_count, uint32_t xcds,
                                               uint32_t coop, const uint32_t* __restrict__ parent, uint32_t bound,
                                               uint32_t ordered) {
    __shared__ uint32_t s_id[VAR == 4 ? 1 : SR * TB];
    __shared__ uint32_t s_cnt;
    __shared__ uint64_t s_end[RTX_SEG_MAX];
    uint32_t s_low[stack_private_words<SR>()];
    TraversalStack<TB, SR> stk{s_id, threadIdx.x, s_low};
    const int nl = lights->count;
    if (threadIdx.x == 0) {
        s_cnt = 0;
        uint64_t acc = 0;
        for (int k = 0; k < segs.n; ++k) {
            acc += (uint64_t)(*segs.s[k].count) * (uint64_t)nl;
            s_end[k] = acc;
        }
    }
    __syncthreads();
    const uint64_t items = segs.n ? s_end[segs.n - 1] : 0;
    uint32_t traced = 0;
    const XcdSpan sp = xcd_span(items, xcds, TB);
    // wave-uniform loop: every lane reaches the traversal call (lanes past the
    // end or with nothing to trace pass start = false)
    for (uint64_t base = sp.begin; base < sp.end; base += sp.stride) {
        const uint64_t it = base + threadIdx.x;
        const bool valid = it < sp.end;
        uint32_t rr = 0;
        int l = 0;
        uint32_t* fl = nullptr;
        RayC rc{};
        float tmax = 1e37f;
        bool shadowed = false, trace = false;
        if (valid) {
            int sg = 0;
            while (it >= s_end[sg]) ++sg;
            const uint64_t local = it - (sg ? s_end[sg - 1] : 0);
            const ShadowSeg& S = segs.s[sg];
            const uint64_t nh = (s_end[sg] - (sg ? s_end[sg - 1] : 0)) / (uint64_t)nl;
            rr = S.list[local % nh];  // light-major: a wave traces one light's rays from neighbouring hits
            l = (int)(local / nh);
            fl = S.flags;
            const rtx_light& L = lights->l[l];
            const v3 p = xyz(S.p4[rr]);
            v3 dir = mk3(0.f, 0.f, 0.f);
            trace = true;
            if (L.type == RTX_LIGHT_DIR) {
                dir = -mk3(L.dir[0], L.dir[1], L.dir[2]);
            } else if (L.type == RTX_LIGHT_SPOT) {
                dir = normalize(mk3(L.pos[0], L.pos[1], L.pos[2]) - p);
                if (dot(-dir, mk3(L.dir[0], L.dir[1], L.dir[2])) < L.angle) {
                    shadowed = true;
                    trace = false;
                }
            } else if (L.type == RTX_LIGHT_POINT) {
                const v3 toL = mk3(L.pos[0], L.pos[1], L.pos[2]) - p;
                tmax = __builtin_sqrtf(dot(toL, toL));
                dir = normalize(toL);
            } else {
                shadowed = true;
                trace = false;
            }
            if (trace) {
                ++traced;
                rc = make_ray(p, dir, 0.01f);
            }
        }
        if (any_var<VAR>(node, delta, tri, rc, tmax, stk, coop, trace, parent, bound, ordered != 0))  // whole wave
            shadowed = true;
        if (valid) set_shadow_bit(&fl[rr], l, shadowed);
    }
    if (shadow_count) {
        if (traced) atomicAdd(&s_cnt, traced);
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt) atomicAdd(shadow_count, (unsigned long long)s_cnt);
    }
}

// ------------------------------------------- persistent traversal (3) --
// Variant 3: persistent waves with a ray queue.  A lane whose ray is done
// does not wait for the rest of its wave: once `refill` lanes of the wave are
// idle, their results are written and the wave takes that many new rays
// from the launch's queue counter with one atomic.  The lanes of a wave then
// stay busy until the queue drains, so a wave costs the sum of its rays'
// steps rather than 64 times the longest ray.  Each ray is still traversed
// alone by the same single-node steps (visiting order, box tests, leaf
// tests, cooperative fat leaves), so its result is the other variants'.

// One step of the single-node traversal for a lane with a ray: returns true
// when the ray's traversal is complete.  A fat leaf reached by this step is
// flagged in (la, lcnt) for the wave's cooperative test.
template <class Stack>
__device__ __forceinline__ bool step_closest(const float4* __restrict__ node, const float4* __restrict__ delta,
                                             const float4* __restrict__ tri, const RayC& r, Closest& best,
                                             Stack& stk, uint32_t& level, uint32_t& curr, uint32_t coop,
                                             uint32_t& la, uint32_t& lcnt) {
    const float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
    float4 dl4 = delta[curr];
    __asm__ volatile("" : "+v"(dl4.x), "+v"(dl4.y), "+v"(dl4.z));  // issued with the node load
    const uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
    if (slab(f0, f1, r, best.t)) {
        if (cnt >= coop) {
            la = a;
            lcnt = cnt;
        } else if (cnt) {
            leaf_closest(tri, a, cnt, r, best);
        } else {
            const bool nl = dot(r.d, xyz(dl4)) > 0.f;
            stk.put(level, nl ? a + 1 : a);
            level++;
            curr = nl ? a : a + 1;
            return false;
        }
    }
    if (level == 0 || level >= STACK) return true;  // no pop at MAX_LEVELS (src/kernel/trace-bvh.cl:361-367)
    curr = stk.get(--level);
    return false;
}

// Any-hit step: returns 1 when a hit was found, 2 when the traversal ended
// without one, 0 to continue.
template <class Stack>
__device__ __forceinline__ int step_any(const float4* __restrict__ node, const float4* __restrict__ delta,
                                        const float4* __restrict__ tri, const RayC& r, float tmax, Stack& stk,
                                        uint32_t& level, uint32_t& curr, uint32_t coop, uint32_t& la,
                                        uint32_t& lcnt) {
    const float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
    float4 dl4 = delta[curr];
    __asm__ volatile("" : "+v"(dl4.x), "+v"(dl4.y), "+v"(dl4.z));
    const uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
    if (slab(f0, f1, r, tmax)) {
        if (cnt >= coop) {
            la = a;
            lcnt = cnt;
        } else if (cnt) {
            if (leaf_any(tri, a, cnt, r, tmax)) return 1;
        } else {
            const bool nl = dot(r.d, xyz(dl4)) > 0.f;
            stk.put(level, nl ? a + 1 : a);
            level++;
            curr = nl ? a : a + 1;
            return 0;
        }
    }
    if (level == 0 || level >= STACK) return 2;  // no pop at MAX_LEVELS
    curr = stk.get(--level);
    return 0;
}

// Per-wave view of the launch's item queue.  The wave takes QUEUE_CHUNK
// consecutive items per atomic (consecutive rays stay together, and the
// single counter sees one atomic per chunk) and hands them to its idle lanes
// from a wave-uniform cursor.
constexpr uint32_t QUEUE_CHUNK = 64;
struct WaveQueue {
    uint64_t cur = 0, end = 0;
    bool last = false;  // no further chunk exists
    __device__ __forceinline__ bool exhausted() const { return last && cur >= end; }
};

// Item for this lane if it is idle and the queue still has one, else ~0.
__device__ __forceinline__ uint64_t wq_take(uint32_t* queue, WaveQueue& q, uint64_t idle, uint64_t n) {
    if (q.cur >= q.end && !q.last) {
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(queue, QUEUE_CHUNK);
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
        q.cur = base;
        q.end = (uint64_t)base + QUEUE_CHUNK < n ? (uint64_t)base + QUEUE_CHUNK : n;
        if (q.end < q.cur) q.end = q.cur;
        if ((uint64_t)base + QUEUE_CHUNK >= n) q.last = true;
    }
    const uint64_t avail = q.end - q.cur;
    const uint64_t k = (uint64_t)__builtin_popcountll(idle);
    const uint64_t give = k < avail ? k : avail;
    const uint64_t r = lanes_below(idle);
    const uint64_t item = ((idle >> lane_id()) & 1ull) && r < give ? q.cur + r : ~0ull;
    q.cur += give;
    return item;
}

// Hit-list appends batched per wave in LDS: one atomic per flush.
constexpr uint32_t HITBUF = 128;
__device__ __forceinline__ void hitbuf_flush(uint32_t* buf, uint32_t& cnt, uint32_t* __restrict__ hit_list,
                                             uint32_t* hit_count) {
    if (cnt == 0) return;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(hit_count, cnt);
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
    for (uint32_t i = lane_id(); i < cnt; i += RTX_WAVE) hit_list[base + i] = buf[i];
    cnt = 0;
}

template <int TB>
__global__ __launch_bounds__(TB) void k_trace_p(const float4* __restrict__ node, const float4* __restrict__ delta,
                                                const float4* __restrict__ tri, const float4* __restrict__ nrm,
                                                const float2* __restrict__ uvs, const float4* __restrict__ ori4,
                                                const float4* __restrict__ dir4, const uint32_t* __restrict__ dcount,
                                                uint64_t hcount, uint32_t* __restrict__ flags,
                                                float* __restrict__ tout, int32_t* __restrict__ idout,
                                                float2* __restrict__ bary, float4* __restrict__ n4,
                                                float4* __restrict__ p4, float2* __restrict__ uv2,
                                                uint32_t* __restrict__ hit_list, uint32_t* __restrict__ hit_count,
                                                uint32_t coop, uint32_t refill) {
    __shared__ uint32_t s_id[STACK * TB];
    __shared__ uint32_t s_hits[TB / RTX_WAVE][HITBUF];
    LdsStack<TB> stk{s_id, threadIdx.x, nullptr};
    uint32_t* hbuf = s_hits[threadIdx.x / RTX_WAVE];
    uint32_t hcnt = 0;  // wave-uniform fill of hbuf
    const uint64_t n = ray_count(dcount, hcount);
    uint32_t* queue = hit_count + 1;  // zeroed with the hit count before the launch
    WaveQueue q;
    bool have = false, done = false;
    uint64_t rr = 0;
    uint32_t level = 0, curr = 0;
    Closest best{0.f, 0.f, 0.f, -1};
    RayC rc{};
    for (;;) {
        const uint64_t idle = ballot(!have);
        if (idle == ~0ull || (!q.exhausted() && (uint32_t)__builtin_popcountll(idle) >= refill)) {
            // results of the lanes that finished since the last refill
            const bool hit = done && best.id >= 0;
            const uint64_t hb = ballot(hit);
            if (hb) {
                if (hcnt + RTX_WAVE > HITBUF) hitbuf_flush(hbuf, hcnt, hit_list, hit_count);
                if (hit) {
                    hbuf[hcnt + lanes_below(hb)] = (uint32_t)rr;
                    write_hit_record(tri, nrm, uvs, rc, best, rr, flags, tout, idout, bary, n4, p4, uv2);
                }
                hcnt += (uint32_t)__builtin_popcountll(hb);
            }
            if (done && !hit) write_miss_record(rr, flags, tout, idout, bary, n4, p4, uv2);
            done = false;
            const uint64_t my = wq_take(queue, q, idle, n);
            if (!have && my < n) {
                rr = my;
                const float4 o = ori4[rr], dd = dir4[rr];
                rc = make_ray(xyz(o), xyz(dd), o.w);
                best = Closest{dd.w, 0.f, 0.f, -1};
                level = 0;
                stk.reset();
                curr = 0;
                have = true;
            }
            if (q.exhausted() && !ballot(have)) break;
        }
        uint32_t la = 0, lcnt = 0;
        if (have && step_closest(node, delta, tri, rc, best, stk, level, curr, coop, la, lcnt)) {
            have = false;
            done = true;
        }
        coop_closest(tri, la, lcnt, rc, best, ~0ull);  // whole wave
    }
    hitbuf_flush(hbuf, hcnt, hit_list, hit_count);
}

template <int TB>
__global__ __launch_bounds__(TB) void k_shadow_p(const float4* __restrict__ node, const float4* __restrict__ delta,
                                                 const float4* __restrict__ tri,
                                                 const rtx_lights* __restrict__ lights, ShadowSegs segs,
                                                 unsigned long long* shadow_count, uint32_t* queue, uint32_t coop,
                                                 uint32_t refill) {
    __shared__ uint32_t s_id[STACK * TB];
    __shared__ uint32_t s_cnt;
    __shared__ uint64_t s_end[RTX_SEG_MAX];
    LdsStack<TB> stk{s_id, threadIdx.x, nullptr};
    const int nl = lights->count;
    if (threadIdx.x == 0) {
        s_cnt = 0;
        uint64_t acc = 0;
        for (int k = 0; k < segs.n; ++k) {
            acc += (uint64_t)(*segs.s[k].count) * (uint64_t)nl;
            s_end[k] = acc;
        }
    }
    __syncthreads();
    const uint64_t items = segs.n ? s_end[segs.n - 1] : 0;
    uint32_t traced = 0;
    bool have = false;
    WaveQueue q;
    uint32_t rr = 0, level = 0, curr = 0;
    int l = 0;
    uint32_t* fl = nullptr;
    RayC rc{};
    float tmax = 1e37f;
    for (;;) {
        const uint64_t idle = ballot(!have);
        if (idle == ~0ull || (!q.exhausted() && (uint32_t)__builtin_popcountll(idle) >= refill)) {
            {
                const uint64_t it = wq_take(queue, q, idle, items);
                if (!have && it < items) {
                    int sg = 0;
                    while (it >= s_end[sg]) ++sg;
                    const uint64_t local = it - (sg ? s_end[sg - 1] : 0);
                    const ShadowSeg& S = segs.s[sg];
                    const uint64_t nh = (s_end[sg] - (sg ? s_end[sg - 1] : 0)) / (uint64_t)nl;
                    rr = S.list[local % nh];
                    l = (int)(local / nh);
                    fl = S.flags;
                    const rtx_light& L = lights->l[l];
                    const v3 p = xyz(S.p4[rr]);
                    v3 dir = mk3(0.f, 0.f, 0.f);
                    bool trace = true, shadowed = false;
                    tmax = 1e37f;
                    if (L.type == RTX_LIGHT_DIR) {
                        dir = -mk3(L.dir[0], L.dir[1], L.dir[2]);
                    } else if (L.type == RTX_LIGHT_SPOT) {
                        dir = normalize(mk3(L.pos[0], L.pos[1], L.pos[2]) - p);
                        if (dot(-dir, mk3(L.dir[0], L.dir[1], L.dir[2])) < L.angle) {
                            shadowed = true;
                            trace = false;
                        }
                    } else if (L.type == RTX_LIGHT_POINT) {
                        const v3 toL = mk3(L.pos[0], L.pos[1], L.pos[2]) - p;
                        tmax = __builtin_sqrtf(dot(toL, toL));
                        dir = normalize(toL);
                    } else {
                        shadowed = true;
                        trace = false;
                    }
                    if (!trace) set_shadow_bit(&fl[rr], l, shadowed);
                    if (trace) {
                        ++traced;
                        rc = make_ray(p, dir, 0.01f);
                        level = 0;
                        stk.reset();
                        curr = 0;
                        have = true;
                    }
                }
            }
            if (q.exhausted() && !ballot(have)) break;
        }
        uint32_t la = 0, lcnt = 0;
        if (have) {
            const int st = step_any(node, delta, tri, rc, tmax, stk, level, curr, coop, la, lcnt);
            if (st != 0) {
                set_shadow_bit(&fl[rr], l, st == 1);
                have = false;
            }
        }
        if (coop_any(tri, la, lcnt, rc, tmax, ~0ull)) {  // whole wave; true on owners with a hit
            set_shadow_bit(&fl[rr], l, true);
            have = false;
        }
    }
    if (shadow_count) {
        if (traced) atomicAdd(&s_cnt, traced);
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt) atomicAdd(shadow_count, (unsigned long long)s_cnt);
    }
}

// Launch macro M(variant, block) with the tuned variant and block size.
#define RTX_DISPATCH_TRAVERSAL(M, var, tb)     \
    do {                                       \
        if (tb == 64) {                        \
            if (var == 0) M(0, 64);            \
            else if (var == 1) M(1, 64);       \
            else M(2, 64);                     \
        } else if (tb == 256) {                \
            if (var == 0) M(0, 256);           \
            else if (var == 1) M(1, 256);      \
            else M(2, 256);                    \
        } else {                               \
            if (var == 0) M(0, 128);           \
            else if (var == 1) M(1, 128);      \
            else M(2, 128);                    \
        }                                      \
    } while (0)

// Launch macro M(variant, block, LDS stack entries) for the kernels with a
// ring-stack form: the tuned variant, block size and stack_lds (8, 16, or
// 32 = the whole stack in LDS).
#define RTX_DISPATCH_VB(M, var, tb, S)            \
    do {                                          \
        if (tb == 64) {                           \
            if (var == 0) M(0, 64, S);            \
            else if (var == 1) M(1, 64, S);       \
            else M(2, 64, S);                     \
        } else if (tb == 256) {                   \
            if (var == 0) M(0, 256, S);           \
            else if (var == 1) M(1, 256, S);      \
            else M(2, 256, S);                    \
        } else {                                  \
            if (var == 0) M(0, 128, S);           \
            else if (var == 1) M(1, 128, S);      \
            else M(2, 128, S);                    \
        }                                         \
    } while (0)
#define RTX_DISPATCH_STACK(M, var, tb, sr)                  \
    do {                                                    \
        if (sr == 8) RTX_DISPATCH_VB(M, var, tb, 8);        \
        else if (sr == 16) RTX_DISPATCH_VB(M, var, tb, 16); \
        else RTX_DISPATCH_VB(M, var, tb, 32);               \
    } while (0)
// Scene kernels: also the stackless form (4), which has no stack.
#define RTX_DISPATCH_STACK_NS(M, var, tb, sr)   \
    do {                                        \
        if (var == 4) {                         \
            if (tb == 64) M(4, 64, 32);         \
            else if (tb == 256) M(4, 256, 32);  \
            else M(4, 128, 32);                 \
        } else {                                \
            RTX_DISPATCH_STACK(M, var, tb, sr); \
        }                                       \
    } while (0)

// Persistent traversal grid: enough groups to fill every CU at 20 waves
// (the LDS stack allows 5 per SIMD), fewer when the items need fewer.
uint32_t persist_grid(rtx_ctx* c, uint64_t count, uint64_t cap, int TB) {
    uint64_t n = count == RTX_DEVICE_COUNT ? cap : count;
    uint64_t g = (n + TB - 1) / TB;
    const uint64_t gmax = (uint64_t)c->cus * (1280 / TB);
    if (g > gmax) g = gmax;
    return (uint32_t)(g ? g : 1);
}

// Leaf size from which leaves are tested wave-cooperatively (0 = never).
uint32_t coop_threshold(const rtx_ctx* c) {
    return c->tune.coop_leaf > 0 ? (uint32_t)c->tune.coop_leaf : 0xffffffffu;
}

// ------------------------------------------------------------ instances --
// Multi-root traversal (the reference's trace_multi / shadow_trace_multi,
// src/kernel/trace-bvh.cl:419-463, src/kernel/shadow-trace-bvh.cl:317-373):
// every instance is a built scene with an object-to-world transform `tr` and
// its inverse, held in an InstRoot.

// multiply_point / multiply_vector: rows dotted with (v, 1) or (v, 0), then
// a division by w when |w| > wmin (1e-26 in trace-bvh.cl:83-113, 0.00001 in
// shadow-trace-bvh.cl:107-143).  The dot is the sequential sum.  A division
// by w == 1 returns its dividend exactly, so it is skipped (affine rows).
__device__ __forceinline__ v3 mul_h(const float* __restrict__ M, v3 v, float vw, float wmin) {
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = ((M[4 * i] * v.x + M[4 * i + 1] * v.y) + M[4 * i + 2] * v.z) + M[4 * i + 3] * vw;
    if (fabsf(r[3]) > wmin && r[3] != 1.f) {
        r[0] = r[0] / r[3];
        r[1] = r[1] / r[3];
        r[2] = r[2] / r[3];
    }
    return mk3(r[0], r[1], r[2]);
}

// The record of ray rr after the root loop (complete_transformed_hit_info):
// hit point along the world ray, normal = tr x the interpolated object
// normal, turned to face the ray; ids in the shared numbering.  Hits join
// the hit list.  Called by every lane of the wave.
__device__ __forceinline__ void inst_record(const InstRoot* __restrict__ roots, uint64_t rr, bool active, v3 o, v3 d,
                                            float bt, float bu, float bv, int32_t bid, int32_t broot,
                                            uint32_t* __restrict__ flags, float* __restrict__ tout,
                                            int32_t* __restrict__ idout, float2* __restrict__ bary,
                                            float4* __restrict__ n4, float4* __restrict__ p4,
                                            float2* __restrict__ uv2, uint32_t* __restrict__ hit_list,
                                            uint32_t* __restrict__ hit_count, uint32_t* __restrict__ hit_root) {
    const bool hit = active && bid >= 0;
    const uint64_t hb = ballot(hit);
    uint32_t wbase = 0;
    if (hb) {
        if (lane_id() == (uint32_t)__builtin_ctzll(hb)) wbase = atomicAdd(hit_count, (uint32_t)__builtin_popcountll(hb));
        wbase = __shfl(wbase, __builtin_ctzll(hb), RTX_WAVE);
    }
    if (!active) return;
    if (!hit) {
        write_miss_record(rr, flags, tout, idout, bary, n4, p4, uv2);
        return;
    }
    hit_list[wbase + lanes_below(hb)] = (uint32_t)rr;
    hit_root[rr] = (uint32_t)broot;
    const InstRoot& R = roots[broot];
    const v3 hp = o + d * bt;
    const uint32_t i0 = __float_as_uint(R.tri[3 * (size_t)bid].w);
    const uint32_t i1 = __float_as_uint(R.tri[3 * (size_t)bid + 1].w);
    const uint32_t i2 = __float_as_uint(R.tri[3 * (size_t)bid + 2].w);
    const float w = 1.f - (bu + bv);
    const v3 n0 = xyz(R.nrm[i0]), n1 = xyz(R.nrm[i1]), n2 = xyz(R.nrm[i2]);
    v3 nn = mul_h(R.tr, normalize(w * n0 + bv * n2 + bu * n1), 0.f, 1e-26f);
    uint32_t fl = 1u;
    if (dot(nn, d) > 0.f) {
        fl |= 2u;
        nn = -nn;
    }
    const float2 t0 = R.uv[i0], t1 = R.uv[i1], t2 = R.uv[i2];
    flags[rr] = fl;
    tout[rr] = bt;
    idout[rr] = bid + (int32_t)R.base;
    bary[rr] = make_float2(bu, bv);
    n4[rr] = make_float4(nn.x, nn.y, nn.z, 0.f);
    p4[rr] = make_float4(hp.x, hp.y, hp.z, 0.f);
    uv2[rr] = make_float2(w * t0.x + bv * t2.x + bu * t1.x, w * t0.y + bv * t2.y + bu * t1.y);
}

// Closest hit over the roots: each root traced from the ray's own tMax in
// its frame; the root's hit point taken back to the world gives t (its
// distance from the origin); the first root with the smallest t wins.  The
// record is completed as complete_transformed_hit_info does: hit point
// along the world ray, normal = tr x the interpolated object normal, turned
// to face the ray.  Ids are in the shared numbering.
template <int VAR, int TB>
__global__ __launch_bounds__(TB) void k_trace_inst(const InstRoot* __restrict__ roots, int nroots,
                                                    const float4* __restrict__ ori4, const float4* __restrict__ dir4,
                                                    const uint32_t* __restrict__ dcount, uint64_t hcount,
                                                    uint32_t* __restrict__ flags, float* __restrict__ tout,
                                                    int32_t* __restrict__ idout, float2* __restrict__ bary,
                                                    float4* __restrict__ n4, float4* __restrict__ p4,
                                                    float2* __restrict__ uv2, uint32_t* __restrict__ hit_list,
                                                    uint32_t* __restrict__ hit_count, uint32_t* __restrict__ hit_root,
                                                    uint32_t coop) {
    __shared__ uint32_t s_id[STACK * TB];
    LdsStack<TB> stk{s_id, threadIdx.x, nullptr};
    const uint64_t n = ray_count(dcount, hcount);
    for (uint64_t base = (uint64_t)blockIdx.x * TB; base < n; base += (uint64_t)gridDim.x * TB) {
        const uint64_t rr = base + threadIdx.x;
        const bool active = rr < n;
        v3 o = mk3(0.f, 0.f, 0.f), d = o;
        float tmin = 0.f, tmax = 0.f;
        if (active) {
            const float4 o4 = ori4[rr], d4 = dir4[rr];
            o = xyz(o4);
            d = xyz(d4);
            tmin = o4.w;
            tmax = d4.w;
        }
        float bt = 0.f, bu = 0.f, bv = 0.f;
        int32_t bid = -1, broot = 0;
        for (int i = 0; i < nroots; ++i) {  // wave-uniform
            const InstRoot& R = roots[i];
            const v3 to = mul_h(R.tr_inv, o, 1.f, 1e-26f), td = mul_h(R.tr_inv, d, 0.f, 1e-26f);
            const RayC rc = make_ray(to, td, tmin);
            Closest c{tmax, 0.f, 0.f, -1};
            closest_var<VAR>(R.node, R.delta, R.tri, rc, c, stk, coop, active);  // whole wave
            if (active && c.id >= 0) {
                const v3 hw = mul_h(R.tr, to + td * c.t, 1.f, 1e-26f);
                const v3 dv = hw - o;
                const float tw = __builtin_sqrtf(dot(dv, dv));
                if (bid < 0 || tw < bt) {
                    bt = tw;
                    bid = c.id;
                    bu = c.u;
                    bv = c.v;
                    broot = i;
                }
            }
        }
        inst_record(roots, rr, active, o, d, bt, bu, bv, bid, broot, flags, tout, idout, bary, n4, p4, uv2, hit_list,
                    hit_count, hit_root);
    }
}

// Shadow rays of every (hit, light) pair of a hit list against every root:
// DIR toward -dir; SPOT toward pos, lit without a trace when dot(toward the
// light, spot dir) < angle (that kernel's test); POINT toward pos with tMax =
// distance (the lights extension); any hit in any root sets bit 8 + l.
// The result is an OR over the roots, so their order is free: a wave starts
// at the root its first item's hit came from (`hit_root`, the closest-hit
// kernel's record; any value only reorders), where an occluded point is
// usually blocked by its own object, and stops once every ray is blocked.
template <int VAR, int TB>
__global__ __launch_bounds__(TB) void k_shadow_inst(const InstRoot* __restrict__ roots, int nroots,
                                                     const rtx_lights lights,
                                                     const uint32_t* __restrict__ list,
                                                     const uint32_t* __restrict__ list_count,
                                                     uint32_t* __restrict__ flags, const float4* __restrict__ p4,
                                                     const uint32_t* __restrict__ hit_root,
                                                     unsigned long long* shadow_count, uint32_t coop,
                                                     uint32_t ordered) {
    __shared__ uint32_t s_id[STACK * TB];
    __shared__ uint32_t s_cnt;
    LdsStack<TB> stk{s_id, threadIdx.x, nullptr};
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const int nl = lights.count;
    const uint64_t items = (uint64_t)(*list_count) * (uint64_t)nl;
    uint32_t traced = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * TB; base < items; base += (uint64_t)gridDim.x * TB) {
        const uint64_t it = base + threadIdx.x;
        const bool valid = it < items;
        uint32_t rr = 0;
        int l = 0;
        v3 p = mk3(0.f, 0.f, 0.f), dir = p;
        float tmax = 1e37f;
        bool trace = false;
        if (valid) {
            rr = list[it / (uint64_t)nl];
            l = (int)(it % (uint64_t)nl);
            const rtx_light& L = lights.l[l];
            p = xyz(p4[rr]);
            trace = true;
            if (L.type == RTX_LIGHT_DIR) {
                dir = -mk3(L.dir[0], L.dir[1], L.dir[2]);
            } else if (L.type == RTX_LIGHT_SPOT) {
                dir = normalize(mk3(L.pos[0], L.pos[1], L.pos[2]) - p);
                if (dot(dir, mk3(L.dir[0], L.dir[1], L.dir[2])) < L.angle) trace = false;
            } else if (L.type == RTX_LIGHT_POINT) {
                const v3 toL = mk3(L.pos[0], L.pos[1], L.pos[2]) - p;
                tmax = __builtin_sqrtf(dot(toL, toL));
                dir = normalize(toL);
            } else {
                trace = false;
            }
            if (trace) ++traced;
        }
        bool shadowed = false;
        uint32_t mine = 0;
        if (valid && hit_root) mine = hit_root[rr] % (uint32_t)nroots;
        const int first = __builtin_amdgcn_readfirstlane((int)mine);
        for (int k = 0, i = first; k < nroots; ++k, i = (i + 1 == nroots) ? 0 : i + 1) {  // wave-uniform
            const InstRoot& R = roots[i];
            const bool go = trace && !shadowed;
            if (!ballot(go)) break;
            RayC rc{};
            if (go) rc = make_ray(mul_h(R.tr_inv, p, 1.f, 0.00001f), mul_h(R.tr_inv, dir, 0.f, 0.00001f), 0.01f);
            if (any_var<VAR>(R.node, R.delta, R.tri, rc, tmax, stk, coop, go, nullptr, 0, ordered != 0))
                shadowed = true;  // whole wave
        }
        if (valid) set_shadow_bit(&flags[rr], l, shadowed);
    }
    if (shadow_count) {
        if (traced) atomicAdd(&s_cnt, traced);
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt) atomicAdd(shadow_count, (unsigned long long)s_cnt);
    }
}

// World-space boxes of the roots for culling (one thread per root): the
// root node's box under tr (its 8 corners, rows summed in order), the box's
// magnitude s, and k = 1e-4 |tr| |tr_inv| (infinity norms).  A ray is
// tested against the box widened by k (2 |o| + s): the rounding of the ray's
// transform into the object's frame and of the hit point's way back (a few
// units of 6e-8, times the transforms' condition, times the magnitudes of
// the origin and of the point, which lies within the box) stays far below
// it, so a root whose widened box the ray misses holds no hit the walk
// could find.  Roots under a projective transform (bottom rows other than
// 0 0 0 1), or with a non-finite box, get k = -1: never culled.
__global__ void k_root_boxes(const InstRoot* __restrict__ roots, int n, RootBox* __restrict__ out) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    const InstRoot& R = roots[k];
    const float* M = R.tr;
    const float* Mi = R.tr_inv;
    const bool affine = M[12] == 0.f && M[13] == 0.f && M[14] == 0.f && M[15] == 1.f && Mi[12] == 0.f &&
                        Mi[13] == 0.f && Mi[14] == 0.f && Mi[15] == 1.f;
    const float4 b0 = R.node[0], b1 = R.node[1];
    RootBox b;
    float nm = 1.f, nmi = 1.f;
    for (int r = 0; r < 3; ++r) {
        b.lo[r] = INFINITY;
        b.hi[r] = -INFINITY;
        nm = fmaxf(nm, fabsf(M[4 * r]) + fabsf(M[4 * r + 1]) + fabsf(M[4 * r + 2]) + fabsf(M[4 * r + 3]));
        nmi = fmaxf(nmi, fabsf(Mi[4 * r]) + fabsf(Mi[4 * r + 1]) + fabsf(Mi[4 * r + 2]) + fabsf(Mi[4 * r + 3]));
    }
    for (int c = 0; c < 8; ++c) {
        const float x = (c & 1) ? b1.x : b0.x, y = (c & 2) ? b1.y : b0.y, z = (c & 4) ? b1.z : b0.z;
        for (int r = 0; r < 3; ++r) {
            const float w = ((M[4 * r] * x + M[4 * r + 1] * y) + M[4 * r + 2] * z) + M[4 * r + 3];
            b.lo[r] = fminf(b.lo[r], w);
            b.hi[r] = fmaxf(b.hi[r], w);
        }
    }
    b.s = 0.f;
    bool finite = true;
    for (int r = 0; r < 3; ++r) {
        b.s = fmaxf(b.s, fmaxf(fabsf(b.lo[r]), fabsf(b.hi[r])));
        finite = finite && isfinite(b.lo[r]) && isfinite(b.hi[r]);
    }
    b.k = 1e-4f * nm * nmi;
    if (!affine || !finite || !isfinite(b.k) || !isfinite(b.s)) b.k = -1.f;
    out[k] = b;
}

// Bit k set: the segment [tmin, tmax] of the ray (o, d) meets root k's
// widened world box, or the root is never culled.  n <= 64.  The loop is
// wave-uniform (the boxes are read as scalars).
__device__ __forceinline__ uint64_t root_candidates(const RootBox* __restrict__ wb, int n, v3 o, v3 d, float tmin,
                                                    float tmax) {
    const float io[3] = {o.x, o.y, o.z}, id[3] = {d.x, d.y, d.z};
    float inv[3];
    for (int a = 0; a < 3; ++a) inv[a] = id[a] != 0.f ? 1.f / id[a] : 0.f;
    const float mo = fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z)));
    uint64_t m = 0;
    for (int k = 0; k < n; ++k) {
        const RootBox b = wb[k];
        bool meet = true;
        if (b.k >= 0.f) {
            const float g = b.k * (2.f * mo + b.s);
            float t0 = tmin, t1 = tmax;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float l = b.lo[a] - g, h = b.hi[a] + g;
                if (id[a] == 0.f) {
                    meet = meet && io[a] >= l && io[a] <= h;
                } else {
                    const float ta = (l - io[a]) * inv[a], tb = (h - io[a]) * inv[a];
                    t0 = fmaxf(t0, fminf(ta, tb));
                    t1 = fminf(t1, fmaxf(ta, tb));
                }
            }
            meet = meet && t0 <= t1;
        }
        if (meet) m |= 1ull << k;
    }
    return m;
}

// Lane-roots form of the two kernels above for the single-node walks (SPEC
// as in traverse_closest_node).  The root loop and the node loop are one
// loop: a lane whose walk of a root ends enters its next root in the next
// step, instead of waiting for the wave's slowest walk of that root, which
// serializes incoherent rays (bounces) over the roots.  Each root is walked
// as traverse_closest_node / traverse_any_node walk it (the same visits,
// stack and MAX_LEVELS rule) and the closest-hit kernel merges the roots in
// the same order with the same test, so the records are those of the
// wave-uniform loop.  A fat leaf flagged in a step is tested by the wave
// (with the owner's triangle pointer) before the owner leaves its root.
// Entering a root is one step like any other: a root the ray misses is left
// at its first node test, in the same step as the other lanes' visits.
template <bool SPEC, bool CHUNKS, int TB>
__global__ __launch_bounds__(TB) void k_trace_inst_lr(const InstRoot* __restrict__ roots, int nroots,
                                                       const float4* __restrict__ ori4,
                                                       const float4* __restrict__ dir4,
                                                       const uint32_t* __restrict__ dcount, uint64_t hcount,
                                                       uint32_t* __restrict__ flags, float* __restrict__ tout,
                                                       int32_t* __restrict__ idout, float2* __restrict__ bary,
                                                       float4* __restrict__ n4, float4* __restrict__ p4,
                                                       float2* __restrict__ uv2, uint32_t* __restrict__ hit_list,
                                                       uint32_t* __restrict__ hit_count,
                                                       uint32_t* __restrict__ hit_root, uint32_t coop,
                                                       const RootBox* __restrict__ wbox) {
    __shared__ uint32_t s_id[STACK * TB];
    LdsStack<TB> stk{s_id, threadIdx.x, nullptr};
    const uint64_t n = ray_count(dcount, hcount);
    const bool cull = wbox != nullptr && (CHUNKS || nroots <= 64);
    for (uint64_t base = (uint64_t)blockIdx.x * TB; base < n; base += (uint64_t)gridDim.x * TB) {
        const uint64_t rr = base + threadIdx.x;
        const bool active = rr < n;
        v3 o = mk3(0.f, 0.f, 0.f), d = o;
        float tmin = 0.f, tmax = 0.f;
        if (active) {
            const float4 o4 = ori4[rr], d4 = dir4[rr];
            o = xyz(o4);
            d = xyz(d4);
            tmin = o4.w;
            tmax = d4.w;
        }
        float bt = 0.f, bu = 0.f, bv = 0.f;
        int32_t bid = -1, broot = 0;
        // the lane's current root i: its tree, the ray in its frame, its best
        // hit.  The roots in ascending order; with culling, those in cand,
        // over chunks of 64 roots when CHUNKS (the chunk loop is
        // wave-uniform; sets of up to 64 roots use the kernel without it).
        int i = -1;
        const float4 *node = nullptr, *delta = nullptr, *tri = nullptr;
        RayC rc{};
        Closest c{tmax, 0.f, 0.f, -1};
        uint32_t level = 0, curr = 0;
        auto enter = [&](int k) {
            const InstRoot& R = roots[k];
            node = R.node;
            delta = R.delta;
            tri = R.tri;
            rc = make_ray(mul_h(R.tr_inv, o, 1.f, 1e-26f), mul_h(R.tr_inv, d, 0.f, 1e-26f), tmin);
            c = Closest{tmax, 0.f, 0.f, -1};
            level = 0;
            stk.reset();
            curr = 0;
        };
        const uint64_t live = ballot(true);
        auto walk = [&](int c0, int cn) {
            uint64_t cand = cull ? root_candidates(wbox + c0, cn, o, d, tmin, tmax) : 0;
            i = c0 - 1;
            auto next = [&]() -> bool {
                if (cull) {
                    if (!cand) return false;
                    i = c0 + __builtin_ctzll(cand);
                    cand &= cand - 1;
                    return true;
                }
                return ++i < nroots;
            };
            bool going = active && next();
            if (going) enter(i);
            while (ballot(going)) {
                uint32_t la = 0, lcnt = 0;
                bool done = false;
                if (going) {
                    const float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
                    float4 dl4 = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (SPEC) {
                        dl4 = delta[curr];
                        __asm__ volatile("" : "+v"(dl4.x), "+v"(dl4.y), "+v"(dl4.z));  // see traverse_closest_node
                    }
                    const uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
                    bool pop = true;
                    if (slab(f0, f1, rc, c.t)) {
                        if (cnt >= coop) {
                            la = a;
                            lcnt = cnt;
                        } else if (cnt) {
                            leaf_closest(tri, a, cnt, rc, c);
                        } else {
                            if (!SPEC) dl4 = delta[curr];
                            const bool nl = dot(rc.d, xyz(dl4)) > 0.f;
                            stk.put(level, nl ? a + 1 : a);
                            level++;
                            curr = nl ? a : a + 1;
                            pop = false;
                        }
                    }
                    if (pop) {
                        if (level == 0 || level >= STACK)  // the walk of root i ends (MAX_LEVELS as trace_ray)
                            done = true;
                        else
                            curr = stk.get(--level);
                    }
                }
                coop_closest<true>(tri, la, lcnt, rc, c, live);
                if (done) {
                    if (c.id >= 0) {
                        const v3 hw = mul_h(roots[i].tr, rc.o + rc.d * c.t, 1.f, 1e-26f);
                        const v3 dv = hw - o;
                        const float tw = __builtin_sqrtf(dot(dv, dv));
                        if (bid < 0 || tw < bt) {
                            bt = tw;
                            bid = c.id;
                            bu = c.u;
                            bv = c.v;
                            broot = i;
                        }
                    }
                    if (next())
                        enter(i);
                    else
                        going = false;
                }
            }
        };
        if (CHUNKS) {
            for (int c0 = 0; c0 < nroots; c0 += 64) walk(c0, nroots - c0 < 64 ? nroots - c0 : 64);
        } else {
            walk(0, nroots);
        }
        inst_record(roots, rr, active, o, d, bt, bu, bv, bid, broot, flags, tout, idout, bary, n4, p4, uv2, hit_list,
                    hit_count, hit_root);
    }
}

template <bool SPEC, bool CHUNKS, int TB>
__global__ __launch_bounds__(TB) void k_shadow_inst_lr(const InstRoot* __restrict__ roots, int nroots,
                                                        const rtx_lights lights,
                                                        const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ list_count,
                                                        uint32_t* __restrict__ flags, const float4* __restrict__ p4,
                                                        const uint32_t* __restrict__ hit_root,
                                                        unsigned long long* shadow_count, uint32_t coop,
                                                        const RootBox* __restrict__ wbox, uint32_t ordered) {
    __shared__ uint32_t s_id[STACK * TB];
    __shared__ uint32_t s_cnt;
    LdsStack<TB> stk{s_id, threadIdx.x, nullptr};
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const bool cull = wbox != nullptr && (CHUNKS || nroots <= 64);
    const int nl = lights.count;
    const uint64_t items = (uint64_t)(*list_count) * (uint64_t)nl;
    uint32_t traced = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * TB; base < items; base += (uint64_t)gridDim.x * TB) {
        const uint64_t it = base + threadIdx.x;
        const bool valid = it < items;
        uint32_t rr = 0;
        int l = 0;
        v3 p = mk3(0.f, 0.f, 0.f), dir = p;
        float tmax = 1e37f;
        bool trace = false;
        if (valid) {
            rr = list[it / (uint64_t)nl];
            l = (int)(it % (uint64_t)nl);
            const rtx_light& L = lights.l[l];
            p = xyz(p4[rr]);
            trace = true;
            if (L.type == RTX_LIGHT_DIR) {
                dir = -mk3(L.dir[0], L.dir[1], L.dir[2]);
            } else if (L.type == RTX_LIGHT_SPOT) {
                dir = normalize(mk3(L.pos[0], L.pos[1], L.pos[2]) - p);
                if (dot(dir, mk3(L.dir[0], L.dir[1], L.dir[2])) < L.angle) trace = false;
            } else if (L.type == RTX_LIGHT_POINT) {
                const v3 toL = mk3(L.pos[0], L.pos[1], L.pos[2]) - p;
                tmax = __builtin_sqrtf(dot(toL, toL));
                dir = normalize(toL);
            } else {
                trace = false;
            }
            if (trace) ++traced;
        }
        // the lane's roots: all of them once (k counts them), an OR over the
        // roots, so the order changes no bit: from the one its hit came from,
        // or, with culling, those in cand, its hit's root first (over chunks
        // of 64 roots when CHUNKS; the chunk loop is wave-uniform)
        int i = 0, i0 = 0;
        if (trace && hit_root && nroots > 0) i0 = (int)(hit_root[rr] % (uint32_t)nroots);
        const float4 *node = nullptr, *delta = nullptr, *tri = nullptr;
        RayC rc{};
        uint32_t level = 0, curr = 0;
        auto enter = [&](int j) {
            const InstRoot& R = roots[j];
            node = R.node;
            delta = R.delta;
            tri = R.tri;
            rc = make_ray(mul_h(R.tr_inv, p, 1.f, 0.00001f), mul_h(R.tr_inv, dir, 0.f, 0.00001f), 0.01f);
            level = 0;
            stk.reset();
            curr = 0;
        };
        bool shadowed = false;
        const uint64_t live = ballot(true);
        auto walk = [&](int c0, int cn) {
            uint64_t cand = cull ? root_candidates(wbox + c0, cn, p, dir, 0.01f, tmax) : 0;
            int k = 0;
            auto next = [&]() -> bool {
                if (cull) {
                    if (!cand) return false;
                    const int j0 = i0 - c0;
                    const bool own = k == 0 && j0 >= 0 && j0 < cn && ((cand >> j0) & 1ull);
                    const int j = own ? j0 : __builtin_ctzll(cand);
                    cand &= ~(1ull << j);
                    i = c0 + j;
                    ++k;
                    return true;
                }
                if (k >= nroots) return false;
                i = k == 0 ? i0 : (i + 1 == nroots ? 0 : i + 1);
                ++k;
                return true;
            };
            bool going = trace && !shadowed && next();
            if (going) enter(i);
            while (ballot(going)) {
                uint32_t la = 0, lcnt = 0;
                bool done = false;
                if (going) {
                    const float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
                    float4 dl4 = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (SPEC && ordered) {
                        dl4 = delta[curr];
                        __asm__ volatile("" : "+v"(dl4.x), "+v"(dl4.y), "+v"(dl4.z));  // see traverse_closest_node
                    }
                    const uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
                    bool pop = true;
                    if (slab(f0, f1, rc, tmax)) {
                        if (cnt >= coop) {
                            la = a;
                            lcnt = cnt;
                        } else if (cnt) {
                            if (leaf_any(tri, a, cnt, rc, tmax)) {
                                shadowed = true;
                                going = false;
                                pop = false;
                            }
                        } else {
                            // index order (left child first, no delta fetched)
                            // unless a root's tree can reach MAX_LEVELS
                            bool nl = true;
                            if (ordered) {
                                if (!SPEC) dl4 = delta[curr];
                                nl = dot(rc.d, xyz(dl4)) > 0.f;
                            }
                            stk.put(level, nl ? a + 1 : a);
                            level++;
                            curr = nl ? a : a + 1;
                            pop = false;
                        }
                    }
                    if (pop) {
                        if (level == 0 || level >= STACK)
                            done = true;
                        else
                            curr = stk.get(--level);
                    }
                }
                if (coop_any<true>(tri, la, lcnt, rc, tmax, live)) {
                    shadowed = true;
                    going = false;
                    done = false;
                }
                if (done) {
                    if (next())
                        enter(i);
                    else
                        going = false;
                }
            }
        };
        if (CHUNKS) {
            for (int c0 = 0; c0 < nroots; c0 += 64) {
                if (!ballot(trace && !shadowed)) break;
                walk(c0, nroots - c0 < 64 ? nroots - c0 : 64);
            }
        } else {
            walk(0, nroots);
        }
        if (valid) set_shadow_bit(&flags[rr], l, shadowed);
    }
    if (shadow_count) {
        if (traced) atomicAdd(&s_cnt, traced);
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt) atomicAdd(shadow_count, (unsigned long long)s_cnt);
    }
}

// Rewrites roots [first, first + n) of the device table from kernel
// arguments: stream-ordered, and the host image may change right after the
// launch.
constexpr int ROOT_CHUNK = 16;
struct RootChunk {
    InstRoot r[ROOT_CHUNK];
};
__global__ void k_put_roots(InstRoot* __restrict__ table, uint32_t first, uint32_t n, const RootChunk c) {
    const uint32_t words = n * (uint32_t)(sizeof(InstRoot) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(c.r);
    uint32_t* dst = reinterpret_cast<uint32_t*>(table + first);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

// Traversal grid: one thread per item up to 64 groups per CU (then
// grid-stride), rounded up to a multiple of the XCD count for xcd_span.
uint32_t trace_grid(rtx_ctx* c, uint64_t count, uint64_t cap, int TB) {
    uint64_t n = count == RTX_DEVICE_COUNT ? cap : count;
    uint64_t g = (n + TB - 1) / TB;
    uint64_t gmax = (uint64_t)c->cus * 64;
    if (g > gmax) g = gmax;
    const uint64_t X = (uint64_t)c->tune.xcds;
    g = (g + X - 1) / X * X;
    return (uint32_t)(g ? g : X);
}

// ---------------------------------------------------------- wavefront --
// Stage kernels of a wavefront step (one launch per stage over the frames of
// every segment).  Per ray they run the code of the single-frame kernels
// above, so each frame's records are those of its stand-alone frame.

// First kernel of a step: the step's queue counters; per frame, the ray
// count of this bounce (primary: the pixel count; later bounces: the rays
// the last step's generation wrote, min(scan total, first dropped offset),
// added to the frame's secondary count as k_sec_finalize does), an empty hit
// list and, when this bounce generates, the next bounce's scan total and
// drop offset.
// Also the trace launch's chunk table (64-ray chunks per segment, inclusive
// prefix, and the total) for the in-order trace form.  One block of
// RTX_WF_MAX_SEGS threads, thread i for segment i.
__global__ __launch_bounds__(RTX_WF_MAX_SEGS) void k_wf_init(const WfSeg* __restrict__ segs, uint32_t nseg,
                                                             uint32_t* __restrict__ ctr) {
    __shared__ uint32_t s_w[RTX_WF_MAX_SEGS / RTX_WAVE];
    for (uint32_t i = threadIdx.x; i < WF_Q_ZERO; i += blockDim.x) ctr[i] = 0u;
    const uint32_t i = threadIdx.x;
    uint32_t chunks = 0;
    if (i < nseg) {
        const WfSeg& S = segs[i];
        if (S.primary) {
            S.count[0] = (uint32_t)S.w * (uint32_t)S.h;
            S.stats[0] = 0ull;
            S.stats[1] = 0ull;
            S.stats[2] = 0ull;
        } else {
            const uint32_t total = S.count[1], drop = S.count[2];
            const uint32_t n = total < drop ? total : drop;
            S.count[0] = n;
            S.stats[0] += n;
            if (n) S.stats[2] += 1ull;
        }
        S.list_count[0] = 0u;
        if (S.gen) {
            S.o_count[1] = 0u;
            S.o_count[2] = 0xffffffffu;
        }
        chunks = (S.count[0] + RTX_WAVE - 1) / RTX_WAVE;
    }
    // inclusive scan of the chunk counts over the block
    const uint32_t lane = threadIdx.x & (RTX_WAVE - 1), w = threadIdx.x / RTX_WAVE;
    uint32_t v = chunks;
#pragma unroll
    for (int off = 1; off < RTX_WAVE; off <<= 1) {
        const uint32_t u = __shfl_up(v, off, RTX_WAVE);
        if (lane >= (uint32_t)off) v += u;
    }
    if (lane == RTX_WAVE - 1) s_w[w] = v;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < RTX_WF_MAX_SEGS / RTX_WAVE; ++k) {
        before += k < (int)w ? s_w[k] : 0u;
        total += s_w[k];
    }
    if (i < nseg) ctr[WF_TAB + i] = before + v;
    if (i == 0) ctr[WF_TOTAL] = total;
}

// Primary rays of the frames that start this step (one sample per pixel, the
// default sample table {0, 0, 1}), and their framebuffer clear: ray k is
// the only ray of its pixel, so it zeroes that pixel's accumulators.
__global__ __launch_bounds__(256) void k_wf_raygen(const WfSeg* __restrict__ segs, uint32_t nseg) {
    __shared__ uint32_t s_end[RTX_WF_MAX_SEGS];
    const uint32_t total = wf_chunks<WF_PIXELS_NEW>(segs, nseg, 256u, s_end);
    for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
        const uint32_t sg = wf_find(s_end, nseg, c);
        const WfSeg& S = segs[sg];
        const uint32_t k = (c - wf_first(s_end, sg)) * 256u + threadIdx.x;
        if (k >= (uint32_t)S.w * (uint32_t)S.h) continue;
        int x, y;
        quad_pixel((int)k, S.w, S.h, S.quad, x, y);
        const v3 d = primary_dir(S.cam, x, y, S.w, S.h, 0.f, 0.f);
        S.ori4[k] = make_float4(S.cam.pos[0], S.cam.pos[1], S.cam.pos[2], 0.f);
        S.dir4[k] = make_float4(d.x, d.y, d.z, 1e36f);
        const uint32_t px = (uint32_t)(S.w * y + x);
        S.pixel[k] = (int32_t)px;
        S.contrib[k] = 1.f;
        S.accum[3 * (size_t)px] = 0u;
        S.accum[3 * (size_t)px + 1] = 0u;
        S.accum[3 * (size_t)px + 2] = 0u;
    }
}

// Closest hit over every frame's rays: persistent waves take 64-ray chunks
// (ChunkSched).  A chunk that starts a secondary-generation tile also clears
// that tile's look-back word for this step's generation.
// One 64-ray chunk (chunk cl of segment S) of the closest-hit launch; the
// whole wave calls it.  A chunk that starts a secondary-generation tile also
// clears that tile's look-back word for this step's generation.
template <int VAR, class Stack>
__device__ __forceinline__ void wf_trace_chunk(const WfSeg& S, uint32_t cl, Stack& stk, uint32_t coop,
                                               uint32_t lane) {
    constexpr uint32_t PER_TILE = WF_GEN_TILE / RTX_WAVE;
    if (S.gen && cl % PER_TILE == 0u && lane == 0) S.status[cl / PER_TILE] = 0u;
    const uint64_t rr = (uint64_t)cl * RTX_WAVE + lane;
    const bool active = rr < (uint64_t)S.count[0];
    Closest best{0.f, 0.f, 0.f, -1};
    RayC rc{};
    if (active) {
        const float4 o = gptr(S.ori4)[rr], dd = gptr(S.dir4)[rr];
        rc = make_ray(xyz(o), xyz(dd), o.w);
        best.t = dd.w;
    }
    closest_var<VAR>(gptr(S.node), gptr(S.delta), gptr(S.tri), rc, best, stk, coop, active);  // whole wave
    const bool hit = active && best.id >= 0;
    const uint64_t hb = ballot(hit);
    uint32_t wbase = 0;
    if (hb) {
        if (lane == (uint32_t)__builtin_ctzll(hb))
            wbase = atomicAdd(S.list_count, (uint32_t)__builtin_popcountll(hb));
        wbase = __shfl(wbase, __builtin_ctzll(hb), RTX_WAVE);
    }
    if (!active) return;
    if (!hit) {
        write_miss_record(rr, gptr(S.flags), gptr(S.t), gptr(S.id), gptr(S.bary), gptr(S.n4), gptr(S.p4),
                          gptr(S.uv2));
        return;
    }
    gptr(S.list)[wbase + lanes_below(hb)] = (uint32_t)rr;
    write_hit_record(gptr(S.tri), gptr(S.nrm), gptr(S.uv), rc, best, rr, gptr(S.flags), gptr(S.t), gptr(S.id),
                     gptr(S.bary), gptr(S.n4), gptr(S.p4), gptr(S.uv2));
}

// Closest hit over every frame's rays, persistent form: waves take 64-ray
// chunks (ChunkSched).
template <int VAR, int TB, int SR>
__global__ __launch_bounds__(TB) void k_wf_trace(const WfSeg* __restrict__ segs, uint32_t nseg,
                                                 uint32_t* __restrict__ queue, uint32_t coop, uint32_t grain) {
    __shared__ uint32_t s_id[SR * TB];
    uint32_t s_low[stack_private_words<SR>()];
    TraversalStack<TB, SR> stk{s_id, threadIdx.x, s_low};
    WaveChunks<WF_RAYS> wc;
    wc.init(segs, nseg, RTX_WAVE);
    ChunkSched sch;
    sch.init(wc.total, queue, grain);
    const uint32_t lane = threadIdx.x & (RTX_WAVE - 1);
    for (uint32_t c = 0; sch.next(c);) {
        uint32_t c0 = 0;
        const uint32_t sg = __builtin_amdgcn_readfirstlane(wc.find(c, c0));
        wf_trace_chunk<VAR>(segs[sg], c - c0, stk, coop, lane);
    }
}

// The same, one chunk per wave in launch order (chunk = global wave index)
// over a grid sized by the frames' ray capacities: the dispatcher hands out
// chunks in order, so the waves in flight always cover neighbouring rays,
// with no queue atomics; waves past the step's chunk total leave at once.
template <int VAR, int TB, int SR>
__global__ __launch_bounds__(TB) void k_wf_trace_ord(const WfSeg* __restrict__ segs, uint32_t nseg,
                                                     const uint32_t* __restrict__ ctr, uint32_t coop) {
    __shared__ uint32_t s_id[SR * TB];
    const uint32_t total = ctr[WF_TOTAL];
    const uint32_t c = blockIdx.x * (TB / RTX_WAVE) + threadIdx.x / RTX_WAVE;
    if (c >= total) return;  // whole waves (no block barrier follows)
    uint32_t s_low[stack_private_words<SR>()];
    TraversalStack<TB, SR> stk{s_id, threadIdx.x, s_low};
    WaveChunks<WF_RAYS> wc;
    wc.load(ctr + WF_TAB, nseg, total);
    uint32_t c0 = 0;
    const uint32_t sg = __builtin_amdgcn_readfirstlane(wc.find(c, c0));
    wf_trace_chunk<VAR>(segs[sg], c - c0, stk, coop, threadIdx.x & (RTX_WAVE - 1));
}

// One 64-pair chunk (chunk cl of segment S) of the shadow launch: the
// (hit, light) pairs of k_shadow, light-major (a wave traces one light's
// rays from neighbouring hits).  The whole wave calls it.
template <int VAR, class Stack>
__device__ __forceinline__ void wf_shadow_chunk(const WfSeg& S, uint32_t cl, Stack& stk, uint32_t coop,
                                                uint32_t lane, bool ordered) {
    const int nl = S.lights->count;
    const uint64_t nh = S.list_count[0];
    const uint64_t it = (uint64_t)cl * RTX_WAVE + lane;
    const bool valid = it < nh * (uint64_t)nl;
    uint32_t rr = 0;
    int l = 0;
    RayC rc{};
    float tmax = 1e37f;
    bool shadowed = false, trace = false;
    if (valid) {
        rr = gptr(S.list)[it % nh];
        l = (int)(it / nh);
        const rtx_light& L = gptr(S.lights)->l[l];
        const v3 p = xyz(gptr(S.p4)[rr]);
        v3 dir = mk3(0.f, 0.f, 0.f);
        trace = true;
        if (L.type == RTX_LIGHT_DIR) {
            dir = -mk3(L.dir[0], L.dir[1], L.dir[2]);
        } else if (L.type == RTX_LIGHT_SPOT) {
            dir = normalize(mk3(L.pos[0], L.pos[1], L.pos[2]) - p);
            if (dot(-dir, mk3(L.dir[0], L.dir[1], L.dir[2])) < L.angle) {
                shadowed = true;
                trace = false;
            }
        } else if (L.type == RTX_LIGHT_POINT) {
            const v3 toL = mk3(L.pos[0], L.pos[1], L.pos[2]) - p;
            tmax = __builtin_sqrtf(dot(toL, toL));
            dir = normalize(toL);
        } else {
            shadowed = true;
            trace = false;
        }
        if (trace) rc = make_ray(p, dir, 0.01f);
    }
    if (any_var<VAR>(gptr(S.node), gptr(S.delta), gptr(S.tri), rc, tmax, stk, coop, trace, nullptr, 0,
                     ordered))  // whole wave
        shadowed = true;
    if (valid) set_shadow_bit(gptr(S.flags) + rr, l, shadowed);
    const uint64_t tb = ballot(trace);
    if (lane == 0 && tb) atomicAdd(&S.stats[1], (unsigned long long)__builtin_popcountll(tb));
}

// Shadow rays of every frame's pairs, persistent form (ChunkSched).
template <int VAR, int TB, int SR>
__global__ __launch_bounds__(TB) void k_wf_shadow(const WfSeg* __restrict__ segs, uint32_t nseg,
                                                  uint32_t* __restrict__ queue, uint32_t coop, uint32_t grain,
                                                  uint32_t ordered) {
    __shared__ uint32_t s_id[SR * TB];
    uint32_t s_low[stack_private_words<SR>()];
    TraversalStack<TB, SR> stk{s_id, threadIdx.x, s_low};
    WaveChunks<WF_PAIRS> wc;
    wc.init(segs, nseg, RTX_WAVE);
    ChunkSched sch;
    sch.init(wc.total, queue, grain);
    const uint32_t lane = threadIdx.x & (RTX_WAVE - 1);
    for (uint32_t c = 0; sch.next(c);) {
        uint32_t c0 = 0;
        const uint32_t sg = __builtin_amdgcn_readfirstlane(wc.find(c, c0));
        wf_shadow_chunk<VAR>(segs[sg], c - c0, stk, coop, lane, ordered != 0);
    }
}

// Between the trace and the in-order shadow launch: the shadow chunk table
// (64-pair chunks per segment, inclusive prefix, total) from the hit lists
// the trace left.  One block of RTX_WF_MAX_SEGS threads.
__global__ __launch_bounds__(RTX_WF_MAX_SEGS) void k_wf_plan_shadow(const WfSeg* __restrict__ segs, uint32_t nseg,
                                                                    uint32_t* __restrict__ ctr) {
    __shared__ uint32_t s_w[RTX_WF_MAX_SEGS / RTX_WAVE];
    const uint32_t i = threadIdx.x;
    uint32_t chunks = 0;
    if (i < nseg) chunks = (uint32_t)((wf_items<WF_PAIRS>(segs[i]) + RTX_WAVE - 1) / RTX_WAVE);
    const uint32_t lane = threadIdx.x & (RTX_WAVE - 1), w = threadIdx.x / RTX_WAVE;
    uint32_t v = chunks;
#pragma unroll
    for (int off = 1; off < RTX_WAVE; off <<= 1) {
        const uint32_t u = __shfl_up(v, off, RTX_WAVE);
        if (lane >= (uint32_t)off) v += u;
    }
    if (lane == RTX_WAVE - 1) s_w[w] = v;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < RTX_WF_MAX_SEGS / RTX_WAVE; ++k) {
        before += k < (int)w ? s_w[k] : 0u;
        total += s_w[k];
    }
    if (i < nseg) ctr[WF_STAB + i] = before + v;
    if (i == 0) ctr[WF_STOTAL] = total;
}

// Shadow rays, one chunk per wave in launch order (see k_wf_trace_ord).
template <int VAR, int TB, int SR>
__global__ __launch_bounds__(TB) void k_wf_shadow_ord(const WfSeg* __restrict__ segs, uint32_t nseg,
                                                      const uint32_t* __restrict__ ctr, uint32_t coop,
                                                      uint32_t ordered) {
    __shared__ uint32_t s_id[SR * TB];
    const uint32_t total = ctr[WF_STOTAL];
    const uint32_t c = blockIdx.x * (TB / RTX_WAVE) + threadIdx.x / RTX_WAVE;
    if (c >= total) return;  // whole waves (no block barrier follows)
    uint32_t s_low[stack_private_words<SR>()];
    TraversalStack<TB, SR> stk{s_id, threadIdx.x, s_low};
    WaveChunks<WF_PAIRS> wc;
    wc.load(ctr + WF_STAB, nseg, total);
    uint32_t c0 = 0;
    const uint32_t sg = __builtin_amdgcn_readfirstlane(wc.find(c, c0));
    wf_shadow_chunk<VAR>(segs[sg], c - c0, stk, coop, threadIdx.x & (RTX_WAVE - 1), ordered != 0);
}

}  // namespace

// ------------------------------------------------- wavefront launchers --
void rtx_wf_init(const WfSeg* segs, uint32_t nseg, uint32_t* ctr, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_init, dim3(1), dim3(RTX_WF_MAX_SEGS), 0, st, segs, nseg, ctr);
}

void rtx_wf_raygen(const WfSeg* segs, uint32_t nseg, uint64_t pixels, int cus, hipStream_t st) {
    uint64_t g = (pixels + 255) / 256;
    if (g > (uint64_t)cus * 16) g = (uint64_t)cus * 16;
    hipLaunchKernelGGL(k_wf_raygen, dim3((uint32_t)(g ? g : 1)), dim3(256), 0, st, segs, nseg);
}

// Persistent grid of kernel K: `share` percent (at least one) of the blocks
// that fit on every CU at once (the LDS stack and the registers set how
// many; queried once per kernel).  A share below 100 leaves room on every
// CU for the short launches of another stream, which would otherwise wait
// until the persistent waves drain.
template <auto K>
static uint32_t resident_grid(int block, int cus, int share) {
    static std::atomic<int> per_cu{0};
    int n = per_cu.load(std::memory_order_relaxed);
    if (n == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, K, block, 0) != hipSuccess || n <= 0)
            n = 1280 / block;  // 20 waves per CU
        per_cu.store(n, std::memory_order_relaxed);
    }
    const int m = share >= 100 ? n : std::max(1, n * share / 100);
    return (uint32_t)cus * (uint32_t)m;
}

void rtx_wf_trace(const WfSeg* segs, uint32_t nseg, uint32_t* ctr, int var, int tb, int sr, uint32_t coop,
                  uint32_t grain, int sched, uint64_t max_chunks, int cus, int share, hipStream_t st) {
    if (sched == 0) {
        const uint64_t g = (max_chunks + tb / RTX_WAVE - 1) / (tb / RTX_WAVE);
#define RTX_WF_TRACE_ORD(V, B, S)                                                                                \
    hipLaunchKernelGGL((k_wf_trace_ord<V, B, S>), dim3((uint32_t)(g ? g : 1)), dim3(B), 0, st, segs, nseg, ctr, \
                       coop)
        RTX_DISPATCH_STACK(RTX_WF_TRACE_ORD, var, tb, sr);
#undef RTX_WF_TRACE_ORD
        return;
    }
#define RTX_WF_TRACE(V, B, S)                                                                                   \
    hipLaunchKernelGGL((k_wf_trace<V, B, S>), dim3(resident_grid<k_wf_trace<V, B, S>>(B, cus, share)), dim3(B), 0, st, \
                       segs, nseg, ctr + WF_Q_TRACE, coop, grain)
    RTX_DISPATCH_STACK(RTX_WF_TRACE, var, tb, sr);
#undef RTX_WF_TRACE
}

void rtx_wf_shadow(const WfSeg* segs, uint32_t nseg, uint32_t* ctr, int var, int tb, int sr, uint32_t coop,
                   uint32_t grain, int sched, uint64_t max_chunks, int cus, int share, bool ordered,
                   hipStream_t st) {
    if (sched == 0) {
        hipLaunchKernelGGL(k_wf_plan_shadow, dim3(1), dim3(RTX_WF_MAX_SEGS), 0, st, segs, nseg, ctr);
        const uint64_t g = (max_chunks + tb / RTX_WAVE - 1) / (tb / RTX_WAVE);
#define RTX_WF_SHADOW_ORD(V, B, S)                                                                                \
    hipLaunchKernelGGL((k_wf_shadow_ord<V, B, S>), dim3((uint32_t)(g ? g : 1)), dim3(B), 0, st, segs, nseg, ctr, \
                       coop, ordered ? 1u : 0u)
        RTX_DISPATCH_STACK(RTX_WF_SHADOW_ORD, var, tb, sr);
#undef RTX_WF_SHADOW_ORD
        return;
    }
#define RTX_WF_SHADOW(V, B, S)                                                                                    \
    hipLaunchKernelGGL((k_wf_shadow<V, B, S>), dim3(resident_grid<k_wf_shadow<V, B, S>>(B, cus, share)), dim3(B), 0, st, \
                       segs, nseg, ctr + WF_Q_SHADOW, coop, grain, ordered ? 1u : 0u)
    RTX_DISPATCH_STACK(RTX_WF_SHADOW, var, tb, sr);
#undef RTX_WF_SHADOW
}

// ----------------------------------------------------------- bundles -----
int32_t rtx_rays_create(rtx_ctx* ctx, uint64_t capacity, rtx_rays** out) {
    RTX_REQUIRE(ctx && out && capacity, RTX_EINVAL, "rtx_rays_create: null argument or capacity 0");
    RTX_REQUIRE(capacity < (1ull << 30), RTX_EUNSUPPORTED, "rtx_rays_create: capacity >= 2^30");
    rtx_rays* r = new rtx_rays();
    r->ctx = ctx;
    r->cap = capacity;
    int32_t e = 0;
    if ((e = r->ori4.reserve(ctx, capacity * 16)) || (e = r->dir4.reserve(ctx, capacity * 16)) ||
        (e = r->pixel.reserve(ctx, capacity * 4)) || (e = r->contrib.reserve(ctx, capacity * 4)) ||
        (e = r->count.reserve(ctx, 16)) || (e = r->scan_status.reserve(ctx, ((capacity + 255) / 256 + 1) * 4))) {
        rtx_rays_destroy(r);
        return e;
    }
    uint32_t init[4] = {0, 0, 0xffffffffu, 0};
    if (hipMemcpy(r->count.p, init, 16, hipMemcpyHostToDevice) != hipSuccess) {
        rtx_rays_destroy(r);
        rtx_set_error("rtx_rays_create: count init failed");
        return RTX_EDEVICE;
    }
    *out = r;
    return RTX_OK;
}

int32_t rtx_rays_destroy(rtx_rays* r) {
    RTX_REQUIRE(r, RTX_EINVAL, "rtx_rays_destroy: null");
    (void)hipStreamSynchronize(r->ctx->stream);
    r->ori4.release();
    r->dir4.release();
    r->pixel.release();
    r->contrib.release();
    r->count.release();
    r->scan_status.release();
    r->psamp.release();
    delete r;
    return RTX_OK;
}

uint64_t rtx_rays_capacity(rtx_rays* r) { return r ? r->cap : 0; }

int32_t rtx_rays_count(rtx_rays* r, uint64_t* count) {
    RTX_REQUIRE(r && count, RTX_EINVAL, "rtx_rays_count: null argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    uint32_t c = 0;
    RTX_CHECK_HIP(hipMemcpy(&c, r->count.p, 4, hipMemcpyDeviceToHost));
    *count = c;
    return RTX_OK;
}

int32_t rtx_rays_read(rtx_rays* r, float* ori4, float* dir4, int32_t* pixel, float* contrib, uint64_t n) {
    RTX_REQUIRE(r && n <= r->cap, RTX_EINVAL, "rtx_rays_read: bad argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    if (ori4) RTX_CHECK_HIP(hipMemcpy(ori4, r->ori4.p, n * 16, hipMemcpyDeviceToHost));
    if (dir4) RTX_CHECK_HIP(hipMemcpy(dir4, r->dir4.p, n * 16, hipMemcpyDeviceToHost));
    if (pixel) RTX_CHECK_HIP(hipMemcpy(pixel, r->pixel.p, n * 4, hipMemcpyDeviceToHost));
    if (contrib) RTX_CHECK_HIP(hipMemcpy(contrib, r->contrib.p, n * 4, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_rays_write(rtx_rays* r, const float* ori4, const float* dir4, const int32_t* pixel, const float* contrib,
                       uint64_t n) {
    RTX_REQUIRE(r && ori4 && dir4 && n <= r->cap, RTX_EINVAL, "rtx_rays_write: bad argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(r->ori4.p, ori4, n * 16, hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(r->dir4.p, dir4, n * 16, hipMemcpyHostToDevice));
    std::vector<int32_t> px(n, 0);
    std::vector<float> cb(n, 1.f);
    RTX_CHECK_HIP(hipMemcpy(r->pixel.p, pixel ? pixel : px.data(), n * 4, hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(r->contrib.p, contrib ? contrib : cb.data(), n * 4, hipMemcpyHostToDevice));
    uint32_t c = (uint32_t)n;
    RTX_CHECK_HIP(hipMemcpy(r->count.p, &c, 4, hipMemcpyHostToDevice));
    return RTX_OK;
}

int32_t rtx_hits_create(rtx_ctx* ctx, uint64_t capacity, rtx_hits** out) {
    RTX_REQUIRE(ctx && out && capacity, RTX_EINVAL, "rtx_hits_create: null argument or capacity 0");
    rtx_hits* h = new rtx_hits();
    h->ctx = ctx;
    h->cap = capacity;
    int32_t e = 0;
    if ((e = h->flags.reserve(ctx, capacity * 4)) || (e = h->t.reserve(ctx, capacity * 4)) ||
        (e = h->id.reserve(ctx, capacity * 4)) || (e = h->bary.reserve(ctx, capacity * 8)) ||
        (e = h->n4.reserve(ctx, capacity * 16)) || (e = h->p4.reserve(ctx, capacity * 16)) ||
        (e = h->uv2.reserve(ctx, capacity * 8)) || (e = h->list.reserve(ctx, capacity * 4)) ||
        (e = h->list_count.reserve(ctx, 16))) {
        rtx_hits_destroy(h);
        return e;
    }
    *out = h;
    return RTX_OK;
}

int32_t rtx_hits_destroy(rtx_hits* h) {
    RTX_REQUIRE(h, RTX_EINVAL, "rtx_hits_destroy: null");
    (void)hipStreamSynchronize(h->ctx->stream);
    DevBuf* all[] = {&h->flags, &h->t, &h->id, &h->bary, &h->n4, &h->p4, &h->uv2, &h->list, &h->list_count};
    for (DevBuf* b : all) b->release();
    delete h;
    return RTX_OK;
}

int32_t rtx_hits_read(rtx_hits* h, uint32_t* flags, float* t, int32_t* id, float* bary2, float* n4, float* p4,
                      float* uv2, uint64_t n) {
    RTX_REQUIRE(h && n <= h->cap, RTX_EINVAL, "rtx_hits_read: bad argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    if (flags) RTX_CHECK_HIP(hipMemcpy(flags, h->flags.p, n * 4, hipMemcpyDeviceToHost));
    if (t) RTX_CHECK_HIP(hipMemcpy(t, h->t.p, n * 4, hipMemcpyDeviceToHost));
    if (id) RTX_CHECK_HIP(hipMemcpy(id, h->id.p, n * 4, hipMemcpyDeviceToHost));
    if (bary2) RTX_CHECK_HIP(hipMemcpy(bary2, h->bary.p, n * 8, hipMemcpyDeviceToHost));
    if (n4) RTX_CHECK_HIP(hipMemcpy(n4, h->n4.p, n * 16, hipMemcpyDeviceToHost));
    if (p4) RTX_CHECK_HIP(hipMemcpy(p4, h->p4.p, n * 16, hipMemcpyDeviceToHost));
    if (uv2) RTX_CHECK_HIP(hipMemcpy(uv2, h->uv2.p, n * 8, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_hits_write(rtx_hits* h, const uint32_t* flags, const float* t, const int32_t* id, const float* n4,
                       const float* p4, uint64_t n) {
    RTX_REQUIRE(h && n <= h->cap, RTX_EINVAL, "rtx_hits_write: bad argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    if (flags) RTX_CHECK_HIP(hipMemcpy(h->flags.p, flags, n * 4, hipMemcpyHostToDevice));
    if (t) RTX_CHECK_HIP(hipMemcpy(h->t.p, t, n * 4, hipMemcpyHostToDevice));
    if (id) RTX_CHECK_HIP(hipMemcpy(h->id.p, id, n * 4, hipMemcpyHostToDevice));
    if (n4) RTX_CHECK_HIP(hipMemcpy(h->n4.p, n4, n * 16, hipMemcpyHostToDevice));
    if (p4) RTX_CHECK_HIP(hipMemcpy(h->p4.p, p4, n * 16, hipMemcpyHostToDevice));
    h->list_valid = false;
    return RTX_OK;
}

// ------------------------------------------------------------ entry pts --
int32_t rtx_raygen_primary(rtx_rays* rays, const rtx_camera* cam, int32_t width, int32_t height, int32_t quad_size,
                           int32_t spp, const float* psamples, uint64_t offset, uint64_t count, void* stream) {
    RTX_REQUIRE(rays && cam, RTX_EINVAL, "rtx_raygen_primary: null argument");
    RTX_REQUIRE(width > 0 && height > 0 && quad_size > 0 && spp > 0, RTX_EINVAL, "rtx_raygen_primary: bad size");
    RTX_REQUIRE(count <= rays->cap, RTX_EINVAL, "rtx_raygen_primary: count exceeds bundle capacity");
    RTX_REQUIRE(offset + count <= (uint64_t)width * height * spp, RTX_EINVAL,
                "rtx_raygen_primary: samples beyond width*height*spp");
    RTX_REQUIRE((uint64_t)width * height * spp < (1ull << 31), RTX_EUNSUPPORTED, "rtx_raygen_primary: too many samples");
    hipStream_t st = rays->ctx->pick(stream);
    float one[3] = {0.f, 0.f, 1.f};
    RTX_REQUIRE(psamples || spp == 1, RTX_EINVAL, "rtx_raygen_primary: spp > 1 needs pixel samples");
    const float* hs = psamples ? psamples : one;
    const size_t n_s = psamples ? (size_t)spp : 1;
    // upload the table only when it changes, so that a frame that reuses it
    // (and its hipGraph capture) performs no host-to-device copy
    if (rays->psamp_host.size() != 3 * n_s || memcmp(rays->psamp_host.data(), hs, 12 * n_s) != 0) {
        // the old table may still be read by queued work; the new one must
        // have landed before any kernel on any stream reads it
        RTX_CHECK_HIP(hipDeviceSynchronize());
        int32_t e = rays->psamp.reserve(rays->ctx, 12 * n_s);
        if (e) return e;
        rays->psamp_host.assign(hs, hs + 3 * n_s);
        RTX_CHECK_HIP(hipMemcpyAsync(rays->psamp.p, rays->psamp_host.data(), 12 * n_s, hipMemcpyHostToDevice, st));
        RTX_CHECK_HIP(hipStreamSynchronize(st));
    }
    float* dtab = rays->psamp.as<float>();
    ProfScope ps(rays->ctx, st, "raygen_primary");
    uint32_t grid = rtx_div_up(count ? count : 1, 256);
    hipLaunchKernelGGL(k_raygen_primary, dim3(grid), dim3(256), 0, st, rays->ori4.as<float4>(),
                       rays->dir4.as<float4>(), rays->pixel.as<int32_t>(), rays->contrib.as<float>(),
                       rays->count.as<uint32_t>(), *cam, width, height, quad_size, dtab, offset, count);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

// The tuned traversal form.  The sibling-pair form pushes only children
// whose boxes pass, so it cannot follow the reference's stack count, which
// ends a ray at a pop with MAX_LEVELS entries; that count can only reach
// MAX_LEVELS in trees of 32 levels, which therefore use the single-node form.
static int traversal_variant(const rtx_scene* s, const SceneView& sv) {
    const int var = (int)s->ctx->tune.trace_variant;
    // the stackless form needs the parents of the nodes it reads (a render
    // copy of the pipelined frame holds none): the single-node form instead
    if (var == 4 && !sv.parent) return 1;
    const int depth = s->ext ? s->ext->depth : s->depth;
    return (var == 2 && depth >= STACK) ? 1 : var;
}

// Iteration bound of the stackless walk: a tree of N <= 2T - 1 nodes allows
// at most 3 visits per node.
static uint32_t stackless_bound(const rtx_scene* s) { return 6u * s->T + 8u; }

int32_t rtx_trace(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream) {
    RTX_REQUIRE(s && rays && hits, RTX_EINVAL, "rtx_trace: null argument");
    RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_trace: scene has no BVH (call rtx_lbvh_build)");
    RTX_REQUIRE(count == RTX_DEVICE_COUNT || (count <= rays->cap && count <= hits->cap), RTX_EINVAL,
                "rtx_trace: count exceeds bundle capacity");
    RTX_REQUIRE(hits->cap >= rays->cap || count != RTX_DEVICE_COUNT, RTX_EINVAL,
                "rtx_trace: hit bundle smaller than ray bundle");
    hipStream_t st = s->ctx->pick(stream);
    if (count == 0) return RTX_OK;
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(1), 0, st, hits->list_count.as<uint32_t>());  // hit count, queue
    ProfScope ps(s->ctx, st, "trace");
    const SceneView sv = scene_view(s);
    const int var = traversal_variant(s, sv), tb = (int)s->ctx->tune.trace_block;
    if (var == 3) {
#define RTX_TRACE_P(B)                                                                                              \
    hipLaunchKernelGGL((k_trace_p<B>), dim3(persist_grid(s->ctx, count, rays->cap, B)), dim3(B), 0, st,            \
                       sv.node, sv.delta, sv.tri, sv.nrm,                                                          \
                       s->uv.as<float2>(), rays->ori4.as<float4>(), rays->dir4.as<float4>(),                        \
                       rays->count.as<uint32_t>(), count, hits->flags.as<uint32_t>(), hits->t.as<float>(),          \
                       hits->id.as<int32_t>(), hits->bary.as<float2>(), hits->n4.as<float4>(), hits->p4.as<float4>(), \
                       hits->uv2.as<float2>(), hits->list.as<uint32_t>(), hits->list_count.as<uint32_t>(),           \
                       coop_threshold(s->ctx), (uint32_t)s->ctx->tune.refill)
        if (tb == 64)
            RTX_TRACE_P(64);
        else if (tb == 256)
            RTX_TRACE_P(256);
        else
            RTX_TRACE_P(128);
#undef RTX_TRACE_P
        RTX_CHECK_HIP(hipGetLastError());
        hits->list_valid = true;
        hits->list_for = count;
        return RTX_OK;
    }
#define RTX_TRACE_LAUNCH(V, B, S)                                                                                    \
    hipLaunchKernelGGL((k_trace<V, B, S>), dim3(trace_grid(s->ctx, count, rays->cap, B)), dim3(B), 0, st,              \
                       sv.node, sv.delta, sv.tri, sv.nrm,                                                 \
                       s->uv.as<float2>(), rays->ori4.as<float4>(), rays->dir4.as<float4>(),                        \
                       rays->count.as<uint32_t>(), count, hits->flags.as<uint32_t>(), hits->t.as<float>(),          \
                       hits->id.as<int32_t>(), hits->bary.as<float2>(), hits->n4.as<float4>(), hits->p4.as<float4>(), \
                       hits->uv2.as<float2>(), hits->list.as<uint32_t>(), hits->list_count.as<uint32_t>(),           \
                       (uint32_t)s->ctx->tune.xcds, coop_threshold(s->ctx), sv.parent, stackless_bound(s))
    RTX_DISPATCH_STACK_NS(RTX_TRACE_LAUNCH, var, tb, (int)s->ctx->tune.stack_lds);
#undef RTX_TRACE_LAUNCH
    RTX_CHECK_HIP(hipGetLastError());
    hits->list_valid = true;
    hits->list_for = count;
    return RTX_OK;
}

// Makes hits->list describe the first `count` rays of `rays` (when the hit
// records did not come from rtx_trace over exactly those rays).
static void ensure_hit_list(rtx_rays* rays, rtx_hits* hits, uint64_t count, hipStream_t st) {
    if (hits->list_valid && hits->list_for == count) return;
    hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, st, hits->list_count.as<uint32_t>());
    uint64_t n = count == RTX_DEVICE_COUNT ? rays->cap : count;
    uint32_t g = rtx_div_up(n ? n : 1, 256);
    hipLaunchKernelGGL(k_hit_list, dim3(g > 4096 ? 4096 : g), dim3(256), 0, st, hits->flags.as<uint32_t>(),
                       rays->count.as<uint32_t>(), count, hits->list.as<uint32_t>(), hits->list_count.as<uint32_t>());
    hits->list_valid = true;
    hits->list_for = count;
}

int32_t rtx_trace_shadow_multi(rtx_scene* s, rtx_rays* const* rays, rtx_hits* const* hits, const uint64_t* counts,
                               int nseg, uint64_t* dev_shadow_count, hipStream_t st) {
    RTX_REQUIRE(s && rays && hits && counts && nseg > 0 && nseg <= RTX_SEG_MAX, RTX_EINVAL,
                "rtx_trace_shadow: bad segment list");
    RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_trace_shadow: scene has no BVH");
    RTX_REQUIRE(s->lights.p, RTX_ESTATE, "rtx_trace_shadow: scene has no lights");
    ShadowSegs segs{};
    uint64_t max_items = 0;
    for (int k = 0; k < nseg; ++k) {
        RTX_REQUIRE(rays[k] && hits[k], RTX_EINVAL, "rtx_trace_shadow: null bundle");
        RTX_REQUIRE(counts[k] == RTX_DEVICE_COUNT || counts[k] <= hits[k]->cap, RTX_EINVAL,
                    "rtx_trace_shadow: count too large");
        ensure_hit_list(rays[k], hits[k], counts[k], st);
        segs.s[k] = ShadowSeg{hits[k]->list.as<uint32_t>(), hits[k]->list_count.as<uint32_t>(),
                              hits[k]->flags.as<uint32_t>(), hits[k]->p4.as<float4>()};
        max_items += counts[k] == RTX_DEVICE_COUNT ? rays[k]->cap : counts[k];
    }
    segs.n = nseg;
    const int nl = s->host_lights.count;
    if (nl <= 0 || max_items == 0) return RTX_OK;
    ProfScope ps(s->ctx, st, "trace_shadow");
    const SceneView sv = scene_view(s);
    uint64_t n = max_items * (uint64_t)nl;
    const int var = traversal_variant(s, sv), tb = (int)s->ctx->tune.trace_block;
    if (var == 3) {
        // queue counter: word 2 of the first segment's hit-list counters
        uint32_t* queue = hits[0]->list_count.as<uint32_t>() + 2;
        hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, st, queue);
#define RTX_SHADOW_P(B)                                                                                          \
    hipLaunchKernelGGL((k_shadow_p<B>), dim3(persist_grid(s->ctx, n, n, B)), dim3(B), 0, st, sv.node,           \
                       sv.delta, sv.tri, s->lights.as<rtx_lights>(), segs,                                     \
                       (unsigned long long*)dev_shadow_count, queue, coop_threshold(s->ctx),                   \
                       (uint32_t)s->ctx->tune.refill)
        if (tb == 64)
            RTX_SHADOW_P(64);
        else if (tb == 256)
            RTX_SHADOW_P(256);
        else
            RTX_SHADOW_P(128);
#undef RTX_SHADOW_P
        RTX_CHECK_HIP(hipGetLastError());
        return RTX_OK;
    }
#define RTX_SHADOW_LAUNCH(V, B, S)                                                                              \
    hipLaunchKernelGGL((k_shadow<V, B, S>), dim3(trace_grid(s->ctx, n, n, B)), dim3(B), 0, st, sv.node,    \
                       sv.delta, sv.tri, s->lights.as<rtx_lights>(), segs,                                     \
                       (unsigned long long*)dev_shadow_count, (uint32_t)s->ctx->tune.xcds, \
                       coop_threshold(s->ctx), sv.parent, stackless_bound(s), ordered)
    // the reference's child order only where a ray's stack can reach MAX_LEVELS
    // (the stackless walk has none)
    const int depth = s->ext ? s->ext->depth : s->depth;
    const uint32_t ordered = (s->ctx->tune.shadow_order != 0 || (depth >= STACK && var != 4)) ? 1u : 0u;
    RTX_DISPATCH_STACK_NS(RTX_SHADOW_LAUNCH, var, tb, (int)s->ctx->tune.stack_lds);
#undef RTX_SHADOW_LAUNCH
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_trace_shadow_impl(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, uint64_t* dev_shadow_count,
                              hipStream_t st) {
    RTX_REQUIRE(s && rays && hits, RTX_EINVAL, "rtx_trace_shadow: null argument");
    if (count == 0) return RTX_OK;
    return rtx_trace_shadow_multi(s, &rays, &hits, &count, 1, dev_shadow_count, st);
}

int32_t rtx_trace_shadow(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream) {
    RTX_REQUIRE(s, RTX_EINVAL, "rtx_trace_shadow: null scene");
    return rtx_trace_shadow_impl(s, rays, hits, count, nullptr, s->ctx->pick(stream));
}

// ---------------------------------------------------- accelerator copies --
int32_t rtx_bvh_create(rtx_ctx* ctx, rtx_bvh** out) {
    RTX_REQUIRE(ctx && out, RTX_EINVAL, "rtx_bvh_create: null argument");
    rtx_bvh* b = new rtx_bvh();
    b->ctx = ctx;
    *out = b;
    return RTX_OK;
}

int32_t rtx_bvh_destroy(rtx_bvh* b) {
    if (!b) return RTX_OK;
    rtx_scene::RenderCopy& c = b->copy;
    for (DevBuf* d : {&c.node, &c.delta, &c.tri, &c.nrm, &c.matmap}) d->release();
    delete b;
    return RTX_OK;
}

int32_t rtx_bvh_copy(rtx_bvh* b, rtx_scene* s, void* stream) {
    RTX_REQUIRE(b && s, RTX_EINVAL, "rtx_bvh_copy: null argument");
    RTX_REQUIRE(b->ctx == s->ctx, RTX_EINVAL, "rtx_bvh_copy: the copy and the scene belong to different contexts");
    RTX_REQUIRE(s->built && !s->sah, RTX_ESTATE, "rtx_bvh_copy: scene has no LBVH (call rtx_lbvh_build)");
    RTX_REQUIRE(s->use_set < 0 && s->use_copy < 0 && !s->ext, RTX_ESTATE,
                "rtx_bvh_copy: the scene is being rendered from another buffer set");
    int32_t e = rtx_copy_render_state(s, b->copy, s->ctx->pick(stream), false);
    if (e) {
        b->valid = false;
        return e;
    }
    b->of = s;
    b->V = s->V;
    b->T = s->T;
    b->depth = s->depth;
    b->valid = true;
    return RTX_OK;
}

// Tracer::trace / shadow_trace against an explicit accelerator
// (src/rt/tracer.cpp:90-135): the same launches with the copy's buffers.
static int32_t check_copy(rtx_scene* s, rtx_bvh* b, const char* who) {
    RTX_REQUIRE(s && b, RTX_EINVAL, std::string(who) + ": null argument");
    RTX_REQUIRE(b->valid && b->of == s && b->T == s->T && b->V == s->V, RTX_EINVAL,
                std::string(who) + ": the accelerator copy was not taken from this scene");
    RTX_REQUIRE(!s->ext && s->use_set < 0 && s->use_copy < 0, RTX_ESTATE,
                std::string(who) + ": the scene is being rendered from another buffer set");
    return RTX_OK;
}

int32_t rtx_trace_with(rtx_scene* s, rtx_bvh* b, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream) {
    int32_t e = check_copy(s, b, "rtx_trace_with");
    if (e) return e;
    s->ext = b;
    e = rtx_trace(s, rays, hits, count, stream);
    s->ext = nullptr;
    return e;
}

int32_t rtx_trace_shadow_with(rtx_scene* s, rtx_bvh* b, rtx_rays* rays, rtx_hits* hits, uint64_t count,
                              void* stream) {
    int32_t e = check_copy(s, b, "rtx_trace_shadow_with");
    if (e) return e;
    s->ext = b;
    e = rtx_trace_shadow(s, rays, hits, count, stream);
    s->ext = nullptr;
    return e;
}

// ---------------------------------------------------------- instances --
static_assert(sizeof(InstRoot) == 176, "InstRoot layout");
static_assert(sizeof(RootChunk) + 16 <= 4096, "root chunk exceeds the kernel argument space");

int32_t rtx_instances_create(rtx_ctx* ctx, rtx_scene* const* scenes, const float* tr, const float* tr_inv, int32_t n,
                             rtx_instances** out) {
    RTX_REQUIRE(ctx && scenes && tr && tr_inv && out, RTX_EINVAL, "rtx_instances_create: null argument");
    RTX_REQUIRE(n > 0 && n <= RTX_MAX_INSTANCES, RTX_EINVAL, "rtx_instances_create: instance count out of range");
    for (int i = 0; i < n; ++i)
        RTX_REQUIRE(scenes[i] && scenes[i]->ctx == ctx, RTX_EINVAL,
                    "rtx_instances_create: null scene or scene of another context");
    rtx_instances* in = new rtx_instances();
    in->ctx = ctx;
    in->n = n;
    in->scene.assign(scenes, scenes + n);
    in->tr.assign(tr, tr + 16 * (size_t)n);
    in->tr_inv.assign(tr_inv, tr_inv + 16 * (size_t)n);
    int32_t e = in->roots.reserve(ctx, sizeof(InstRoot) * (size_t)n);
    if (!e) e = in->wbox.reserve(ctx, sizeof(RootBox) * (size_t)n);
    if (e) {
        in->roots.release();
        delete in;
        return e;
    }
    *out = in;
    return RTX_OK;
}

int32_t rtx_instances_destroy(rtx_instances* in) {
    RTX_REQUIRE(in, RTX_EINVAL, "rtx_instances_destroy: null");
    (void)hipDeviceSynchronize();
    in->roots.release();
    in->hit_root.release();
    in->wbox.release();
    if (in->shading) rtx_scene_destroy(in->shading);
    delete in;
    return RTX_OK;
}

static int32_t instance_bases(const rtx_instances* in, std::vector<uint32_t>& base);

int32_t rtx_instances_set_materials(rtx_instances* in, const rtx_material* per_instance) {
    RTX_REQUIRE(in && per_instance, RTX_EINVAL, "rtx_instances_set_materials: null argument");
    in->inst_mats.assign(per_instance, per_instance + in->n);
    in->shading_bases.clear();  // rebuilt by the next frame
    return RTX_OK;
}

// The shading scene of frames over the instances (Scene::create_bvhs,
// src/rt/scene.cpp:652-670): the distinct scenes in order of first use each
// add the material of their first instance, and their triangles take its id
// in the shared numbering (later instances of a scene share it).
int32_t rtx_instances_shading_scene(rtx_instances* in, rtx_scene** out) {
    RTX_REQUIRE(in && out, RTX_EINVAL, "rtx_instances_shading_scene: null argument");
    RTX_REQUIRE((int)in->inst_mats.size() == in->n, RTX_ESTATE,
                "rtx_instances_shading_scene: instances have no materials (rtx_instances_set_materials)");
    std::vector<uint32_t> base;
    int32_t e = instance_bases(in, base);
    if (e) return e;
    if (!in->shading && (e = rtx_scene_create(in->ctx, &in->shading))) return e;
    rtx_scene* sh = in->shading;
    if (base != in->shading_bases) {
        std::vector<rtx_material> mats;
        std::vector<int32_t> mm;
        std::vector<const rtx_scene*> seen;
        for (int i = 0; i < in->n; ++i) {
            if (std::find(seen.begin(), seen.end(), in->scene[i]) != seen.end()) continue;
            seen.push_back(in->scene[i]);
            mats.push_back(in->inst_mats[i]);
            mm.insert(mm.end(), in->scene[i]->T, (int32_t)mats.size() - 1);
        }
        RTX_CHECK_HIP(hipDeviceSynchronize());
        if ((e = rtx_scene_set_materials(sh, mats.data(), (uint32_t)mats.size()))) return e;
        if ((e = sh->matmap.reserve(in->ctx, mm.size() * 4))) return e;
        RTX_CHECK_HIP(hipMemcpy(sh->matmap.p, mm.data(), mm.size() * 4, hipMemcpyHostToDevice));
        sh->T = (uint32_t)mm.size();
        in->shading_bases = base;
    }
    if ((e = rtx_scene_set_lights(sh, &in->lights))) return e;
    *out = sh;
    return RTX_OK;
}

int32_t rtx_instances_set_transforms(rtx_instances* in, const float* tr, const float* tr_inv) {
    RTX_REQUIRE(in && tr && tr_inv, RTX_EINVAL, "rtx_instances_set_transforms: null argument");
    std::copy(tr, tr + 16 * (size_t)in->n, in->tr.begin());
    std::copy(tr_inv, tr_inv + 16 * (size_t)in->n, in->tr_inv.begin());
    return RTX_OK;
}

int32_t rtx_instances_set_lights(rtx_instances* in, const rtx_lights* lights) {
    RTX_REQUIRE(in && lights, RTX_EINVAL, "rtx_instances_set_lights: null argument");
    RTX_REQUIRE(lights->count >= 0 && lights->count <= RTX_MAX_LIGHTS, RTX_EINVAL,
                "rtx_instances_set_lights: count out of range");
    in->lights = *lights;
    return RTX_OK;
}

// Shared triangle numbering: the distinct scenes in order of first use, each
// taking the next block of ids (the offsets Scene::create_bvhs gives the
// meshes' index ranges, src/rt/scene.cpp:639-700).
static int32_t instance_bases(const rtx_instances* in, std::vector<uint32_t>& base) {
    base.assign(in->n, 0);
    std::vector<const rtx_scene*> seen;
    std::vector<uint32_t> first;
    uint64_t next = 0;
    for (int i = 0; i < in->n; ++i) {
        size_t k = 0;
        while (k < seen.size() && seen[k] != in->scene[i]) ++k;
        if (k == seen.size()) {
            seen.push_back(in->scene[i]);
            first.push_back((uint32_t)next);
            next += in->scene[i]->T;
        }
        base[i] = first[k];
    }
    RTX_REQUIRE(next < (1ull << 31), RTX_EUNSUPPORTED, "rtx_instances: more than 2^31 triangles in all");
    return RTX_OK;
}

int32_t rtx_instances_tri_base(rtx_instances* in, uint32_t* base) {
    RTX_REQUIRE(in && base, RTX_EINVAL, "rtx_instances_tri_base: null argument");
    std::vector<uint32_t> b;
    int32_t e = instance_bases(in, b);
    if (e) return e;
    std::copy(b.begin(), b.end(), base);
    return RTX_OK;
}

// Brings the device root table up to date for work queued on `st`: the
// scenes' current buffers (a rebuild or a new buffer set may move them) and
// the transforms.  Only chunks that differ from what the device holds are
// rewritten.
static int32_t refresh_roots(rtx_instances* in, hipStream_t st, int* var) {
    std::vector<uint32_t> base;
    int32_t e = instance_bases(in, base);
    if (e) return e;
    in->table.resize(in->n);
    bool deep = false;
    for (int i = 0; i < in->n; ++i) {
        rtx_scene* s = in->scene[i];
        RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_trace_instances: an instance's scene has no BVH");
        const SceneView sv = scene_view(s);
        InstRoot& R = in->table[i];
        R.node = sv.node;
        R.delta = sv.delta;
        R.tri = sv.tri;
        R.nrm = sv.nrm;
        R.uv = s->uv.as<float2>();
        R.base = base[i];
        R.ntri = s->T;
        std::copy(&in->tr[16 * (size_t)i], &in->tr[16 * (size_t)i] + 16, R.tr);
        std::copy(&in->tr_inv[16 * (size_t)i], &in->tr_inv[16 * (size_t)i] + 16, R.tr_inv);
        deep = deep || s->depth >= STACK;
    }
    in->deep = deep;
    // the sibling-pair form cannot follow the MAX_LEVELS rule (32-level
    // trees) and the persistent form has no multi-root kernel: both use the
    // single-node form with its early delta load
    const int v = (int)in->ctx->tune.trace_variant;
    *var = (v == 2 && !deep) ? 2 : (v == 0 ? 0 : 1);
    if (in->shipped.size() != in->table.size()) in->shipped.assign(in->table.size(), InstRoot{});
    const bool fresh = in->shipped.empty() || in->shipped[0].node == nullptr;
    for (int c0 = 0; c0 < in->n; c0 += ROOT_CHUNK) {
        const int cn = std::min(ROOT_CHUNK, in->n - c0);
        const size_t bytes = sizeof(InstRoot) * (size_t)cn;
        if (!fresh && memcmp(&in->shipped[c0], &in->table[c0], bytes) == 0) continue;
        RootChunk ch{};
        memcpy(ch.r, &in->table[c0], bytes);
        hipLaunchKernelGGL(k_put_roots, dim3(1), dim3(256), 0, st, in->roots.as<InstRoot>(), (uint32_t)c0,
                           (uint32_t)cn, ch);
        RTX_CHECK_HIP(hipGetLastError());
        memcpy(&in->shipped[c0], &in->table[c0], bytes);
    }
    return RTX_OK;
}

// The culling boxes of the lane-roots kernels (inst_lane_roots 2), computed
// on the stream from the current root table and root boxes; nullptr: no
// culling.
static const RootBox* root_boxes(rtx_instances* in, bool lane_roots, hipStream_t st) {
    if (!lane_roots || in->ctx->tune.inst_lane_roots != 2) return nullptr;
    hipLaunchKernelGGL(k_root_boxes, dim3((in->n + 63) / 64), dim3(64), 0, st, in->roots.as<InstRoot>(), in->n,
                       in->wbox.as<RootBox>());
    return in->wbox.as<RootBox>();
}

int32_t rtx_trace_instances(rtx_instances* in, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream) {
    RTX_REQUIRE(in && rays && hits, RTX_EINVAL, "rtx_trace_instances: null argument");
    RTX_REQUIRE(count == RTX_DEVICE_COUNT || (count <= rays->cap && count <= hits->cap), RTX_EINVAL,
                "rtx_trace_instances: count exceeds bundle capacity");
    RTX_REQUIRE(hits->cap >= rays->cap || count != RTX_DEVICE_COUNT, RTX_EINVAL,
                "rtx_trace_instances: hit bundle smaller than ray bundle");
    hipStream_t st = in->ctx->pick(stream);
    if (count == 0) return RTX_OK;
    int var = 1;
    int32_t e = refresh_roots(in, st, &var);
    if (e) return e;
    if (in->hit_root.bytes < hits->cap * 4) {
        // the buffer may still be read by queued shadow work
        RTX_CHECK_HIP(hipDeviceSynchronize());
        if ((e = in->hit_root.reserve(in->ctx, hits->cap * 4))) return e;
    }
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(1), 0, st, hits->list_count.as<uint32_t>());
    ProfScope ps(in->ctx, st, "trace_instances");
    const int tb = (int)in->ctx->tune.trace_block;
    const bool lane_roots = var != 2 && in->ctx->tune.inst_lane_roots != 0;
#define RTX_TRACE_INST(V, B)                                                                                        \
    do {                                                                                                        \
        if (lane_roots) {                                                                                           \
            hipLaunchKernelGGL((chunks ? (V == 1 ? k_trace_inst_lr<true, true, B> : k_trace_inst_lr<false, true, B>)  \
                                      : (V == 1 ? k_trace_inst_lr<true, false, B> : k_trace_inst_lr<false, false, B>)),\
                               dim3(trace_grid(in->ctx, count, rays->cap, B)), dim3(B), 0, st, TRACE_INST_ARGS,     \
                               boxes);                                                                              \
        } else {                                                                                                    \
            hipLaunchKernelGGL((k_trace_inst<V, B>), dim3(trace_grid(in->ctx, count, rays->cap, B)), dim3(B), 0, st,\
                               TRACE_INST_ARGS);                                                                    \
        }                                                                                                           \
    } while (0)
#define TRACE_INST_ARGS                                                                                             \
    in->roots.as<InstRoot>(), in->n, rays->ori4.as<float4>(), rays->dir4.as<float4>(), rays->count.as<uint32_t>(), \
        count, hits->flags.as<uint32_t>(), hits->t.as<float>(), hits->id.as<int32_t>(), hits->bary.as<float2>(),    \
        hits->n4.as<float4>(), hits->p4.as<float4>(), hits->uv2.as<float2>(), hits->list.as<uint32_t>(),            \
        hits->list_count.as<uint32_t>(), in->hit_root.as<uint32_t>(), coop_threshold(in->ctx)
    const RootBox* boxes = root_boxes(in, lane_roots, st);
    const bool chunks = boxes != nullptr && in->n > 64;  // culling over chunks of 64 roots
    RTX_DISPATCH_TRAVERSAL(RTX_TRACE_INST, var, tb);
#undef TRACE_INST_ARGS
#undef RTX_TRACE_INST
    RTX_CHECK_HIP(hipGetLastError());
    hits->list_valid = true;
    hits->list_for = count;
    return RTX_OK;
}

int32_t rtx_trace_shadow_instances(rtx_instances* in, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream) {
    RTX_REQUIRE(in, RTX_EINVAL, "rtx_trace_shadow_instances: null argument");
    return rtx_trace_shadow_instances_impl(in, rays, hits, count, nullptr, in->ctx->pick(stream));
}

int32_t rtx_trace_shadow_instances_impl(rtx_instances* in, rtx_rays* rays, rtx_hits* hits, uint64_t count,
                                        uint64_t* dev_shadow_count, hipStream_t st) {
    RTX_REQUIRE(in && rays && hits, RTX_EINVAL, "rtx_trace_shadow_instances: null argument");
    RTX_REQUIRE(count == RTX_DEVICE_COUNT || count <= hits->cap, RTX_EINVAL,
                "rtx_trace_shadow_instances: count too large");
    if (count == 0 || in->lights.count <= 0) return RTX_OK;
    int var = 1;
    int32_t e = refresh_roots(in, st, &var);
    if (e) return e;
    ensure_hit_list(rays, hits, count, st);
    ProfScope ps(in->ctx, st, "trace_shadow_instances");
    const uint64_t items = (count == RTX_DEVICE_COUNT ? rays->cap : count) * (uint64_t)in->lights.count;
    // root order hint: the winning roots of the last closest-hit call, when
    // the buffer covers every ray of this bundle (stale values only reorder)
    const uint32_t* order = in->hit_root.bytes >= hits->cap * 4 ? in->hit_root.as<uint32_t>() : nullptr;
    const int tb = (int)in->ctx->tune.trace_block;
    const bool lane_roots = var != 2 && in->ctx->tune.inst_lane_roots != 0;
#define RTX_SHADOW_INST(V, B)                                                                                       \
    do {                                                                                                        \
        if (lane_roots) {                                                                                           \
            hipLaunchKernelGGL((chunks ? (V == 1 ? k_shadow_inst_lr<true, true, B> : k_shadow_inst_lr<false, true, B>)\
                                      : (V == 1 ? k_shadow_inst_lr<true, false, B> : k_shadow_inst_lr<false, false, B>)),\
                               dim3(trace_grid(in->ctx, items, items, B)), dim3(B), 0, st, SHADOW_INST_ARGS, boxes,  \
                               ordered);                                                                            \
        } else {                                                                                                    \
            hipLaunchKernelGGL((k_shadow_inst<V, B>), dim3(trace_grid(in->ctx, items, items, B)), dim3(B), 0, st,   \
                               SHADOW_INST_ARGS, ordered);                                                          \
        }                                                                                                           \
    } while (0)
#define SHADOW_INST_ARGS                                                                                            \
    in->roots.as<InstRoot>(), in->n, in->lights, hits->list.as<uint32_t>(), hits->list_count.as<uint32_t>(),        \
        hits->flags.as<uint32_t>(), hits->p4.as<float4>(), order, (unsigned long long*)dev_shadow_count,            \
        coop_threshold(in->ctx)
    const RootBox* boxes = root_boxes(in, lane_roots, st);
    const bool chunks = boxes != nullptr && in->n > 64;  // culling over chunks of 64 roots
    // shadow rays in index order unless a root's tree has 32 levels (an OR
    // over the roots of order-free any-hit walks)
    const uint32_t ordered = (in->ctx->tune.shadow_order != 0 || in->deep) ? 1u : 0u;
    RTX_DISPATCH_TRAVERSAL(RTX_SHADOW_INST, var, tb);
#undef SHADOW_INST_ARGS
#undef RTX_SHADOW_INST
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}
