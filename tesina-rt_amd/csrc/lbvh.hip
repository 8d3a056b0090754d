// LBVH build on gfx950: prim AABB + global AABB, Morton codes, onesweep
// radix sort, permutation of the scene buffers, level-ordered hierarchy
// emission, bottom-up AABB refit.
//
// Reference: BVHBuilder::build_lbvh / refit_lbvh (src/rt/bvh-builder.cpp:201-850)
// with the kernels of src/kernel/bvh-builder.cl and bvh-builder-sort.cl.
//
// Result contract (checked against the oracle): identical Morton codes,
// the permutation of a stable sort by (code, input position), nodes in
// canonical level order (root, then each iteration's children block ordered
// by parent position, left before right — what process_task produces when
// its work-groups run in order), identical node AABBs.
#include <cstring>

#include "devmath.hpp"
#include "lookback.hpp"
#include "rtx_impl.hpp"

namespace {

constexpr int BS = 256;
constexpr int SORT_ITEMS = 16;               // the smallest tile (sizes the look-back words)
constexpr int SORT_TILE = BS * SORT_ITEMS;
constexpr int BOUNDS_GRID = 512;
constexpr uint32_t FAT_LEAF = 16;
constexpr int HIER_TOP = 8;    // levels 0..8 (<= 256 nodes) emitted by one block
constexpr int REFIT_TOP = 10;  // levels 10..0 (<= 1024 nodes) refit by one block

// Counter slots in scene->tickets.
constexpr int TK_BOUNDS = 0;
constexpr int TK_SORT = 1;    // up to 8 passes (4-bit digits)
constexpr int TK_LEVEL = 16;  // up to 64 levels
constexpr int TK_COUNT = 80;

// Layout of scene->levels (uint32): [d] = first node of level d,
// [64 + d] = node count of level d, [128] = total node count.
// [129] = number of fat leaves (more than FAT_LEAF triangles) in scene->fat.
constexpr int LV_OFF = 0, LV_CNT = 64, LV_TOTAL = 128, LV_FAT = 129, LV_WORDS = 192;

// ------------------------------------------------------------- helpers --
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, uint32_t lane) {
    for (int off = 1; off < RTX_WAVE; off <<= 1) {
        uint32_t y = __shfl_up(x, off, RTX_WAVE);
        if (lane >= (uint32_t)off) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* s_w, uint32_t* total) {
    uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(x, lane);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BS / RTX_WAVE; ++i) {
        uint32_t v = s_w[i];
        base += (i < (int)w) ? v : 0u;
        tot += v;
    }
    __syncthreads();
    *total = tot;
    return base + inc - x;
}

// -------------------------------------------------------------- kernels --
__global__ __launch_bounds__(BS) void k_build_init(uint32_t* zero_a, uint32_t na, uint32_t* zero_b, uint32_t nb,
                                                   uint32_t* zero_c, uint32_t nc, uint32_t* levels, uint32_t* bstart,
                                                   uint32_t* bend, uint32_t* parent, uint32_t T, int set_root) {
    uint32_t stride = gridDim.x * BS;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < na; i += stride) zero_a[i] = 0;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < nb; i += stride) zero_b[i] = 0;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < nc; i += stride) zero_c[i] = 0;
    if (set_root && blockIdx.x == 0) {
        // every word is written once, by one thread, with its final value:
        // the root's count and the node total are 1, the rest 0 (a zeroing
        // pass followed by a store of the root words from another wave would
        // race with it)
        for (uint32_t i = threadIdx.x; i < LV_WORDS; i += BS)
            levels[i] = (i == (uint32_t)(LV_CNT + 0) || i == (uint32_t)LV_TOTAL) ? 1u : 0u;
        if (threadIdx.x == 0) {
            bstart[0] = 0;
            bend[0] = T;
            parent[0] = 0;
        }
    }
}

// build_primitive_bbox + max_local_bbox (src/kernel/bvh-builder.cl:100-130,
// 318-391): per-triangle AABB, its centroid (hi+lo)*0.5 for the Morton
// stage, and the union of all AABBs (last block to finish reduces the
// per-block partials).
__global__ __launch_bounds__(BS) void k_prim_bounds(const uint32_t* __restrict__ idx, const float4* __restrict__ pos,
                                                    float4* __restrict__ cen, float* __restrict__ partials,
                                                    uint32_t T) {
    __shared__ float s_r[6][BS / RTX_WAVE];
    v3 lo = mk3(INFINITY, INFINITY, INFINITY), hi = mk3(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < T; i += gridDim.x * BS) {
        uint32_t i0 = idx[3 * (size_t)i], i1 = idx[3 * (size_t)i + 1], i2 = idx[3 * (size_t)i + 2];
        v3 v0 = xyz(pos[i0]), v1 = xyz(pos[i1]), v2 = xyz(pos[i2]);
        v3 bhi = v0, blo = v0;
        bhi = vmax(bhi, v1);
        bhi = vmax(bhi, v2);
        blo = vmin(blo, v1);
        blo = vmin(blo, v2);
        v3 c = (bhi + blo) * 0.5f;
        cen[i] = make_float4(c.x, c.y, c.z, 0.f);
        lo = vmin(lo, blo);
        hi = vmax(hi, bhi);
    }
    float r[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], __shfl_xor(r[k], off, RTX_WAVE));
#pragma unroll
        for (int k = 3; k < 6; ++k) r[k] = fmaxf(r[k], __shfl_xor(r[k], off, RTX_WAVE));
    }
    uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 6; ++k) s_r[k][w] = r[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int ww = 1; ww < BS / RTX_WAVE; ++ww) {
            for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], s_r[k][ww]);
            for (int k = 3; k < 6; ++k) r[k] = fmaxf(r[k], s_r[k][ww]);
        }
        for (int k = 0; k < 6; ++k) partials[6 * blockIdx.x + k] = r[k];
    }
}

// Union of the per-block partial AABBs (one block).
__global__ __launch_bounds__(BS) void k_gbox_reduce(const float* __restrict__ partials, uint32_t nparts,
                                                    float* __restrict__ gbox6) {
    __shared__ float s_r[6][BS / RTX_WAVE];
    float q[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < nparts; b += BS) {
        for (int k = 0; k < 3; ++k) q[k] = fminf(q[k], partials[6 * b + k]);
        for (int k = 3; k < 6; ++k) q[k] = fmaxf(q[k], partials[6 * b + k]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = fminf(q[k], __shfl_xor(q[k], off, RTX_WAVE));
#pragma unroll
        for (int k = 3; k < 6; ++k) q[k] = fmaxf(q[k], __shfl_xor(q[k], off, RTX_WAVE));
    }
    uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 6; ++k) s_r[k][w] = q[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int ww = 1; ww < BS / RTX_WAVE; ++ww) {
            for (int k = 0; k < 3; ++k) q[k] = fminf(q[k], s_r[k][ww]);
            for (int k = 3; k < 6; ++k) q[k] = fmaxf(q[k], s_r[k][ww]);
        }
        for (int k = 0; k < 6; ++k) gbox6[k] = q[k];
    }
}

// morton_encode (src/kernel/bvh-builder.cl:181-252): centroid normalized to
// the global AABB with a reciprocal then a multiply, clamped with
// fmax(0, fmin(1, s)), scaled by 2^ceil(depth/3), truncated; bit i of the
// code takes axis (off+1+i)%3, source bit (i+off)/3.  Also accumulates the
// 8-bit digit histograms of every radix pass.
// Bit k of v (k < 11) to bit 3k.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x7ffu;
    x = (x | (x << 16)) & 0x00000000ff0000ffull;
    x = (x | (x << 8)) & 0x000000f00f00f00full;
    x = (x | (x << 4)) & 0x00000c30c30c30c3ull;
    x = (x | (x << 2)) & 0x0000249249249249ull;
    return x;
}

__global__ __launch_bounds__(BS) void k_morton(const float4* __restrict__ cen, const float* __restrict__ gbox6,
                                               uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                               uint32_t* __restrict__ hist, uint32_t T, int depth, int passes,
                                               int rbits) {
    // digit histograms of every pass: pass p, digit d at p * 2^rbits + d
    // (4 passes of 8 bits or 8 of 4 bits: at most 1024 words)
    __shared__ uint32_t s_h[1024];
    for (int i = threadIdx.x; i < 1024; i += BS) s_h[i] = 0;
    __syncthreads();
    const uint32_t rmask = (1u << rbits) - 1u;
    v3 glo = mk3(gbox6[0], gbox6[1], gbox6[2]);
    v3 ghi = mk3(gbox6[3], gbox6[4], gbox6[5]);
    v3 inv = recip(ghi - glo);
    float scale = (float)(1u << ((depth + 2) / 3));
    const int offset = depth % 3 ? 3 - (depth % 3) : 0;
    const uint64_t dmask = depth >= 64 ? ~0ull : ((1ull << depth) - 1ull);
    const uint32_t lane = threadIdx.x & 63u;
    // all lanes of a wave step together (the histogram is aggregated per wave)
    for (uint32_t base = blockIdx.x * BS; base < T; base += gridDim.x * BS) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < T;
        uint32_t code = 0;
        if (active) {
            v3 c = xyz(cen[i]);
            v3 sl = (c - glo) * inv;
            sl = vmax(mk3(0.f, 0.f, 0.f), vmin(mk3(1.f, 1.f, 1.f), sl));
            sl = sl * scale;
            // bit b of the code is bit (b + offset) / 3 of axis (offset + 1 + b) % 3:
            // interleave y, z, x at bits 3k, 3k + 1, 3k + 2, shift by offset, keep depth bits
            const uint64_t f = spread3((uint32_t)sl.y) | (spread3((uint32_t)sl.z) << 1) |
                               (spread3((uint32_t)sl.x) << 2);
            code = (uint32_t)((f >> offset) & dmask);
            keys[i] = code;
            vals[i] = i;
        }
        const uint64_t act = ballot(active);
        const uint32_t lead = (uint32_t)__builtin_ctzll(act);
        for (int p = 0; p < passes; ++p) {
            const uint32_t d = (code >> (rbits * p)) & rmask;
            const uint32_t d0 = __shfl(d, (int)lead, RTX_WAVE);
            // input in the last build's order is nearly sorted: whole waves share a digit
            if (ballot(active && d == d0) == act) {
                if (lane == lead) atomicAdd(&s_h[(p << rbits) + d0], (uint32_t)__builtin_popcountll(act));
            } else if (active) {
                atomicAdd(&s_h[(p << rbits) + d], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (passes << rbits); i += BS) {
        const uint32_t c = s_h[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// One stable LSD pass over `bits` <= RB bits at `shift` (onesweep: in-tile
// rank by wave64 ballot matching, per-digit decoupled look-back across tiles,
// LDS-staged scatter) over tiles of BS * ITEMS pairs.  Replaces the bitonic
// network of src/kernel/bvh-builder-sort.cl:26-232.  RB = 8: 256 digits a
// pass; RB = 4: 16 digits, twice the passes, longer runs per digit in the
// scatter (the reference's radix prototype, src/kernel/sort.cl:491-628, uses
// 4-bit digits).  The look-back word of tile t, digit d is status[t * 2^RB + d].
template <int RB, int ITEMS>
__global__ __launch_bounds__(BS) void k_sort_pass(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                  uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                  uint32_t n, int shift, int bits, const uint32_t* __restrict__ hist,
                                                  uint32_t* status, uint32_t* ticket) {
    constexpr int TILE = BS * ITEMS;
    constexpr uint32_t DIG = 1u << RB;
    static_assert(DIG <= (uint32_t)BS, "one thread per digit");
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_wcnt[BS / RTX_WAVE][DIG];
    __shared__ uint32_t s_tofs[DIG];
    __shared__ uint32_t s_gofs[DIG];
    __shared__ uint32_t s_w[BS / RTX_WAVE];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    if (tid < DIG) {
#pragma unroll
        for (int i = 0; i < BS / RTX_WAVE; ++i) s_wcnt[i][tid] = 0;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * TILE;
    const uint32_t dmask = (1u << bits) - 1u;

    uint32_t k[ITEMS], v[ITEMS], r[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        uint32_t e = base + (w * ITEMS + j) * RTX_WAVE + lane;
        bool ok = e < n;
        k[j] = ok ? kin[e] : 0xffffffffu;
        v[j] = ok ? vin[e] : 0u;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        uint32_t e = base + (w * ITEMS + j) * RTX_WAVE + lane;
        bool ok = e < n;
        uint32_t d = (k[j] >> shift) & dmask;
        uint64_t peers = ballot(ok);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if (b < bits) {
                bool bit = (d >> b) & 1u;
                uint64_t bal = ballot(bit);
                peers &= bit ? bal : ~bal;
            }
        }
        uint32_t below = lanes_below(peers);
        uint32_t cnt = (uint32_t)__builtin_popcountll(peers);
        uint32_t basec = ok ? s_wcnt[w][d] : 0u;
        r[j] = basec + below;
        if (ok && below == 0) s_wcnt[w][d] = basec + cnt;
    }
    __syncthreads();
    // per-digit tile totals, wave offsets (exclusive, in wave order)
    uint32_t tot = 0;
    if (tid < DIG) {
        uint32_t c[BS / RTX_WAVE];
#pragma unroll
        for (int i = 0; i < BS / RTX_WAVE; ++i) c[i] = s_wcnt[i][tid];
#pragma unroll
        for (int i = 0; i < BS / RTX_WAVE; ++i) {
            s_wcnt[i][tid] = tot;
            tot += c[i];
        }
    }
    uint32_t dummy;
    uint32_t tofs = block_excl_scan(tot, s_w, &dummy);
    if (tid < DIG) s_tofs[tid] = tofs;
    // global start of each digit = exclusive scan of the pass histogram
    uint32_t hv = tid <= dmask ? hist[tid] : 0u;
    uint32_t dbase = block_excl_scan(hv, s_w, &dummy);
    if (tid <= dmask) {
        uint32_t prefix = lookback_thread(status + tid, tile, tot, DIG);
        s_gofs[tid] = dbase + prefix - tofs;  // dest = s_gofs[d] + local position
    }
    __syncthreads();
    // stage the tile in LDS in sorted order
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        uint32_t e = base + (w * ITEMS + j) * RTX_WAVE + lane;
        if (e < n) {
            uint32_t d = (k[j] >> shift) & dmask;
            uint32_t lp = s_tofs[d] + s_wcnt[w][d] + r[j];
            s_keys[lp] = k[j];
            s_vals[lp] = v[j];
        }
    }
    __syncthreads();
    uint32_t nvalid = n - base < (uint32_t)TILE ? n - base : (uint32_t)TILE;
    for (uint32_t i = tid; i < nvalid; i += BS) {
        uint32_t key = s_keys[i];
        uint32_t dst = s_gofs[(key >> shift) & dmask] + i;
        kout[dst] = key;
        vout[dst] = s_vals[i];
    }
}

// Copy of the buffers the build permutes in place (the reference copies each
// to an aux buffer first, src/rt/bvh-builder.cpp:481-527).
__global__ __launch_bounds__(BS) void k_stage_copy(const uint32_t* __restrict__ idx, uint32_t* __restrict__ idx_c,
                                                   const int32_t* __restrict__ mm, int32_t* __restrict__ mm_c,
                                                   const uint32_t* __restrict__ orig, uint32_t* __restrict__ orig_c,
                                                   uint32_t T) {
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < T; i += gridDim.x * BS) {
        idx_c[3 * (size_t)i] = idx[3 * (size_t)i];
        idx_c[3 * (size_t)i + 1] = idx[3 * (size_t)i + 1];
        idx_c[3 * (size_t)i + 2] = idx[3 * (size_t)i + 2];
        mm_c[i] = mm[i];
        orig_c[i] = orig[i];
    }
}

// rearrange_indices / rearrange_map (src/kernel/bvh-builder.cl:49-72) plus
// the sorted triangle records used by traversal and refit, and the
// original-id map (composition of every build's permutation).
__global__ __launch_bounds__(BS) void k_rearrange(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ idx,
                                                  uint32_t* __restrict__ idx_out, const int32_t* __restrict__ mm,
                                                  int32_t* __restrict__ mm_out, const uint32_t* __restrict__ orig,
                                                  uint32_t* __restrict__ orig_out, const float4* __restrict__ pos,
                                                  float4* __restrict__ tri, uint32_t T) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= T) return;
    uint32_t p = perm[i];
    uint32_t a = idx[3 * (size_t)p], b = idx[3 * (size_t)p + 1], c = idx[3 * (size_t)p + 2];
    idx_out[3 * (size_t)i] = a;
    idx_out[3 * (size_t)i + 1] = b;
    idx_out[3 * (size_t)i + 2] = c;
    mm_out[i] = mm[p];
    orig_out[i] = orig[p];
    float4 pa = pos[a], pb = pos[b], pc = pos[c];
    tri[3 * (size_t)i] = make_float4(pa.x, pa.y, pa.z, __uint_as_float(a));
    tri[3 * (size_t)i + 1] = make_float4(pb.x, pb.y, pb.z, __uint_as_float(b));
    tri[3 * (size_t)i + 2] = make_float4(pc.x, pc.y, pc.z, __uint_as_float(c));
}

// Triangle records from the current index buffer, identity order (refit
// path, Q14 of src/rt/bvh-builder.cpp:779-817).
__global__ __launch_bounds__(BS) void k_tri_records(const uint32_t* __restrict__ idx, const float4* __restrict__ pos,
                                                    float4* __restrict__ tri, uint32_t T) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= T) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t a = idx[3 * (size_t)i + k];
        float4 p = pos[a];
        tri[3 * (size_t)i + k] = make_float4(p.x, p.y, p.z, __uint_as_float(a));
    }
}

// One iteration of process_task (src/kernel/bvh-builder.cl:582-671) over the
// nodes of level d, in order.  Children of the k-th splitting node of the
// level go to off_next + 2k (ordered compaction by decoupled look-back),
// which is the order in-order work-groups of the reference produce.
// Binary radix tree over the sorted codes (Karras 2012, "Maximizing
// parallelism in the construction of BVHs, octrees and k-d trees"), keys
// made unique by their position.  For internal node i it stores the other
// end j of its range and the split gamma (left = [min, gamma], right =
// [gamma+1, max]).  The reference splits a range [s, e) with differing end
// codes at the highest differing bit, at the first index with that bit set
// (find_middle, src/kernel/bvh-builder.cl:484-552): the same split as the
// radix-tree node whose range is [s, e-1], and that node's index is s or e-1.
__device__ __forceinline__ int rt_delta(const uint32_t* __restrict__ codes, int T, int i, int j) {
    if (j < 0 || j >= T) return -1;
    const uint32_t x = codes[i] ^ codes[j];
    return x ? __builtin_clz(x) : 32 + __builtin_clz((uint32_t)(i ^ j));
}

__global__ __launch_bounds__(BS) void k_radix_tree(const uint32_t* __restrict__ codes, uint32_t* __restrict__ kj,
                                                   uint32_t* __restrict__ kg, int T) {
    const int i = (int)(blockIdx.x * BS + threadIdx.x);
    if (i >= T - 1) return;
    const int d = rt_delta(codes, T, i, i + 1) - rt_delta(codes, T, i, i - 1) > 0 ? 1 : -1;
    const int dmin = rt_delta(codes, T, i, i - d);
    int lmax = 2;
    while (rt_delta(codes, T, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (rt_delta(codes, T, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = rt_delta(codes, T, i, j);
    int sp = 0;
    for (int t = l;;) {
        t = (t + 1) >> 1;
        if (rt_delta(codes, T, i, i + (sp + t) * d) > dnode) sp += t;
        if (t <= 1) break;
    }
    const int gamma = i + sp * d + (d < 0 ? -1 : 0);
    kj[i] = (uint32_t)j;
    kg[i] = (uint32_t)gamma;
}

// build_leaf_bbox (src/kernel/bvh-builder.cl:262-288) for a leaf of at most
// FAT_LEAF triangles, by one thread: the vertices of four triangles are
// loaded before any is folded in (a short leaf costs a few memory round
// trips, not one per vertex).  Min/max is exact in any order.
__device__ __forceinline__ void leaf_bounds(const float4* __restrict__ tri, float4* __restrict__ node, uint32_t n,
                                            uint32_t a, uint32_t cnt) {
    v3 lo = xyz(tri[3 * (size_t)a]), hi = lo;
    const uint32_t end = a + cnt;
    for (uint32_t t = a; t < end; t += 4) {
        float4 p[12];
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            const uint32_t tq = t + (uint32_t)(q / 3) < end ? t + (uint32_t)(q / 3) : a;
            p[q] = tri[3 * (size_t)tq + (q % 3)];
        }
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            lo = vmin(lo, xyz(p[q]));
            hi = vmax(hi, xyz(p[q]));
        }
    }
    node[2 * (size_t)n] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(a));
    node[2 * (size_t)n + 1] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(cnt));
}

// Leaves of more than FAT_LEAF triangles (long runs of equal codes, up to
// hundreds of triangles) are listed during emission and bounded here, one
// wave per leaf, every lane folding four vertices per round.
__global__ __launch_bounds__(BS) void k_fat_leaves(const float4* __restrict__ tri, float4* __restrict__ node,
                                                   const uint32_t* __restrict__ fat, const uint32_t* __restrict__ levels) {
    const uint32_t nfat = levels[LV_FAT];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wpb = BS / RTX_WAVE;
    for (uint32_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < nfat; i += gridDim.x * wpb) {
        const uint32_t n = fat[i];
        const uint32_t a = __float_as_uint(node[2 * (size_t)n].w), cnt = __float_as_uint(node[2 * (size_t)n + 1].w);
        const size_t v0 = 3 * (size_t)a, v1 = 3 * (size_t)(a + cnt);
        float q[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (size_t v = v0 + lane; v < v1; v += 4 * RTX_WAVE) {
            float4 p[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) p[j] = tri[v + j * RTX_WAVE < v1 ? v + j * RTX_WAVE : v];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                q[0] = fminf(q[0], p[j].x); q[1] = fminf(q[1], p[j].y); q[2] = fminf(q[2], p[j].z);
                q[3] = fmaxf(q[3], p[j].x); q[4] = fmaxf(q[4], p[j].y); q[5] = fmaxf(q[5], p[j].z);
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) q[k] = fminf(q[k], __shfl_xor(q[k], off, RTX_WAVE));
#pragma unroll
            for (int k = 3; k < 6; ++k) q[k] = fmaxf(q[k], __shfl_xor(q[k], off, RTX_WAVE));
        }
        if (lane == 0) {
            node[2 * (size_t)n] = make_float4(q[0], q[1], q[2], __uint_as_float(a));
            node[2 * (size_t)n + 1] = make_float4(q[3], q[4], q[5], __uint_as_float(cnt));
        }
    }
}

struct HierArgs {
    const uint32_t* __restrict__ codes;
    const uint32_t* __restrict__ kj;  // radix tree: other end of node i's range
    const uint32_t* __restrict__ kg;  // radix tree: split of node i
    uint32_t* __restrict__ bstart;
    uint32_t* __restrict__ bend;
    float4* __restrict__ node;
    uint32_t* __restrict__ parent;
    const float4* __restrict__ tri;
    uint32_t* levels;
    uint32_t* __restrict__ fat_list;
    uint32_t T;
    int min_leaf;
};

// One tile (BS consecutive nodes) of level d, by the whole block; the tile's
// exclusive prefix of splitting nodes comes from the look-back over the
// level's tiles (status).  Ends with a block barrier.
__device__ __forceinline__ void hier_tile(const HierArgs& A, uint32_t* status, int d, bool force_leaf, uint32_t off,
                                          uint32_t cnt, uint32_t tile, uint32_t ntiles, uint32_t* s_w,
                                          uint32_t* s_prefix) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t k = tile * BS + tid;
    const bool active = k < cnt;
    const uint32_t n = off + k;
    uint32_t start = 0, end = 0, split = 0;
    if (active) {
        start = A.bstart[n];
        end = A.bend[n];
        // Leaf iff count <= min_leaf + 2 (unsigned test, so a single triangle
        // falls through) or the first and last codes are equal: find_middle
        // advances its level while first and last share the level's bit, so it
        // ends at their highest differing bit, and runs out of bits only when
        // the codes are equal (every node's level bit is above all bits in
        // which its codes differ).
        if (!force_leaf && end - start - 2u > (uint32_t)A.min_leaf && A.codes[start] != A.codes[end - 1]) {
            const uint32_t last = end - 1;
            const uint32_t other = (last <= A.T - 2u) ? last : start;
            split = (A.kj[start] == last ? A.kg[start] : A.kg[other]) + 1u;
        }
    }
    const bool is_split = split != 0;
    const uint64_t bal = ballot(is_split);
    const uint32_t excl = lanes_below(bal);
    if (lane == 0) s_w[w] = (uint32_t)__builtin_popcountll(bal);
    __syncthreads();
    uint32_t wofs = 0, agg = 0;
#pragma unroll
    for (int i = 0; i < BS / RTX_WAVE; ++i) {
        wofs += (i < (int)w) ? s_w[i] : 0u;
        agg += s_w[i];
    }
    if (w == 0) {
        const uint32_t prefix = lookback_wave(status, tile, agg, 1u);
        if (lane == 0) *s_prefix = prefix;
        if (lane == 0 && tile == ntiles - 1) {
            const uint32_t total = prefix + agg;
            A.levels[LV_OFF + d + 1] = off + cnt;
            A.levels[LV_CNT + d + 1] = 2u * total;
            A.levels[LV_TOTAL] = off + cnt + 2u * total;
        }
    }
    __syncthreads();
    // Leaves of more than FAT_LEAF triangles go to the fat list (bounded by
    // k_fat_leaves, one wave each); the node gets its range now.
    const bool fat = active && !is_split && end - start > FAT_LEAF;
    const uint64_t fm = ballot(fat);
    if (fm) {
        uint32_t fbase = 0;
        if (lane == (uint32_t)__builtin_ctzll(fm))
            fbase = atomicAdd(&A.levels[LV_FAT], (uint32_t)__builtin_popcountll(fm));
        fbase = __shfl(fbase, __builtin_ctzll(fm), RTX_WAVE);
        if (fat) {
            A.fat_list[fbase + lanes_below(fm)] = n;
            A.node[2 * (size_t)n] = make_float4(0.f, 0.f, 0.f, __uint_as_float(start));
            A.node[2 * (size_t)n + 1] = make_float4(0.f, 0.f, 0.f, __uint_as_float(end - start));
        }
    }
    if (active && is_split) {
        const uint32_t rank = *s_prefix + wofs + excl;
        const uint32_t c = off + cnt + 2u * rank;
        A.bstart[c] = start;
        A.bend[c] = split;
        A.bstart[c + 1] = split;
        A.bend[c + 1] = end;
        A.parent[c] = n;
        A.parent[c + 1] = n;
        uint32_t* node_u32 = reinterpret_cast<uint32_t*>(A.node);
        node_u32[8 * (size_t)n + 3] = c;
        node_u32[8 * (size_t)n + 7] = 0u;
    } else if (active && !fat) {
        // build_leaf_bbox (src/kernel/bvh-builder.cl:262-288) for this leaf
        leaf_bounds(A.tri, A.node, n, start, end - start);
    }
    __syncthreads();
}

// Levels 0..dtop (at most 2^d <= BS nodes each, one tile) by one block, one
// level after the other: the top of the tree costs one launch, not dtop+1.
__global__ __launch_bounds__(BS) void k_hier_top(HierArgs A, uint32_t* status, uint32_t status_stride, int dtop,
                                                 int depth) {
    __shared__ uint32_t s_prefix;
    __shared__ uint32_t s_w[BS / RTX_WAVE];
    for (int d = 0; d <= dtop; ++d) {
        const uint32_t off = A.levels[LV_OFF + d], cnt = A.levels[LV_CNT + d];
        if (cnt == 0) {
            if (threadIdx.x == 0) {
                A.levels[LV_OFF + d + 1] = off;
                A.levels[LV_CNT + d + 1] = 0;
            }
            __syncthreads();
            continue;
        }
        hier_tile(A, status + (size_t)d * status_stride, d, d == depth, off, cnt, 0u, 1u, s_w, &s_prefix);
    }
}

// One iteration of process_task (src/kernel/bvh-builder.cl:582-671) over the
// nodes of level d, in order.  Children of the k-th splitting node of the
// level go to off_next + 2k (ordered compaction by decoupled look-back),
// which is the order in-order work-groups of the reference produce.  Blocks
// are persistent: each takes tile tickets until the level is done.
__global__ __launch_bounds__(BS) void k_hier_level(HierArgs A, uint32_t* status, uint32_t* ticket, int d,
                                                   int force_leaf) {
    __shared__ uint32_t s_tile, s_prefix;
    __shared__ uint32_t s_w[BS / RTX_WAVE];
    const uint32_t off = A.levels[LV_OFF + d];
    const uint32_t cnt = A.levels[LV_CNT + d];
    if (cnt == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            A.levels[LV_OFF + d + 1] = off;
            A.levels[LV_CNT + d + 1] = 0;
        }
        return;
    }
    const uint32_t ntiles = (cnt + BS - 1) / BS;
    if (blockIdx.x >= ntiles) return;  // only blocks with work take tickets
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles) return;
        hier_tile(A, status, d, force_leaf != 0, off, cnt, tile, ntiles, s_w, &s_prefix);
    }
}

// build_node_bbox (src/kernel/bvh-builder.cl:290-311) for the internal nodes
// of level d: union of the two children (computed by the launch for level
// d+1, or as leaves), plus the near-child delta right.lo - left.lo used by
// traversal.  One launch per level, deepest first, as the reference's host
// loop (src/rt/bvh-builder.cpp:683-695).
__global__ __launch_bounds__(BS) void k_refit_level(float4* __restrict__ node, float4* __restrict__ delta,
                                                    const uint32_t* __restrict__ levels, int d) {
    const uint32_t off = levels[LV_OFF + d], cnt = levels[LV_CNT + d];
    for (uint32_t k = blockIdx.x * BS + threadIdx.x; k < cnt; k += gridDim.x * BS) {
        uint32_t n = off + k;
        float4 f0 = node[2 * (size_t)n], f1 = node[2 * (size_t)n + 1];
        if (__float_as_uint(f1.w) != 0u) continue;  // leaf: bounded at emission
        uint32_t l = __float_as_uint(f0.w);
        float4 L0 = node[2 * (size_t)l], L1 = node[2 * (size_t)l + 1];
        float4 R0 = node[2 * (size_t)l + 2], R1 = node[2 * (size_t)l + 3];
        v3 plo = vmin(xyz(L0), xyz(R0));
        v3 phi = vmax(xyz(L1), xyz(R1));
        node[2 * (size_t)n] = make_float4(plo.x, plo.y, plo.z, f0.w);
        node[2 * (size_t)n + 1] = make_float4(phi.x, phi.y, phi.z, f1.w);
        v3 dl = xyz(R0) - xyz(L0);
        delta[n] = make_float4(dl.x, dl.y, dl.z, 0.f);
    }
}

// Internal-node AABBs of levels dhi..0 (at most 2^d nodes each) by one
// block, deepest first.
__global__ __launch_bounds__(BS) void k_refit_top(float4* __restrict__ node, float4* __restrict__ delta,
                                                  const uint32_t* __restrict__ levels, int dhi) {
    for (int d = dhi; d >= 0; --d) {
        const uint32_t off = levels[LV_OFF + d], cnt = levels[LV_CNT + d];
        for (uint32_t k = threadIdx.x; k < cnt; k += BS) {
            const uint32_t n = off + k;
            const float4 f0 = node[2 * (size_t)n], f1 = node[2 * (size_t)n + 1];
            if (__float_as_uint(f1.w) != 0u) continue;
            const uint32_t l = __float_as_uint(f0.w);
            const float4 L0 = node[2 * (size_t)l], L1 = node[2 * (size_t)l + 1];
            const float4 R0 = node[2 * (size_t)l + 2], R1 = node[2 * (size_t)l + 3];
            const v3 plo = vmin(xyz(L0), xyz(R0));
            const v3 phi = vmax(xyz(L1), xyz(R1));
            node[2 * (size_t)n] = make_float4(plo.x, plo.y, plo.z, f0.w);
            node[2 * (size_t)n + 1] = make_float4(phi.x, phi.y, phi.z, f1.w);
            const v3 dl = xyz(R0) - xyz(L0);
            delta[n] = make_float4(dl.x, dl.y, dl.z, 0.f);
        }
        __syncthreads();
    }
}

// Leaf AABBs of an existing topology (refit path).
__global__ __launch_bounds__(BS) void k_leaf_bounds(const float4* __restrict__ tri, float4* __restrict__ node,
                                                    const uint32_t* __restrict__ levels) {
    const uint32_t N = levels[LV_TOTAL];
    for (uint32_t n = blockIdx.x * BS + threadIdx.x; n < N; n += gridDim.x * BS) {
        float4 f0 = node[2 * (size_t)n], f1 = node[2 * (size_t)n + 1];
        uint32_t cnt = __float_as_uint(f1.w);
        if (cnt && cnt <= FAT_LEAF) leaf_bounds(tri, node, n, __float_as_uint(f0.w), cnt);
    }
}

}  // namespace

// ------------------------------------------------------------------ host --
static int32_t lbvh_alloc(rtx_scene* s, uint32_t T, int depth) {
    rtx_ctx* c = s->ctx;
    size_t ncap = 2 * (size_t)T;
    uint32_t passes = (uint32_t)(depth + 7) / 8;
    uint32_t sort_tiles = rtx_div_up(T, SORT_TILE);
    uint32_t hier_tiles = rtx_div_up(T, BS);
    int32_t e = 0;
    if ((e = s->b().node.reserve(c, ncap * 32 + 64))) return e;
    if ((e = s->b().delta.reserve(c, ncap * 16))) return e;
    if ((e = s->b().parent.reserve(c, ncap * 4))) return e;
    if ((e = s->b().bstart.reserve(c, ncap * 4))) return e;
    if ((e = s->b().bend.reserve(c, ncap * 4))) return e;
    if ((e = s->b().tri.reserve(c, (size_t)T * 48))) return e;
    if ((e = s->b().codes.reserve(c, (size_t)T * 4))) return e;
    if ((e = s->b().perm.reserve(c, (size_t)T * 4))) return e;
    if ((e = s->b().keys_alt.reserve(c, (size_t)T * 4))) return e;
    if ((e = s->b().vals_alt.reserve(c, (size_t)T * 4))) return e;
    if ((e = s->b().centroid.reserve(c, (size_t)T * 16))) return e;
    if ((e = s->b().partials.reserve(c, BOUNDS_GRID * 6 * 4))) return e;
    if ((e = s->b().gbox.reserve(c, 8 * 4))) return e;
    if ((e = s->b().levels.reserve(c, LV_WORDS * 4))) return e;
    if ((e = s->b().tickets.reserve(c, TK_COUNT * 4))) return e;
    if ((e = s->b().sort_hist.reserve(c, 4 * 256 * 4))) return e;
    if ((e = s->b().sort_status.reserve(c, (size_t)passes * sort_tiles * 256 * 4))) return e;
    if ((e = s->b().hier_status.reserve(c, (size_t)(depth + 1) * hier_tiles * 4))) return e;
    if ((e = s->b().fat.reserve(c, ((size_t)T / (FAT_LEAF + 1) + 1) * 4))) return e;
    return 0;
}

// The build's scratch (device memory beyond its inputs and outputs): the
// scene's buffer set holds it for rtx_lbvh_build; the allocation-free build
// carves it from a fixed array instead.
struct BuildScratch {
    uint32_t *bstart, *bend;        // per-node triangle range during emission
    uint32_t *keys_alt, *vals_alt;  // radix ping-pong, then the radix-tree arrays
    float4* centroid;
    float* partials;
    uint32_t *sort_hist, *sort_status, *tickets, *hier_status;
};

static BuildScratch scratch_of(BuildSet& b) {
    return BuildScratch{b.bstart.as<uint32_t>(),   b.bend.as<uint32_t>(),        b.keys_alt.as<uint32_t>(),
                        b.vals_alt.as<uint32_t>(), b.centroid.as<float4>(),      b.partials.as<float>(),
                        b.sort_hist.as<uint32_t>(), b.sort_status.as<uint32_t>(), b.tickets.as<uint32_t>(),
                        b.hier_status.as<uint32_t>()};
}

static void fat_leaf_bounds(rtx_scene* s, hipStream_t st, rtx_ctx* pc) {
    ProfScope ps(pc, st, "lbvh_fat_leaves");
    uint32_t g = rtx_div_up((uint64_t)s->T / (FAT_LEAF + 1) + 1, BS / RTX_WAVE);
    hipLaunchKernelGGL(k_fat_leaves, dim3(g > 1024 ? 1024 : g), dim3(BS), 0, st, s->b().tri.as<float4>(), s->nodes(),
                       s->b().fat.as<uint32_t>(), s->b().levels.as<uint32_t>());
}

// Internal-node AABBs, deepest level first (levels depth-1 .. 0; level
// `depth` holds only leaves).
static void refit_internal(rtx_scene* s, hipStream_t st, int depth, rtx_ctx* pc) {
    const uint32_t gmax = (uint32_t)s->ctx->cus * (uint32_t)s->ctx->tune.build_grid;
    for (int d = depth - 1; d > REFIT_TOP; --d) {
        ProfScope ps(pc, st, "lbvh_refit_level");
        uint64_t maxn = d < 32 ? (1ull << d) : (uint64_t)s->T;
        if (maxn > 2ull * s->T) maxn = 2ull * s->T;
        uint32_t g = rtx_div_up(maxn, BS);
        hipLaunchKernelGGL(k_refit_level, dim3(g > gmax ? gmax : g), dim3(BS), 0, st, s->nodes(),
                           s->b().delta.as<float4>(), s->b().levels.as<uint32_t>(), d);
    }
    const int dhi = depth - 1 < REFIT_TOP ? depth - 1 : REFIT_TOP;
    if (dhi >= 0) {
        ProfScope ps(pc, st, "lbvh_refit_top");
        hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(BS), 0, st, s->nodes(), s->b().delta.as<float4>(),
                           s->b().levels.as<uint32_t>(), dhi);
    }
}

// Depth and leaf size of a build, clamped as the reference's RendererConfig
// does (src/rt/bvh-builder.cpp:194-199); codes are 32-bit here.
static int32_t build_params(rtx_scene* s, const rtx_lbvh_params* p, int* depth_out, int* min_leaf_out) {
    RTX_REQUIRE(s && s->ctx, RTX_EINVAL, "rtx_lbvh_build: null scene");
    RTX_REQUIRE(s->T > 0 && s->V > 0, RTX_ESTATE, "rtx_lbvh_build: scene has no triangles");
    RTX_REQUIRE(s->T < (1u << 29), RTX_EUNSUPPORTED, "rtx_lbvh_build: more than 2^29 triangles");
    int depth = p ? p->depth : 30;
    int min_leaf = p ? p->min_leaf : 1;
    RTX_REQUIRE(depth <= 32, RTX_EUNSUPPORTED, "rtx_lbvh_build: bvh_depth > 32 needs 64-bit Morton keys");
    *depth_out = depth < 1 ? 1 : depth;
    *min_leaf_out = min_leaf < 1 ? 1 : (min_leaf > 256 ? 256 : min_leaf);
    return RTX_OK;
}

// Sort phase: build state reset, primitive and global bounds, Morton codes,
// radix sort, and the in-place permutation of the scene's index, material
// and id buffers with the triangle records gathered from the positions.  The
// permutation is what chains one build to the next.
// `W` is the scratch; `pc` the context that profiles the launches (null:
// none).  Returns true when the sorted codes and permutation were left in
// W.keys_alt / W.vals_alt (an odd pass count).
static bool build_sort_phase(rtx_scene* s, int depth, hipStream_t st, const BuildScratch& W, rtx_ctx* pc) {
    const uint32_t T = s->T;
    const int rb = (int)s->ctx->tune.sort_bits, items = (int)s->ctx->tune.sort_items;
    const int passes = (depth + rb - 1) / rb;
    const uint32_t sort_tiles = rtx_div_up(T, (uint64_t)BS * items);
    const uint32_t hier_tiles = rtx_div_up(T, BS);
    uint32_t* tickets = W.tickets;
    uint32_t* levels = s->b().levels.as<uint32_t>();
    {
        ProfScope ps(pc, st, "lbvh_init");
        uint32_t nb = (uint32_t)(passes * (size_t)sort_tiles << rb);
        uint32_t nc = (uint32_t)((size_t)(depth + 1) * hier_tiles);
        uint32_t grid = rtx_div_up((uint64_t)(nb > nc ? nb : nc), BS);
        grid = grid < 1 ? 1 : (grid > 2048 ? 2048 : grid);
        // zero: tickets + histograms (contiguous? no: separate buffers)
        hipLaunchKernelGGL(k_build_init, dim3(grid), dim3(BS), 0, st, tickets, (uint32_t)TK_COUNT, W.sort_status, nb,
                           W.hier_status, nc, levels, W.bstart, W.bend, s->b().parent.as<uint32_t>(), T, 1);
        hipLaunchKernelGGL(k_build_init, dim3(1), dim3(BS), 0, st, W.sort_hist, 4u * 256u,
                           (uint32_t*)nullptr, 0u, (uint32_t*)nullptr, 0u, levels, (uint32_t*)nullptr,
                           (uint32_t*)nullptr, (uint32_t*)nullptr, T, 0);
    }
    {
        ProfScope ps(pc, st, "lbvh_prim_bounds");
        uint32_t grid = rtx_div_up(T, BS);
        grid = grid > BOUNDS_GRID ? BOUNDS_GRID : grid;
        hipLaunchKernelGGL(k_prim_bounds, dim3(grid), dim3(BS), 0, st, s->idx.as<uint32_t>(), s->pos.as<float4>(),
                           W.centroid, W.partials, T);
        hipLaunchKernelGGL(k_gbox_reduce, dim3(1), dim3(BS), 0, st, W.partials, grid,
                           s->b().gbox.as<float>());
    }
    {
        ProfScope ps(pc, st, "lbvh_morton");
        uint32_t grid = rtx_div_up(T, BS);
        grid = grid > 1024 ? 1024 : grid;
        hipLaunchKernelGGL(k_morton, dim3(grid), dim3(BS), 0, st, W.centroid, s->b().gbox.as<float>(),
                           s->b().codes.as<uint32_t>(), s->b().perm.as<uint32_t>(), W.sort_hist, T, depth, passes,
                           rb);
    }
    uint32_t *kin = s->b().codes.as<uint32_t>(), *vin = s->b().perm.as<uint32_t>();
    uint32_t *kout = W.keys_alt, *vout = W.vals_alt;
    for (int pass = 0; pass < passes; ++pass) {
        ProfScope ps(pc, st, "lbvh_sort_pass");
        const int shift = rb * pass;
        const int bits = depth - shift < rb ? depth - shift : rb;
        uint32_t* const hist = W.sort_hist + ((size_t)pass << rb);
        uint32_t* const status = W.sort_status + ((size_t)pass * sort_tiles << rb);
#define RTX_SORT_PASS(R, I)                                                                                    \
    hipLaunchKernelGGL((k_sort_pass<R, I>), dim3(sort_tiles), dim3(BS), 0, st, kin, vin, kout, vout, T, shift, \
                       bits, hist, status, tickets + TK_SORT + pass)
        if (rb == 4 && items == 32)
            RTX_SORT_PASS(4, 32);
        else if (rb == 4)
            RTX_SORT_PASS(4, 16);
        else if (items == 32)
            RTX_SORT_PASS(8, 32);
        else
            RTX_SORT_PASS(8, 16);
#undef RTX_SORT_PASS
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    const bool in_alt = kin != s->b().codes.as<uint32_t>();  // odd pass count
    {
        // permute in place: copy, then gather back into the scene's buffers
        // (no handle swap, so a captured frame chains builds as eager ones do)
        ProfScope ps(pc, st, "lbvh_rearrange");
        uint32_t g = rtx_div_up(T, BS);
        hipLaunchKernelGGL(k_stage_copy, dim3(g > 4096 ? 4096 : g), dim3(BS), 0, st, s->idx.as<uint32_t>(),
                           s->idx_alt.as<uint32_t>(), s->matmap.as<int32_t>(), s->matmap_alt.as<int32_t>(),
                           s->orig.as<uint32_t>(), s->orig_alt.as<uint32_t>(), T);
        hipLaunchKernelGGL(k_rearrange, dim3(rtx_div_up(T, BS)), dim3(BS), 0, st, vin,
                           s->idx_alt.as<uint32_t>(), s->idx.as<uint32_t>(), s->matmap_alt.as<int32_t>(),
                           s->matmap.as<int32_t>(), s->orig_alt.as<uint32_t>(), s->orig.as<uint32_t>(),
                           s->pos.as<float4>(), s->b().tri.as<float4>(), T);
    }
    return in_alt;
}

// Tree phase: radix tree, level-ordered hierarchy, fat-leaf bounds and the
// bottom-up refit.  It reads only the buffer set's own sorted codes and
// triangle records, so it can run beside the next build's sort phase when
// that one writes another set.
static void build_tree_phase(rtx_scene* s, int depth, int min_leaf, hipStream_t st, const BuildScratch& W,
                             rtx_ctx* pc) {
    const uint32_t T = s->T;
    const uint32_t hier_tiles = rtx_div_up(T, BS);
    uint32_t* tickets = W.tickets;
    uint32_t* levels = s->b().levels.as<uint32_t>();
    {
        ProfScope ps(pc, st, "lbvh_radix_tree");
        // radix-tree arrays reuse the sort's alternate buffers
        hipLaunchKernelGGL(k_radix_tree, dim3(rtx_div_up(T, BS)), dim3(BS), 0, st, s->b().codes.as<uint32_t>(),
                           W.keys_alt, W.vals_alt, (int)T);
    }
    {
        HierArgs A{s->b().codes.as<uint32_t>(), W.keys_alt, W.vals_alt, W.bstart, W.bend, s->nodes(),
                   s->b().parent.as<uint32_t>(), s->b().tri.as<float4>(), levels, s->b().fat.as<uint32_t>(), T,
                   min_leaf};
        const int dtop = depth < HIER_TOP ? depth : HIER_TOP;
        {
            ProfScope ps(pc, st, "lbvh_hier_top");
            hipLaunchKernelGGL(k_hier_top, dim3(1), dim3(BS), 0, st, A, W.hier_status, hier_tiles, dtop, depth);
        }
        const uint32_t gmax = (uint32_t)s->ctx->cus * (uint32_t)s->ctx->tune.build_grid;
        for (int d = dtop + 1; d <= depth; ++d) {
            ProfScope ps(pc, st, "lbvh_hier_level");
            uint64_t maxn = d < 32 ? (1ull << d) : (uint64_t)T;
            if (maxn > 2ull * T) maxn = 2ull * T;
            uint32_t grid = rtx_div_up(maxn, BS);
            hipLaunchKernelGGL(k_hier_level, dim3(grid > gmax ? gmax : grid), dim3(BS), 0, st, A,
                               W.hier_status + (size_t)d * hier_tiles, tickets + TK_LEVEL + d, d, d == depth ? 1 : 0);
        }
    }
    fat_leaf_bounds(s, st, pc);
    refit_internal(s, st, depth, pc);
}

// Sorted codes and permutation left in the scratch's alternates by an odd
// pass count go to the buffer set: by a handle swap when the set owns the
// alternates, by a copy otherwise.
static void settle_sorted(rtx_scene* s, bool in_alt) {
    if (!in_alt) return;
    std::swap(s->b().codes, s->b().keys_alt);
    std::swap(s->b().perm, s->b().vals_alt);
}

int32_t rtx_lbvh_build(rtx_scene* s, const rtx_lbvh_params* p, void* stream) {
    int depth = 30, min_leaf = 1;
    int32_t e = build_params(s, p, &depth, &min_leaf);
    if (e) return e;
    hipStream_t st = s->ctx->pick(stream);
    if ((e = lbvh_alloc(s, s->T, depth))) return e;
    {
        ProfScope whole(s->ctx, st, "LBVH_BUILD");
        settle_sorted(s, build_sort_phase(s, depth, st, scratch_of(s->b()), s->ctx));
        build_tree_phase(s, depth, min_leaf, st, scratch_of(s->b()), s->ctx);
    }
    RTX_CHECK_HIP(hipGetLastError());
    s->built = true;
    s->sah = false;
    s->depth = depth;
    s->min_leaf = min_leaf;
    return RTX_OK;
}

int32_t rtx_lbvh_build_split(rtx_scene* s, const rtx_lbvh_params* p, hipStream_t sort_stream,
                             hipEvent_t sorted, hipStream_t tree_stream) {
    int depth = 30, min_leaf = 1;
    int32_t e = build_params(s, p, &depth, &min_leaf);
    if (e) return e;
    if ((e = lbvh_alloc(s, s->T, depth))) return e;
    settle_sorted(s, build_sort_phase(s, depth, sort_stream, scratch_of(s->b()), s->ctx));
    RTX_CHECK_HIP(hipEventRecord(sorted, sort_stream));
    RTX_CHECK_HIP(hipStreamWaitEvent(tree_stream, sorted, 0));
    build_tree_phase(s, depth, min_leaf, tree_stream, scratch_of(s->b()), s->ctx);
    RTX_CHECK_HIP(hipGetLastError());
    s->built = true;
    s->sah = false;
    s->depth = depth;
    s->min_leaf = min_leaf;
    return RTX_OK;
}

int32_t rtx_lbvh_reserve(rtx_scene* s, int depth) { return lbvh_alloc(s, s->T, depth); }

// ------------------------------------------------ allocation-free build --
// rtx_lbvh_build_noalloc runs the same kernels as rtx_lbvh_build with the
// build scratch carved from `scratch`, a fixed array that every GPU holds for
// the life of the process, so that it calls no allocator.  Sized for
// RTX_LBVH_NOALLOC_MAX_TRIS triangles at any depth <= 32 (4 sort passes of
// 8-bit digits or 8 of 4-bit digits, whose look-back words fit the same space; 33
// levels).
namespace {
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_T = RTX_LBVH_NOALLOC_MAX_TRIS;
constexpr size_t NA_SORT_TILES = (NA_T + SORT_TILE - 1) / SORT_TILE;
constexpr size_t NA_HIER_TILES = (NA_T + BS - 1) / BS;
constexpr size_t NA_BYTES = 2 * al256(2 * NA_T * 4)                  // bstart, bend
                          + 2 * al256(NA_T * 4)                      // keys_alt, vals_alt
                          + al256(NA_T * 16)                         // centroid
                          + al256((size_t)BOUNDS_GRID * 6 * 4)       // partials
                          + al256(4 * 256 * 4)                       // sort_hist
                          + al256(4 * NA_SORT_TILES * 256 * 4)       // sort_status
                          + al256((size_t)TK_COUNT * 4)              // tickets
                          + al256(33 * NA_HIER_TILES * 4);           // hier_status
}  // namespace

static_assert(NA_BYTES == RTX_LBVH_NOALLOC_SCRATCH_BYTES, "include/rtx.h states the scratch size");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// Once per context (one device, current when rtx_init calls this): the
// device address of this device's copy of `scratch`.  Host code uses only
// this address, never the array's name.
int32_t rtx_lbvh_scratch_init(rtx_ctx* c) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        rtx_set_error(std::string("rtx_init: hipGetSymbolAddress(build scratch): ") + hipGetErrorString(e));
        return RTX_EDEVICE;
    }
    c->build_scratch = ptr;
    return RTX_OK;
}

static BuildScratch carve_scratch(void* base, uint32_t T) {
    char* p = static_cast<char*>(base);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += al256(bytes);
        return q;
    };
    const size_t sort_tiles = rtx_div_up(T, SORT_TILE), hier_tiles = rtx_div_up(T, BS);
    BuildScratch W;
    W.bstart = reinterpret_cast<uint32_t*>(take(2 * (size_t)T * 4));
    W.bend = reinterpret_cast<uint32_t*>(take(2 * (size_t)T * 4));
    W.keys_alt = reinterpret_cast<uint32_t*>(take((size_t)T * 4));
    W.vals_alt = reinterpret_cast<uint32_t*>(take((size_t)T * 4));
    W.centroid = reinterpret_cast<float4*>(take((size_t)T * 16));
    W.partials = reinterpret_cast<float*>(take((size_t)BOUNDS_GRID * 6 * 4));
    W.sort_hist = reinterpret_cast<uint32_t*>(take(4 * 256 * 4));
    W.sort_status = reinterpret_cast<uint32_t*>(take(4 * sort_tiles * 256 * 4));
    W.tickets = reinterpret_cast<uint32_t*>(take((size_t)TK_COUNT * 4));
    W.hier_status = reinterpret_cast<uint32_t*>(take(33 * hier_tiles * 4));
    return W;
}

namespace {
__global__ void k_copy_pairs(const uint32_t* __restrict__ ka, const uint32_t* __restrict__ va, uint32_t* __restrict__ kb,
                             uint32_t* __restrict__ vb, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        kb[i] = ka[i];
        vb[i] = va[i];
    }
}
}  // namespace

int32_t rtx_lbvh_build_noalloc(rtx_scene* s, const rtx_lbvh_params* p, void* stream) {
    int depth = 30, min_leaf = 1;
    int32_t e = build_params(s, p, &depth, &min_leaf);
    if (e) return e;
    RTX_REQUIRE(s->T <= RTX_LBVH_NOALLOC_MAX_TRIS, RTX_EINVAL,
                "rtx_lbvh_build_noalloc: more than RTX_LBVH_NOALLOC_MAX_TRIS (1048576) triangles");
    RTX_REQUIRE(s->ctx->build_scratch, RTX_ESTATE, "rtx_lbvh_build_noalloc: context has no build scratch");
    // outputs: the current buffer set, reserved beforehand (rtx_lbvh_build)
    const size_t T = s->T, ncap = 2 * T;
    BuildSet& b = s->b();
    RTX_REQUIRE(b.node.bytes >= ncap * 32 + 64 && b.delta.bytes >= ncap * 16 && b.parent.bytes >= ncap * 4 &&
                    b.tri.bytes >= T * 48 && b.codes.bytes >= T * 4 && b.perm.bytes >= T * 4 && b.gbox.bytes >= 32 &&
                    b.levels.bytes >= LV_WORDS * 4 && b.fat.bytes >= (T / (FAT_LEAF + 1) + 1) * 4,
                RTX_ESTATE,
                "rtx_lbvh_build_noalloc: the scene's tree buffers are not reserved for its triangle count "
                "(run rtx_lbvh_build once first)");
    hipStream_t st = s->ctx->pick(stream);
    const BuildScratch W = carve_scratch(s->ctx->build_scratch, s->T);
    if (build_sort_phase(s, depth, st, W, nullptr)) {
        uint32_t g = rtx_div_up(T, BS);
        hipLaunchKernelGGL(k_copy_pairs, dim3(g > 4096 ? 4096 : g), dim3(BS), 0, st, W.keys_alt, W.vals_alt,
                           b.codes.as<uint32_t>(), b.perm.as<uint32_t>(), (uint32_t)T);
    }
    build_tree_phase(s, depth, min_leaf, st, W, nullptr);
    RTX_CHECK_HIP(hipGetLastError());
    s->built = true;
    s->sah = false;
    s->depth = depth;
    s->min_leaf = min_leaf;
    return RTX_OK;
}

// Scene::create_aggregate_bvh with the SAH accelerator (BVH::construct_and_map,
// src/rt/bvh.cpp:126-175): the tree is built on the host from the scene's
// current triangle order, the index buffer, material map and upload ids are
// permuted into its order, and the nodes go to the current buffer set in the
// layout every traversal form reads (children adjacent as the builder
// allocates them, deltas, parents, triangle records).
int32_t rtx_sah_build(rtx_scene* s, void* stream) {
    RTX_REQUIRE(s && s->ctx, RTX_EINVAL, "rtx_sah_build: null scene");
    RTX_REQUIRE(s->T > 0 && s->V > 0, RTX_ESTATE, "rtx_sah_build: scene has no triangles");
    hipStream_t st = s->ctx->pick(stream);
    const uint32_t T = s->T, V = s->V;
    RTX_CHECK_HIP(hipDeviceSynchronize());
    std::vector<float4> p4(V);
    std::vector<uint32_t> idx(3 * (size_t)T), orig(T);
    std::vector<int32_t> mm(T);
    RTX_CHECK_HIP(hipMemcpy(p4.data(), s->pos.p, V * sizeof(float4), hipMemcpyDeviceToHost));
    RTX_CHECK_HIP(hipMemcpy(idx.data(), s->idx.p, idx.size() * 4, hipMemcpyDeviceToHost));
    RTX_CHECK_HIP(hipMemcpy(mm.data(), s->matmap.p, T * 4, hipMemcpyDeviceToHost));
    RTX_CHECK_HIP(hipMemcpy(orig.data(), s->orig.p, T * 4, hipMemcpyDeviceToHost));
    std::vector<float> p3(3 * (size_t)V);
    for (uint32_t v = 0; v < V; ++v) {
        p3[3 * (size_t)v] = p4[v].x;
        p3[3 * (size_t)v + 1] = p4[v].y;
        p3[3 * (size_t)v + 2] = p4[v].z;
    }
    std::vector<rtx_node_ref> nodes;
    std::vector<uint32_t> order;
    uint32_t height = 0;
    int32_t e = rtx_sah_construct(p3.data(), V, idx.data(), T, nodes, order, &height);
    if (e) return e;
    const uint32_t N = (uint32_t)nodes.size();
    std::vector<uint32_t> idx2(idx.size()), orig2(T);
    std::vector<int32_t> mm2(T);
    for (uint32_t i = 0; i < T; ++i) {
        for (int k = 0; k < 3; ++k) idx2[3 * (size_t)i + k] = idx[3 * (size_t)order[i] + k];
        mm2[i] = mm[order[i]];
        orig2[i] = orig[order[i]];
    }
    std::vector<float4> nd(2 * (size_t)N), dl(N, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<uint32_t> par(N), lv(LV_WORDS, 0u), depth_of(N, 0u);
    for (uint32_t n = 0; n < N; ++n) {
        const rtx_node_ref& r = nodes[n];
        const bool leaf = r.leaf != 0;
        if (!leaf) RTX_REQUIRE(r.b == r.a + 1, RTX_EUNSUPPORTED, "rtx_sah_build: children not adjacent");
        const uint32_t w0 = r.a, w1 = leaf ? r.b - r.a : 0u;
        float f0, f1;
        std::memcpy(&f0, &w0, 4);
        std::memcpy(&f1, &w1, 4);
        nd[2 * (size_t)n] = make_float4(r.lo[0], r.lo[1], r.lo[2], f0);
        nd[2 * (size_t)n + 1] = make_float4(r.hi[0], r.hi[1], r.hi[2], f1);
        if (!leaf) {
            const rtx_node_ref &L = nodes[r.a], &R = nodes[r.a + 1];
            dl[n] = make_float4(R.lo[0] - L.lo[0], R.lo[1] - L.lo[1], R.lo[2] - L.lo[2], 0.f);
            depth_of[r.a] = depth_of[r.a + 1] = depth_of[n] + 1;  // children follow their parent
        }
        par[n] = r.parent;
        if (depth_of[n] < 64) lv[LV_CNT + depth_of[n]]++;
    }
    lv[LV_TOTAL] = N;
    if ((e = lbvh_alloc(s, T, 30))) return e;
    BuildSet& b = s->b();
    RTX_CHECK_HIP(hipMemcpy(s->idx.p, idx2.data(), idx2.size() * 4, hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(s->matmap.p, mm2.data(), T * 4, hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(s->orig.p, orig2.data(), T * 4, hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(s->nodes(), nd.data(), nd.size() * sizeof(float4), hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(b.delta.p, dl.data(), dl.size() * sizeof(float4), hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(b.parent.p, par.data(), N * 4, hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(b.levels.p, lv.data(), LV_WORDS * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_tri_records, dim3(rtx_div_up(T, BS)), dim3(BS), 0, st, s->idx.as<uint32_t>(),
                       s->pos.as<float4>(), b.tri.as<float4>(), T);
    RTX_CHECK_HIP(hipGetLastError());
    RTX_CHECK_HIP(hipStreamSynchronize(st));
    s->built = true;
    s->sah = true;
    // the deepest node's depth: trees of 32 or more levels keep to the
    // traversal forms that follow the MAX_LEVELS rule
    s->depth = (int)std::min<uint32_t>(height, 63);
    return RTX_OK;
}

int32_t rtx_lbvh_refit(rtx_scene* s, void* stream) {
    RTX_REQUIRE(s && s->ctx, RTX_EINVAL, "rtx_lbvh_refit: null scene");
    RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_lbvh_refit: no BVH built");
    RTX_REQUIRE(!s->sah, RTX_ESTATE, "rtx_lbvh_refit: the tree is an SAH tree (refit follows LBVH levels)");
    hipStream_t st = s->ctx->pick(stream);
    const uint32_t T = s->T;
    {
        ProfScope ps(s->ctx, st, "lbvh_tri_records");
        hipLaunchKernelGGL(k_tri_records, dim3(rtx_div_up(T, BS)), dim3(BS), 0, st, s->idx.as<uint32_t>(),
                           s->pos.as<float4>(), s->b().tri.as<float4>(), T);
    }
    {
        ProfScope ps(s->ctx, st, "lbvh_leaf_bounds");
        uint32_t ncap = (uint32_t)(2 * (size_t)T);
        uint32_t g = rtx_div_up(ncap, BS);
        hipLaunchKernelGGL(k_leaf_bounds, dim3(g > 4096 ? 4096 : g), dim3(BS), 0, st, s->b().tri.as<float4>(),
                           s->nodes(), s->b().levels.as<uint32_t>());
    }
    fat_leaf_bounds(s, st, s->ctx);
    refit_internal(s, st, s->depth, s->ctx);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_lbvh_info(rtx_scene* s, uint32_t* node_count, uint32_t* level_counts64) {
    RTX_REQUIRE(s && s->ctx, RTX_EINVAL, "rtx_lbvh_info: null scene");
    RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_lbvh_info: no BVH built");
    uint32_t lv[LV_WORDS];
    RTX_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(lv, s->b().levels.p, sizeof lv, hipMemcpyDeviceToHost));
    if (node_count) *node_count = lv[LV_TOTAL];
    if (level_counts64)
        for (int i = 0; i < 64; ++i) level_counts64[i] = (i < 63 && i < s->depth) ? lv[LV_CNT + i + 1] : 0u;
    return RTX_OK;
}

int32_t rtx_read_nodes_ref(rtx_scene* s, rtx_node_ref* out, uint32_t cap, uint32_t* count) {
    uint32_t N = 0;
    int32_t e = rtx_lbvh_info(s, &N, nullptr);
    if (e) return e;
    if (count) *count = N;
    if (!out) return RTX_OK;
    RTX_REQUIRE(cap >= N, RTX_EINVAL, "rtx_read_nodes_ref: capacity below node count");
    std::vector<float4> nd(2 * (size_t)N);
    std::vector<uint32_t> par(N);
    RTX_CHECK_HIP(hipMemcpy(nd.data(), s->nodes(), nd.size() * sizeof(float4), hipMemcpyDeviceToHost));
    RTX_CHECK_HIP(hipMemcpy(par.data(), s->b().parent.p, N * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t n = 0; n < N; ++n) {
        const float4 a = nd[2 * (size_t)n], b = nd[2 * (size_t)n + 1];
        rtx_node_ref& o = out[n];
        o.lo[0] = a.x; o.lo[1] = a.y; o.lo[2] = a.z; o.lo[3] = 0.f;
        o.hi[0] = b.x; o.hi[1] = b.y; o.hi[2] = b.z; o.hi[3] = 0.f;
        uint32_t w0, w1;
        std::memcpy(&w0, &a.w, 4);
        std::memcpy(&w1, &b.w, 4);
        if (w1 == 0) {  // internal
            o.a = w0;
            o.b = w0 + 1;
            o.leaf = 0;
        } else {
            o.a = w0;
            o.b = w0 + w1;
            o.leaf = 1;
        }
        o.parent = n == 0 ? 0u : par[n];
        o.split_axis = 0;
        o.pad[0] = o.pad[1] = 0;
    }
    return RTX_OK;
}
