// Device helpers of the wavefront stage kernels: which items a segment
// holds for a stage, the per-block table of chunk ends over the segments,
// and the segment of a chunk.
//
// A stage's items are the concatenation of every segment's items, cut into
// chunks that never straddle two segments (a segment's last chunk may be
// partial), so the wave or block that takes a chunk works for one frame and
// one scene.  Segments are listed grouped by scene, so consecutive chunks
// read the same BVH.
#pragma once

#include "devmath.hpp"
#include "rtx_impl.hpp"

enum WfWhat { WF_RAYS = 0, WF_PAIRS, WF_GEN_TILES, WF_PIXELS_NEW, WF_PIXELS_DONE };

// Items of segment S for a stage; the counts are those the previous kernels
// of the step left on the device.
template <int WHAT>
__device__ __forceinline__ uint64_t wf_items(const WfSeg& S) {
    if (WHAT == WF_RAYS) return S.count[0];
    if (WHAT == WF_PAIRS) return (uint64_t)S.list_count[0] * (uint64_t)S.lights->count;
    if (WHAT == WF_GEN_TILES) return S.gen ? ((uint64_t)S.count[0] + 255u) / 256u : 0u;
    if (WHAT == WF_PIXELS_NEW) return S.primary ? (uint64_t)S.w * (uint64_t)S.h : 0u;
    return S.resolve ? (uint64_t)S.w * (uint64_t)S.h : 0u;
}

// Wave 0 fills s_end[0, nseg) with the inclusive prefix of ceil(items / ch)
// over the segments; every thread of the block returns the chunk total.
// Contains a block barrier: call it from every thread, before any other use
// of s_end.
template <int WHAT>
__device__ uint32_t wf_chunks(const WfSeg* __restrict__ segs, uint32_t nseg, uint32_t ch, uint32_t* s_end) {
    if (threadIdx.x < RTX_WAVE) {
        const uint32_t lane = threadIdx.x;
        uint32_t carry = 0;
        for (uint32_t base = 0; base < nseg; base += RTX_WAVE) {
            const uint32_t i = base + lane;
            uint32_t v = i < nseg ? (uint32_t)((wf_items<WHAT>(segs[i]) + ch - 1) / ch) : 0u;
#pragma unroll
            for (int off = 1; off < RTX_WAVE; off <<= 1) {
                const uint32_t u = __shfl_up(v, off, RTX_WAVE);
                if (lane >= (uint32_t)off) v += u;
            }
            if (i < nseg) s_end[i] = carry + v;
            carry += __shfl(v, RTX_WAVE - 1, RTX_WAVE);
        }
    }
    __syncthreads();
    return nseg ? s_end[nseg - 1] : 0u;
}

// The segment holding chunk c < total: the first i with s_end[i] > c.
__device__ __forceinline__ uint32_t wf_find(const uint32_t* s_end, uint32_t nseg, uint32_t c) {
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_end[mid] > c)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// First chunk of segment i.
__device__ __forceinline__ uint32_t wf_first(const uint32_t* s_end, uint32_t i) { return i ? s_end[i - 1] : 0u; }
