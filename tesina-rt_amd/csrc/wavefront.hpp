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

// Input rays per tile of the secondary generation (256 threads x 4 rays).
constexpr uint32_t WF_GEN_TILE = 1024;

// Items of segment S for a stage; the counts are those the previous kernels
// of the step left on the device.
template <int WHAT>
__device__ __forceinline__ uint64_t wf_items(const WfSeg& S) {
    if (WHAT == WF_RAYS) return S.count[0];
    if (WHAT == WF_PAIRS) return (uint64_t)S.list_count[0] * (uint64_t)S.lights->count;
    if (WHAT == WF_GEN_TILES) return S.gen ? ((uint64_t)S.count[0] + WF_GEN_TILE - 1) / WF_GEN_TILE : 0u;
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

// The same table held by one wave in registers: lane i of end[r] is the end
// of segment 64 r + i (lanes past the last segment hold the total), so the
// persistent traversal kernels need no LDS for it (their LDS stack alone
// sets how many blocks fit on a CU) and no block barrier.
template <int WHAT>
struct WaveChunks {
    static constexpr int R = RTX_WF_MAX_SEGS / RTX_WAVE;
    uint32_t end[R];
    uint32_t total;
    __device__ __forceinline__ void init(const WfSeg* __restrict__ segs, uint32_t nseg, uint32_t ch) {
        const uint32_t lane = lane_id();
        uint32_t carry = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t i = (uint32_t)r * RTX_WAVE + lane;
            uint32_t v = i < nseg ? (uint32_t)((wf_items<WHAT>(segs[i]) + ch - 1) / ch) : 0u;
#pragma unroll
            for (int off = 1; off < RTX_WAVE; off <<= 1) {
                const uint32_t u = __shfl_up(v, off, RTX_WAVE);
                if (lane >= (uint32_t)off) v += u;
            }
            end[r] = carry + v;
            carry += __shfl(v, RTX_WAVE - 1, RTX_WAVE);
        }
        total = __builtin_amdgcn_readfirstlane(carry);
    }
    // The table as the step's first kernel left it: ends[i] for i < nseg.
    __device__ __forceinline__ void load(const uint32_t* __restrict__ ends, uint32_t nseg, uint32_t total_) {
        const uint32_t lane = lane_id();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t i = (uint32_t)r * RTX_WAVE + lane;
            end[r] = i < nseg ? ends[i] : total_;
        }
        total = total_;
    }
    // Segment of chunk c < total (wave-uniform) and its first chunk.
    __device__ __forceinline__ uint32_t find(uint32_t c, uint32_t& first) const {
        uint32_t prev = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint64_t m = ballot(end[r] > c);
            if (m) {
                const int k = __builtin_ctzll(m);
                first = k ? (uint32_t)__builtin_amdgcn_readlane((int)end[r], k - 1) : prev;
                return (uint32_t)(r * RTX_WAVE + k);
            }
            prev = (uint32_t)__builtin_amdgcn_readlane((int)end[r], RTX_WAVE - 1);
        }
        first = prev;
        return 0u;  // not reached for c < total
    }
};

// Chunk schedule of a persistent traversal launch: a queue per XCD.
// Work-groups are dispatched to the 8 XCDs round robin (group g on XCD
// g mod 8), each XCD with its own L2.  The chunks are split into 8
// contiguous ranges; the waves of XCD x take chunks of range x in order from
// its own counter, so each L2 sees a compact window of neighbouring chunks
// (one scene at a time) and each counter takes an eighth of the atomics
// (device-scope atomics on one address serialize, about 10 ns each).  A wave
// whose range is drained takes chunks of the next ranges.  Every wave calls
// next() until it returns false; the chunks taken partition [0, total).
constexpr int WF_XCDS = 8;
constexpr int WF_QSTRIDE = 16;  // counters 64 bytes apart
struct ChunkSched {
    uint32_t total, grain, lo, hi;
    int home, k;
    uint32_t* queue;  // WF_XCDS counters, WF_QSTRIDE words apart
    // grain: consecutive chunks taken per atomic
    __device__ __forceinline__ void init(uint32_t total_, uint32_t* q, uint32_t grain_) {
        total = total_;
        grain = grain_ ? grain_ : 1u;
        lo = hi = 0;
        home = (int)(blockIdx.x % WF_XCDS);
        k = 0;
        queue = q;
    }
    // (Issuing the next chunk's atomic ahead was measured slower: the
    // atomic's round trip is longer than an L2-hit node load, and the
    // in-order memory counter then holds the chunk's first loads behind it.)
    __device__ __forceinline__ bool next(uint32_t& c) {
        if (lo < hi) {
            c = lo++;
            return true;
        }
        while (k < WF_XCDS) {
            const int x = (home + k) % WF_XCDS;
            const uint32_t b = (uint32_t)((uint64_t)total * x / WF_XCDS);
            const uint32_t e = (uint32_t)((uint64_t)total * (x + 1) / WF_XCDS);
            uint32_t t = 0;
            if ((threadIdx.x & (RTX_WAVE - 1)) == 0) t = atomicAdd(&queue[x * WF_QSTRIDE], grain);
            t = __builtin_amdgcn_readfirstlane(t);
            if (t < e - b) {
                lo = b + t;
                hi = e - lo < grain ? e : lo + grain;
                c = lo++;
                return true;
            }
            ++k;  // range x is drained
        }
        return false;
    }
};
