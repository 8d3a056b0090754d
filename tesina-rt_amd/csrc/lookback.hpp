// Decoupled look-back over tiles acquired in order (single-pass ordered
// scan).  Each tile publishes its aggregate, accumulates its predecessors'
// values back to the nearest inclusive prefix, then publishes its own
// inclusive prefix.  A status word carries flag and value together
// (RTX_LB_AGG / RTX_LB_INC | 30-bit value), so relaxed device-scope atomics
// suffice.  The word of tile t lives at status[t * stride].
#pragma once

#include "devmath.hpp"
#include "rtx_impl.hpp"

// One thread walks back on its own (used where every thread of a block runs
// its own look-back, e.g. one per radix digit).  It reads the words of W
// predecessors (k, k-1, ..., k-W+1) in one round trip and consumes them in
// order up to the first that is not ready or holds an inclusive prefix, so a
// run of aggregate-only predecessors costs one round trip per W tiles rather
// than per tile.  Words below tile 0 are never consumed: tile 0's word becomes
// inclusive, and the walk stops there.
template <int W = 8>
__device__ __forceinline__ uint32_t lookback_thread(uint32_t* status, uint32_t tile, uint32_t agg,
                                                    uint32_t stride) {
    if (tile == 0) {
        st_relaxed(&status[0], RTX_LB_INC | agg);
        return 0;
    }
    st_relaxed(&status[(size_t)tile * stride], RTX_LB_AGG | agg);
    uint32_t acc = 0;
    int k = (int)tile - 1;
    for (;;) {
        uint32_t s[W];
#pragma unroll
        for (int j = 0; j < W; ++j) s[j] = k - j >= 0 ? ld_relaxed(&status[(size_t)(k - j) * stride]) : 0u;
        int used = 0;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (!(s[j] & (RTX_LB_AGG | RTX_LB_INC))) break;  // not ready: wait for it
            acc += s[j] & RTX_LB_VAL;
            if (s[j] & RTX_LB_INC) {
                st_relaxed(&status[(size_t)tile * stride], RTX_LB_INC | (acc + agg));
                return acc;
            }
            ++used;
        }
        if (used == 0) __builtin_amdgcn_s_sleep(1);
        k -= used;
    }
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, RTX_WAVE);
    return v;
}

// All 64 lanes of one wave walk back together: lane i reads predecessor
// hi - i, so one round trip covers 64 tiles.  The window's values are summed
// up to the nearest inclusive prefix when every word before it is ready;
// otherwise the ready words in front of the first empty one are consumed
// and the window moves on.  Returns the exclusive prefix on every lane.
__device__ __forceinline__ uint32_t lookback_wave(uint32_t* status, uint32_t tile, uint32_t agg, uint32_t stride) {
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0) st_relaxed(&status[0], RTX_LB_INC | agg);
        return 0;
    }
    if (lane == 0) st_relaxed(&status[(size_t)tile * stride], RTX_LB_AGG | agg);
    uint32_t acc = 0;
    int64_t hi = (int64_t)tile - 1;  // nearest predecessor not yet consumed
    for (;;) {
        const int64_t k = hi - (int64_t)lane;
        const uint32_t s = k >= 0 ? ld_relaxed(&status[(size_t)k * stride]) : 0u;
        const uint64_t inc = ballot((s & RTX_LB_INC) != 0u);
        const uint64_t empty = ballot((s & (RTX_LB_AGG | RTX_LB_INC)) == 0u);
        const int first_inc = inc ? __builtin_ctzll(inc) : 64;
        const int first_empty = empty ? __builtin_ctzll(empty) : 64;
        if (first_inc < first_empty) {  // lanes 0..first_inc all ready: done
            acc += wave_sum_u32((int)lane <= first_inc ? (s & RTX_LB_VAL) : 0u);
            break;
        }
        if (first_empty > 0) {  // consume the ready aggregates in front
            acc += wave_sum_u32((int)lane < first_empty ? (s & RTX_LB_VAL) : 0u);
            hi -= first_empty;
        } else {
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (lane == 0) st_relaxed(&status[(size_t)tile * stride], RTX_LB_INC | (acc + agg));
    return acc;
}
