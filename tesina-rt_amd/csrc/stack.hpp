// Per-lane traversal stacks of the BVH walks.
//
// The walk pushes the far child's node id at the current level and descends;
// a pop returns the entry of the level below.  STACK is the reference's
// MAX_LEVELS (src/kernel/trace-bvh.cl:330): a pop at level >= STACK ends the
// ray (src/kernel/trace-bvh.cl:361-367), so entries pushed at levels >= STACK
// are never read.  Both stacks keep exactly these semantics; they differ only
// in where the entries live.
//
// The methods are host-callable too, so that tests/cpp/stack_test.cpp can
// run the very same code on the host against a plain array, with every
// access bounds-checked.
#pragma once

#include <cstdint>
#include <type_traits>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define RTX_STACK_FN __host__ __device__ __forceinline__
// the spill words: private memory, addressed with scratch instructions;
// volatile keeps them in memory rather than in registers indexed per lane
using StackPrivate = volatile __attribute__((address_space(5))) uint32_t*;
#else
#define RTX_STACK_FN inline
using StackPrivate = volatile uint32_t*;
#endif

constexpr int STACK = 32;

// The whole stack in LDS, [entry][lane] (bank-conflict free): STACK * BLOCK
// words per block.  put(lv, id) takes the unmasked push level; get(lv) the
// level lv < STACK.  reset() starts a ray.
template <int BLOCK>
struct LdsStack {
    uint32_t* id;
    uint32_t lane;
    RTX_STACK_FN LdsStack(uint32_t* lds, uint32_t l, uint32_t*) : id(lds), lane(l) {}
    RTX_STACK_FN void reset() {}
    RTX_STACK_FN void put(uint32_t lv, uint32_t i) { id[(lv & (STACK - 1)) * BLOCK + lane] = i; }
    RTX_STACK_FN uint32_t get(uint32_t lv) { return id[(lv & (STACK - 1)) * BLOCK + lane]; }
};

// The same stack with only its top R levels in LDS: levels [floor, level)
// sit in an R-entry ring in LDS (R * BLOCK words per block), levels below
// floor in STACK - R words of the lane's private memory.  A push that finds
// the ring full first moves the ring's oldest entry (level floor) out; a pop
// below floor reads private memory and leaves the ring empty.  The LDS per
// wave drops from STACK to R entries, so more waves fit on a CU; private
// memory is touched only when a ray's path goes more than R levels below the
// last level it popped back to.
//
// Bounds, given the walk's protocol (put at the current level, then level+1;
// get at level-1 only when 0 < level < STACK; reset at each ray):
//   floor <= level at all times, and put keeps level - floor <= R;
//   a spill writes low[floor] with floor = lv - R <= STACK - 1 - R;
//   a reload reads low[lv] with lv < floor <= STACK - R;
//   the ring index is (level mod R) < R.
template <int BLOCK, int R>
struct RingStack {
    static_assert(R > 0 && R < STACK && (R & (R - 1)) == 0, "ring size: a power of two below STACK");
    uint32_t* id;
    uint32_t lane;
    uint32_t floor = 0;
    StackPrivate low;
    RTX_STACK_FN RingStack(uint32_t* lds, uint32_t l, uint32_t* priv) : id(lds), lane(l), low((StackPrivate)priv) {}
    RTX_STACK_FN void reset() { floor = 0; }
    RTX_STACK_FN void put(uint32_t lv, uint32_t i) {
        if (lv >= (uint32_t)STACK) return;
        if (lv - floor == (uint32_t)R) {
            low[floor] = id[(floor & (R - 1)) * BLOCK + lane];
            ++floor;
        }
        id[(lv & (R - 1)) * BLOCK + lane] = i;
    }
    RTX_STACK_FN uint32_t get(uint32_t lv) {
        if (lv >= floor) return id[(lv & (R - 1)) * BLOCK + lane];
        floor = lv;
        return low[lv];
    }
};

// Stack of a block of BLOCK threads with SR entries per lane in LDS (SR =
// STACK: the whole stack), and the words of its private part.
template <int BLOCK, int SR>
using TraversalStack = typename std::conditional<(SR >= STACK), LdsStack<BLOCK>, RingStack<BLOCK, SR>>::type;
template <int SR>
constexpr int stack_private_words() {
    return SR >= STACK ? 1 : STACK - SR;
}
