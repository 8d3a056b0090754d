// Device-memory arena of the wavefront renderer and the admission queue of
// its requests.  Host-only code (no HIP calls): offsets into one
// preallocated device buffer, so that admitting and finishing frames never
// calls the allocator.
//
// FrameLayout: where a frame's buffers sit inside its region.  A frame of P
// pixels and b bounces keeps two ray buffers (this bounce's and the next
// one's), one set of hit records, its framebuffer and counters.  Bounce b's
// rays live in buffer b % 2, whose capacity is the largest of the
// stand-alone renderer's bundles for those bounces ((P << b) + 2 for b > 0),
// so the secondary-ray capacity rule drops exactly what the stand-alone
// frame drops.
#pragma once

#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <vector>

namespace rtxwf {

constexpr uint64_t ALIGN = 256;
constexpr uint64_t NONE = ~0ull;

inline uint64_t align_up(uint64_t b) { return (b + ALIGN - 1) & ~(ALIGN - 1); }

// Ray capacity of bounce `bounce` of a frame of P pixels: P for primary rays,
// else (P << bounce) + 2, or the caller's capacity when smaller, at most
// 2^30 - 1 (as rtx_renderer_create sizes its bundles).
inline uint64_t bounce_capacity(uint64_t P, int bounce, uint64_t user_cap = 0) {
    uint64_t cap = bounce == 0 ? P : (P << bounce) + 2;
    if (bounce > 0 && user_cap && user_cap < cap) cap = user_cap;
    const uint64_t lim = (1ull << 30) - 1;
    return cap > lim ? lim : cap;
}

struct FrameLayout {
    uint64_t P = 0, cap = 0;  // cap: largest bounce's rays (hit records)
    uint64_t bcap[2] = {};    // ray buffer k holds the even (0) or odd (1) bounces
    // byte offsets inside the frame's region
    uint64_t ray[2][4] = {};  // ori4, dir4, pixel, contrib of ray buffer 0 and 1
    uint64_t count[2] = {};   // 4 x uint32 per ray buffer
    uint64_t flags = 0, t = 0, id = 0, bary = 0, n4 = 0, p4 = 0, uv2 = 0, list = 0, list_count = 0;
    uint64_t status = 0, accum = 0, rgba = 0, stats = 0;
    uint64_t bytes = 0;

    FrameLayout() = default;
    FrameLayout(uint64_t pixels, int bounces, uint64_t user_cap = 0) {
        P = pixels;
        cap = 0;
        for (int b = 0; b <= bounces; ++b) {
            const uint64_t c = bounce_capacity(P, b, user_cap);
            if (c > cap) cap = c;
            if (c > bcap[b & 1]) bcap[b & 1] = c;
        }
        uint64_t o = 0;
        auto take = [&](uint64_t n) {
            const uint64_t at = o;
            o += align_up(n);
            return at;
        };
        for (int k = 0; k < 2; ++k) {
            ray[k][0] = take(bcap[k] * 16);
            ray[k][1] = take(bcap[k] * 16);
            ray[k][2] = take(bcap[k] * 4);
            ray[k][3] = take(bcap[k] * 4);
            count[k] = take(16);
        }
        flags = take(cap * 4);
        t = take(cap * 4);
        id = take(cap * 4);
        bary = take(cap * 8);
        n4 = take(cap * 16);
        p4 = take(cap * 16);
        uv2 = take(cap * 8);
        list = take(cap * 4);
        list_count = take(16);
        status = take(((cap + 255) / 256 + 1) * 4);
        accum = take(P * 12);
        rgba = take(P * 4);
        stats = take(32);
        bytes = o;
    }
};

// First-fit allocator over [0, size); blocks are ALIGN-aligned and freed
// blocks merge with free neighbours.
class Arena {
public:
    void reset(uint64_t size) {
        size_ = size & ~(ALIGN - 1);
        free_.clear();
        used_.clear();
        used_bytes_ = 0;
        if (size_) free_[0] = size_;
    }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_bytes_; }
    uint64_t largest_free() const {
        uint64_t m = 0;
        for (auto& f : free_) m = f.second > m ? f.second : m;
        return m;
    }
    // Offset of a new block of `bytes`, or NONE when no free run is large enough.
    uint64_t alloc(uint64_t bytes) {
        const uint64_t n = align_up(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const uint64_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_[off] = n;
            used_bytes_ += n;
            return off;
        }
        return NONE;
    }
    // Returns false for an offset that is not an allocated block.
    bool release(uint64_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        uint64_t start = off, len = u->second;
        used_bytes_ -= len;
        used_.erase(u);
        auto next = free_.lower_bound(start);
        if (next != free_.end() && next->first == start + len) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
        return true;
    }
    size_t blocks() const { return used_.size(); }

private:
    uint64_t size_ = 0, used_bytes_ = 0;
    std::map<uint64_t, uint64_t> free_;  // start -> length
    std::map<uint64_t, uint64_t> used_;  // start -> length
};

// Requests waiting for memory, served first come first served: the head
// waits until it fits (later requests do not overtake it, so a large request
// is not starved), and a request larger than the whole arena is rejected at
// once.
struct Admission {
    struct Item {
        uint64_t id, bytes;
    };
    std::deque<Item> queue;

    void push(uint64_t id, uint64_t bytes) { queue.push_back(Item{id, bytes}); }
    // Admits from the head while fewer than `max_inflight` are in flight and
    // try_add(item) finds room for it (returns false: no room now).
    // admitted: ids in order; rejected: ids larger than `arena_size`.
    template <class TryAdd>
    void admit(uint64_t arena_size, size_t inflight, size_t max_inflight, TryAdd try_add,
               std::vector<uint64_t>& admitted, std::vector<uint64_t>& rejected) {
        while (!queue.empty()) {
            const Item it = queue.front();
            if (align_up(it.bytes) > arena_size) {
                rejected.push_back(it.id);
                queue.pop_front();
                continue;
            }
            if (inflight >= max_inflight || !try_add(it)) return;
            admitted.push_back(it.id);
            queue.pop_front();
            ++inflight;
        }
    }
};

}  // namespace rtxwf
