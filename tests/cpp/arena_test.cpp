// Host test of the wavefront arena (first-fit blocks that merge on release)
// and the admission queue (first come first served; the head waits for
// room; requests larger than the arena are rejected), with a tiny budget.
// Prints "ok" and exits 0, or names the first failed check and exits 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "arena.hpp"

using namespace rtxwf;

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::printf("FAILED line %d: %s\n", __LINE__, #c);     \
            std::exit(1);                                          \
        }                                                          \
    } while (0)

int main() {
    // ---- allocator
    Arena a;
    a.reset(10 * ALIGN + 17);  // rounds down to whole units
    CHECK(a.size() == 10 * ALIGN);
    uint64_t x = a.alloc(1), y = a.alloc(3 * ALIGN), z = a.alloc(2 * ALIGN + 1);
    CHECK(x == 0 && y == ALIGN && z == 4 * ALIGN);
    CHECK(a.used() == 7 * ALIGN && a.largest_free() == 3 * ALIGN);
    CHECK(a.alloc(4 * ALIGN) == NONE);   // 3 units left
    CHECK(a.release(y));
    CHECK(!a.release(y));                // not allocated any more
    CHECK(!a.release(5));                // not a block start
    CHECK(a.largest_free() == 3 * ALIGN);
    CHECK(a.release(x));                 // merges with y's run: 4 units at 0
    CHECK(a.largest_free() == 4 * ALIGN);
    uint64_t w = a.alloc(4 * ALIGN);     // first fit: the merged run at 0
    CHECK(w == 0);
    CHECK(a.release(z) && a.release(w));
    CHECK(a.used() == 0 && a.largest_free() == 10 * ALIGN && a.blocks() == 0);

    // ---- frame layout: buffers inside the region, in order, aligned
    FrameLayout L(256 * 256, 3);
    CHECK(L.P == 65536 && L.cap == (65536ull << 3) + 2);
    CHECK(L.bcap[0] == (65536ull << 2) + 2 && L.bcap[1] == L.cap);
    CHECK(L.ray[0][0] == 0 && L.ray[1][0] > L.ray[0][3] && L.stats + 32 <= L.bytes);
    CHECK(L.accum % ALIGN == 0 && L.status % ALIGN == 0);
    FrameLayout Lc(256 * 256, 3, 1000);  // user capacity below (P << b) + 2
    CHECK(Lc.cap == 65536 && Lc.bcap[1] == 1000 && Lc.bcap[0] == 65536);
    CHECK(bounce_capacity(1u << 29, 3) == (1ull << 30) - 1);

    // ---- admission with a tiny budget: 8 units
    Arena b;
    b.reset(8 * ALIGN);
    Admission q;
    q.push(1, 3 * ALIGN);
    q.push(2, 4 * ALIGN);
    q.push(3, 9 * ALIGN);  // larger than the arena: rejected
    q.push(4, 2 * ALIGN);
    q.push(5, 1 * ALIGN);
    std::vector<uint64_t> adm, rej;
    std::vector<uint64_t> offs(6, NONE);
    auto try_add = [&](const Admission::Item& it) {
        const uint64_t o = b.alloc(it.bytes);
        if (o == NONE) return false;
        offs[it.id] = o;
        return true;
    };
    q.admit(b.size(), 0, 8, try_add, adm, rej);
    // 1 and 2 fit (7 units); 3 is rejected on reaching the head; 4 needs 2
    // units with 1 left: it waits at the head, and 5 (1 unit) does not
    // overtake it
    CHECK(adm.size() == 2 && adm[0] == 1 && adm[1] == 2);
    CHECK(rej.size() == 1 && rej[0] == 3);
    CHECK(q.queue.size() == 2 && q.queue.front().id == 4);
    adm.clear();
    rej.clear();
    q.admit(b.size(), 2, 8, try_add, adm, rej);
    CHECK(adm.empty() && rej.empty());  // still no room
    CHECK(b.release(offs[1]));          // 3 units back
    q.admit(b.size(), 1, 8, try_add, adm, rej);
    CHECK(adm.size() == 2 && adm[0] == 4 && adm[1] == 5 && q.queue.empty());
    // the in-flight limit holds requests back even with room
    Arena c;
    c.reset(100 * ALIGN);
    Admission q2;
    for (uint64_t i = 0; i < 5; ++i) q2.push(i, ALIGN);
    std::vector<uint64_t> adm2, rej2;
    q2.admit(c.size(), 1, 3, [&](const Admission::Item& it) { return c.alloc(it.bytes) != NONE; }, adm2, rej2);
    CHECK(adm2.size() == 2 && q2.queue.size() == 3);
    std::printf("ok\n");
    return 0;
}
