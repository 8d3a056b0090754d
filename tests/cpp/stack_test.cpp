// Host test of the traversal stacks of tesina-rt_amd/csrc/stack.hpp, the code
// the kernels use.  Push/pop sequences that follow the walk's protocol (put at
// the current level then level+1; get at level-1 only while 0 < level <
// STACK, else the ray ends; reset at every ray) drive a whole-LDS stack and
// ring stacks of 8 and 16 entries side by side: every pop must return the
// same id.  The buffers are heap arrays of exactly the sizes the kernels
// declare (R * BLOCK LDS words, STACK - R private words), so under
// AddressSanitizer any access outside them aborts the run.
//
// Sequences: every sequence of up to 20 operations (exhaustive), and long
// random walks that go past level STACK, end rays early and reset.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "stack.hpp"

namespace {

constexpr int BLOCK = 64;

struct Pair {
    uint32_t* lds_full;
    uint32_t* lds_ring;
    uint32_t* priv;
    uint32_t lane;
};

int failures = 0;

template <int R>
struct Walk {
    LdsStack<BLOCK> full;
    RingStack<BLOCK, R> ring;
    uint32_t level = 0, next_id = 1;
    uint64_t pops = 0;
    explicit Walk(const Pair& p) : full(p.lds_full, p.lane, nullptr), ring(p.lds_ring, p.lane, p.priv) {}
    void start() {
        full.reset();
        ring.reset();
        level = 0;
    }
    void push() {
        const uint32_t v = next_id++;
        full.put(level, v);
        ring.put(level, v);
        level++;
    }
    // a pop, or the end of the ray where the walk would end it
    void pop() {
        if (level == 0 || level >= (uint32_t)STACK) {
            start();
            return;
        }
        --level;
        const uint32_t a = full.get(level), b = ring.get(level);
        ++pops;
        if (a != b) {
            if (failures++ < 10) std::printf("R=%d level %u: ring stack %u, whole stack %u\n", R, level, b, a);
        }
    }
};

template <int R>
void exhaustive(const Pair& p, int len) {
    for (uint32_t bits = 0; bits < (1u << len); ++bits) {
        Walk<R> w(p);
        w.start();
        for (int k = 0; k < len; ++k) {
            if ((bits >> k) & 1u)
                w.push();
            else
                w.pop();
        }
        while (w.level > 0 && w.level < (uint32_t)STACK) w.pop();  // unwind every entry
    }
}

template <int R>
void random_walks(const Pair& p, uint64_t seed, int walks, int ops) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    uint64_t pops = 0;
    uint32_t max_level = 0;
    for (int n = 0; n < walks; ++n) {
        Walk<R> w(p);
        w.start();
        const double push_p = 0.45 + 0.2 * u(rng);  // some walks dive past STACK
        for (int k = 0; k < ops; ++k) {
            const double x = u(rng);
            if (x < 0.002)
                w.start();  // the lane's ray ends early (an any-hit)
            else if (x < 0.002 + push_p)
                w.push();
            else
                w.pop();
            if (w.level > max_level) max_level = w.level;
        }
        pops += w.pops;
    }
    if (max_level <= (uint32_t)STACK || pops < 1000) {
        std::printf("R=%d: random walks too shallow (max level %u, %llu pops)\n", R, max_level,
                    (unsigned long long)pops);
        failures++;
    }
}

template <int R>
void run(uint32_t lane) {
    Pair p{new uint32_t[STACK * BLOCK], new uint32_t[R * BLOCK], new uint32_t[STACK - R], lane};
    exhaustive<R>(p, 20);
    random_walks<R>(p, 1234 + R + lane, 400, 4000);
    delete[] p.lds_full;
    delete[] p.lds_ring;
    delete[] p.priv;
}

}  // namespace

int main() {
    for (uint32_t lane : {0u, 37u, (uint32_t)BLOCK - 1}) {
        run<8>(lane);
        run<16>(lane);
    }
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
