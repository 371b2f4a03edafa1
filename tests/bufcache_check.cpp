// Host-only check of the server's block cache (opencliper_amd/csrc/ocl_bufcache.hpp) over a fake device: a seeded
// random stream of request admissions (first come first served under the budget, as the server admits) and
// completions.  After every step: the blocks of live requests add up to `held`; held + cached bytes stay within the
// budget; the fake device holds exactly those bytes (nothing leaked or freed twice); a reused block is at most 4/3 of
// the request.  At the end every block is given back and cleared.  Exit 0 and "ok ..." on success, 1 otherwise.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <set>
#include <vector>

#include "ocl_bufcache.hpp"

struct FakeDev {
    static constexpr int kOom = 2;
    size_t live = 0, capacity = 0;
    long syncs = 0;
    std::set<void*> blocks;
    std::vector<std::pair<void*, size_t>>* sizes = nullptr;
    int alloc(void** p, size_t n) {
        if (live + n > capacity) return kOom;
        *p = std::malloc(n);
        if (!*p) return kOom;
        blocks.insert(*p);
        sizes->push_back({*p, n});
        live += n;
        return 0;
    }
    int release(void* p) {
        if (!blocks.count(p)) return 3;            // freed twice or never allocated
        size_t n = 0;
        for (auto& e : *sizes)
            if (e.first == p) { n = e.second; e.first = nullptr; break; }
        blocks.erase(p);
        live -= n;
        std::free(p);
        return 0;
    }
    int sync() { ++syncs; return 0; }
};

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_rand() {           // splitmix64
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

#define CHECK(cond, ...)                                            \
    do {                                                            \
        if (!(cond)) {                                              \
            std::printf("step %ld: ", step);                        \
            std::printf(__VA_ARGS__);                               \
            std::printf("\n");                                      \
            return 1;                                               \
        }                                                           \
    } while (0)

int main(int argc, char** argv) {
    const long steps = argc > 1 ? std::atol(argv[1]) : 20000;
    if (argc > 2) rng_state = std::strtoull(argv[2], nullptr, 10);
    // request sizes in 256-byte units, as the server rounds them: some close together (reuse within 4/3), some far
    const size_t kinds[] = {4000, 4100, 4900, 6000, 1200, 900, 12000, 2500};
    const size_t limit = 40000 * 256;
    std::vector<std::pair<void*, size_t>> sizes;
    BlockCache<FakeDev> cache;
    cache.limit = limit;
    cache.dev.capacity = 4 * limit;             // the device itself has room: only the cache's own bound applies
    cache.dev.sizes = &sizes;
    struct Req { void* p; size_t got; };
    std::vector<Req> live;
    std::deque<size_t> queue;
    size_t used = 0;
    long step = 0;
    for (step = 0; step < steps; ++step) {
        if (next_rand() % 3) queue.push_back(kinds[next_rand() % 8] * 256);
        // admission in arrival order while the request fits the budget
        while (!queue.empty() && used + queue.front() <= limit) {
            const size_t need = queue.front();
            void* p = nullptr;
            size_t got = 0;
            const int e = cache.take(&p, need, &got);
            CHECK(e == 0, "take(%zu) failed with %d", need, e);
            CHECK(got >= need && got <= need + need / 3, "block of %zu bytes for a request of %zu", got, need);
            std::memset(p, 0x5A, need);         // the request writes its whole block (ASan checks the bounds)
            live.push_back({p, got});
            used += got;
            queue.pop_front();
        }
        // completions
        if (!live.empty() && next_rand() % 2) {
            const size_t i = next_rand() % live.size();
            cache.give(live[i].p, live[i].got);
            used -= live[i].got;
            live.erase(live.begin() + (long)i);
        }
        size_t sum = 0;
        for (const Req& r : live) sum += r.got;
        CHECK(sum == cache.held, "live blocks %zu bytes, held %zu", sum, cache.held);
        CHECK(sum == used, "live blocks %zu bytes, admitted %zu", sum, used);
        CHECK(cache.held + cache.bytes <= limit, "held %zu + cached %zu over the budget %zu", cache.held,
              cache.bytes, limit);
        CHECK(cache.dev.live == cache.held + cache.bytes, "device holds %zu, cache accounts %zu", cache.dev.live,
              cache.held + cache.bytes);
    }
    for (const Req& r : live) cache.give(r.p, r.got);
    cache.clear();
    CHECK(cache.held == 0 && cache.bytes == 0 && cache.dev.live == 0, "left over: held %zu cached %zu device %zu",
          cache.held, cache.bytes, cache.dev.live);
    CHECK(cache.hits > 0 && cache.misses > 0, "hits %ld misses %ld", cache.hits, cache.misses);
    std::printf("ok steps %ld hits %ld misses %ld syncs %ld\n", steps, cache.hits, cache.misses, cache.dev.syncs);
    return 0;
}
