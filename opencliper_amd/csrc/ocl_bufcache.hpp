// Device memory a server keeps between requests, so that admission and completion call neither hipMalloc (about
// 0.3 ms per buffer) nor hipFree (which waits for the whole device) while requests come and go.  Each request takes
// one block and carves its work buffers from it; a block given back is reused by a later request that needs at least
// three quarters of it (the smallest such block).  `held` counts the blocks of live requests and `bytes` the cached
// ones; held + bytes stays within `limit` (the server's HBM budget): a reuse moves a block from one count to the
// other, and a new block first frees cached ones, largest first, until it fits.  Used by one thread only.
//
// Dev supplies the device calls (all returning 0 on success): alloc(void**, size_t) -> 0, or kOom when memory is
// exhausted; release(void*); sync() waits for the work that may still use cached blocks before any is freed.  Every
// user of the blocks is on one stream, so a block given back may be handed out again at once (stream order).
#pragma once
#include <cstddef>
#include <iterator>
#include <map>

template <class Dev>
struct BlockCache {
    Dev dev;
    std::multimap<size_t, void*> free;
    size_t bytes = 0, held = 0, limit = 0;
    long hits = 0, misses = 0;

    // a block of at least n bytes in *p, its size in *got
    int take(void** p, size_t n, size_t* got) {
        auto it = free.lower_bound(n);
        if (it != free.end() && it->first <= n + n / 3 && held + it->first <= limit) {
            *p = it->second;
            *got = it->first;
            bytes -= it->first;
            held += it->first;
            free.erase(it);
            hits++;
            return 0;
        }
        if (held + bytes + n > limit && !free.empty()) {
            int e = dev.sync();
            if (e) return e;
            while (held + bytes + n > limit && !free.empty()) {
                auto big = std::prev(free.end());
                e = dev.release(big->second);
                bytes -= big->first;
                free.erase(big);
                if (e) return e;
            }
        }
        const int e = dev.alloc(p, n);
        if (e == 0) {
            *got = n;
            held += n;
            misses++;
        }
        return e;
    }
    void give(void* p, size_t n) {
        free.emplace(n, p);
        bytes += n;
        held -= n;
    }
    void clear() {
        if (free.empty()) return;
        (void)dev.sync();
        for (auto& kv : free) (void)dev.release(kv.second);
        free.clear();
        bytes = 0;
    }
};
