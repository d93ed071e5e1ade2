// Randomised check of rpx::BlockHeap over malloc: tables consistent after
// every operation, live blocks never overlap (each is filled with its id and
// verified when freed), chunks stay within the limit whenever no live block
// needs more, every chunk is returned at the end.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <map>
#include <random>
#include <vector>
#include "rpx_heap.h"

static long g_gets = 0, g_puts = 0, g_fail_after = -1;

struct Malloc {
    void *get(uint64_t bytes) {
        if (g_fail_after == 0) return nullptr;
        if (g_fail_after > 0) --g_fail_after;
        ++g_gets;
        return aligned_alloc(4096, (bytes + 4095) / 4096 * 4096);
    }
    void put(void *p) {
        ++g_puts;
        free(p);
    }
};

#define REQUIRE(c)                                                   \
    do {                                                             \
        if (!(c)) {                                                  \
            fprintf(stderr, "FAIL line %d: %s (seed %u, op %d)\n",   \
                    __LINE__, #c, seed, op);                         \
            return 1;                                                \
        }                                                            \
    } while (0)

static int run(unsigned seed) {
    std::mt19937_64 rng(seed);
    const uint64_t keep = 64ull << 20, min_chunk = 8ull << 20;
    int op = -1;
    {
        rpx::BlockHeap<Malloc> h;
        h.set_limits(keep, min_chunk);
        struct Live { uint64_t bytes; unsigned char tag; };
        std::map<char *, Live> live;
        uint64_t live_bytes = 0;
        for (op = 0; op < 4000; ++op) {
            const bool do_alloc = live.empty() || (rng() % 100) < 55;
            if (do_alloc) {
                // sizes from 1 byte to 24 MiB, log-uniform
                const double e = std::uniform_real_distribution<double>(0.0, 24.5)(rng);
                const uint64_t bytes = std::max<uint64_t>(1, (uint64_t)std::exp2(e));
                if (live_bytes + bytes > (200ull << 20)) continue;
                char *p = (char *)h.alloc(bytes);
                REQUIRE(p != nullptr);
                REQUIRE(((uintptr_t)p) % 256 == 0);
                const unsigned char tag = (unsigned char)(1 + rng() % 255);
                memset(p, tag, bytes);
                auto nx = live.lower_bound(p);
                if (nx != live.end()) REQUIRE(p + bytes <= nx->first);
                if (nx != live.begin()) {
                    auto pv = std::prev(nx);
                    REQUIRE(pv->first + pv->second.bytes <= p);
                }
                live[p] = {bytes, tag};
                live_bytes += bytes;
            } else {
                auto it = live.begin();
                std::advance(it, rng() % live.size());
                const Live l = it->second;
                for (uint64_t k = 0; k < l.bytes; k += 1 + l.bytes / 64) REQUIRE((unsigned char)it->first[k] == l.tag);
                REQUIRE((unsigned char)it->first[l.bytes - 1] == l.tag);
                REQUIRE(h.free(it->first));
                live_bytes -= l.bytes;
                live.erase(it);
            }
            REQUIRE(h.check());
            REQUIRE(h.used_bytes() >= live_bytes);
            // above the limit only while live blocks hold what is there
            if (h.chunk_bytes() > keep) {
                h.release_free_chunks(keep);
                REQUIRE(h.check());
            }
        }
        REQUIRE(!h.free((void *)0x1000));   // not handed out
        for (auto &kv : live) REQUIRE(h.free(kv.first));
        REQUIRE(h.check());
        REQUIRE(h.used_bytes() == 0);
        REQUIRE(h.n_free_blocks() == h.n_chunks());   // each chunk merged back into one block
        h.release_free_chunks(0);
        REQUIRE(h.n_chunks() == 0 && h.chunk_bytes() == 0);
    }
    REQUIRE(g_gets == g_puts);
    // a backend that refuses: alloc returns nullptr, the heap stays usable
    {
        rpx::BlockHeap<Malloc> h;
        h.set_limits(keep, min_chunk);
        void *a = h.alloc(1 << 20);
        REQUIRE(a != nullptr);
        g_fail_after = 0;
        REQUIRE(h.alloc(32ull << 20) == nullptr);
        void *b = h.alloc(1 << 20);   // served from the first chunk
        REQUIRE(b != nullptr);
        g_fail_after = -1;
        REQUIRE(h.check());
        REQUIRE(h.free(a) && h.free(b));
    }
    REQUIRE(g_gets == g_puts);
    return 0;
}

int main(int argc, char **argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 4;
    for (int s = 0; s < n; ++s)
        if (run(1234u + 7u * (unsigned)s)) return 1;
    printf("heap ok: %d seeds\n", n);
    return 0;
}
