// Host check of the device block cache (horkxg_amd/csrc/hip/block_cache.hpp)
// over a fake segment source that hands out unmapped addresses. Exits non-zero with a
// message on the first violated property.
#include "block_cache.hpp"

#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

using gpart_hip::BlockCache;

#define REQUIRE(cond)                                                       \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            std::exit(1);                                                   \
        }                                                                   \
    } while (0)

// Hands out addresses from an unmapped range: the cache never dereferences them.
struct FakeDevice {
    std::size_t capacity;
    std::size_t used = 0;
    std::uintptr_t next = std::uintptr_t{1} << 44;
    std::map<void*, std::size_t> live;
    int drains = 0;
    gpart_hip::SegmentSource source() {
        gpart_hip::SegmentSource s;
        s.allocate = [this](std::size_t bytes) -> void* {
            if (used + bytes > capacity) return nullptr;
            void* p = reinterpret_cast<void*>(next);
            next += (bytes + 4095) / 4096 * 4096 + 4096;
            used += bytes;
            live.emplace(p, bytes);
            return p;
        };
        s.release = [this](void* p) {
            auto it = live.find(p);
            REQUIRE(it != live.end());
            used -= it->second;
            live.erase(it);
        };
        s.drain = [this] { ++drains; };
        return s;
    }
};

// Live blocks lie inside one segment each and never overlap.
void check_layout(const BlockCache& cache, const std::map<char*, std::size_t>& blocks) {
    const auto& segs = cache.segments();
    char* prev_end = nullptr;
    for (const auto& [p, size] : blocks) {
        REQUIRE(cache.block_size(p) == size);
        auto seg = segs.upper_bound(p);
        REQUIRE(seg != segs.begin());
        --seg;
        REQUIRE(p >= seg->first && p + size <= seg->first + seg->second);
        REQUIRE((p - seg->first) % BlockCache::kAlign == 0);
        REQUIRE(prev_end == nullptr || p >= prev_end);
        prev_end = p + size;
    }
}

void test_random_sequences() {
    const std::size_t MiB = std::size_t{1} << 20;
    for (unsigned seed = 0; seed < 6; ++seed) {
        FakeDevice dev{std::size_t{256} << 30};
        BlockCache cache(dev.source());
        std::mt19937_64 rng(seed);
        std::map<char*, std::size_t> blocks;
        std::vector<char*> order;
        std::uint64_t live = 0;
        for (int step = 0; step < 4000; ++step) {
            const bool alloc = order.empty() || rng() % 100 < 50;
            if (alloc) {
                std::size_t bytes;
                switch (rng() % 4) {
                    case 0: bytes = 1 + rng() % 4096; break;
                    case 1: bytes = 1 + rng() % MiB; break;
                    case 2: bytes = MiB + rng() % (64 * MiB); break;
                    default: bytes = (rng() % 3 == 0) ? (1 + rng() % 512) * MiB : 1 + rng() % (8 * MiB); break;
                }
                char* p = static_cast<char*>(cache.allocate(bytes));
                const std::size_t size = cache.block_size(p);
                REQUIRE(size >= bytes);
                REQUIRE(size - bytes < BlockCache::kAlign + (bytes > BlockCache::kSmallLimit ? BlockCache::kLargeMinSplit : 0));
                REQUIRE(blocks.emplace(p, size).second);
                order.push_back(p);
                live += size;
            } else {
                const std::size_t i = rng() % order.size();
                char* p = order[i];
                order[i] = order.back();
                order.pop_back();
                live -= blocks.at(p);
                blocks.erase(p);
                cache.free(p);
            }
            REQUIRE(cache.stats().live_bytes == live);
            REQUIRE(cache.stats().live_blocks == static_cast<std::int64_t>(blocks.size()));
            REQUIRE(cache.stats().reserved_bytes == dev.used);
            REQUIRE(cache.stats().live_bytes + cache.free_bytes() == cache.stats().reserved_bytes);
            if (step % 250 == 0) check_layout(cache, blocks);
        }
        check_layout(cache, blocks);
        for (char* p : order) cache.free(p);
        REQUIRE(cache.stats().live_bytes == 0);
        // fully coalesced: one free block per segment, all of them released
        REQUIRE(cache.free_blocks() == cache.stats().segments);
        const auto segments = cache.stats().segments;
        REQUIRE(cache.release_free_segments() == segments);
        REQUIRE(cache.stats().reserved_bytes == 0 && dev.used == 0 && dev.live.empty());
        REQUIRE(dev.drains == 1);
    }
}

void test_best_fit_and_coalescing() {
    const std::size_t MiB = std::size_t{1} << 20;
    FakeDevice dev{std::size_t{64} << 30};
    BlockCache cache(dev.source());
    char* a = static_cast<char*>(cache.allocate(8 * MiB));
    char* b = static_cast<char*>(cache.allocate(3 * MiB));
    char* c = static_cast<char*>(cache.allocate(16 * MiB));
    char* d = static_cast<char*>(cache.allocate(4 * MiB));
    REQUIRE(cache.stats().segments == 1);  // all inside the first 1 GiB segment
    REQUIRE(b == a + 8 * MiB && c == b + 3 * MiB && d == c + 16 * MiB);
    cache.free(b);
    cache.free(d);  // merges with the segment's tail
    // best fit: the 3 MiB hole, not the tail
    char* e = static_cast<char*>(cache.allocate(2 * MiB));
    REQUIRE(e == b);
    cache.free(e);
    cache.free(c);  // b + c coalesce with the tail into one block after a
    char* f = static_cast<char*>(cache.allocate(19 * MiB));
    REQUIRE(f == b);
    cache.free(f);
    cache.free(a);
    REQUIRE(cache.free_blocks() == 1);
    // a request above the segment size gets a segment of its own, rounded to 2 MiB
    char* g = static_cast<char*>(cache.allocate((std::size_t{3} << 30) + 1));
    REQUIRE(cache.stats().segments == 2);
    REQUIRE(cache.segments().at(g) == (std::size_t{3} << 30) + 2 * MiB);
    cache.free(g);
    // small requests come from small segments
    char* s = static_cast<char*>(cache.allocate(1000));
    REQUIRE(cache.segments().count(s) == 1 && cache.segments().at(s) == BlockCache::kSmallSegment);
    REQUIRE(cache.block_size(s) == 1024);
    cache.free(s);
    REQUIRE(cache.stats().segment_allocations == 3);
    // unknown and double frees are refused
    bool threw = false;
    try { cache.free(s); } catch (const std::invalid_argument&) { threw = true; }
    REQUIRE(threw);
    threw = false;
    try { cache.free(s + 512); } catch (const std::invalid_argument&) { threw = true; }
    REQUIRE(threw);
}

void test_out_of_memory() {
    const std::size_t GiB = std::size_t{1} << 30;
    FakeDevice dev{4 * GiB + (std::size_t{768} << 20)};
    BlockCache cache(dev.source());
    void* a = cache.allocate(3 * GiB);        // segment of 3 GiB
    cache.free(a);                            // cached, segment kept
    void* b = cache.allocate(GiB / 2);        // served from the 3 GiB segment
    cache.free(b);
    REQUIRE(cache.stats().segment_allocations == 1);
    // 4 GiB does not fit next to the cached 3 GiB: the free segment is released
    void* c = cache.allocate(4 * GiB);
    REQUIRE(cache.stats().out_of_memory_releases == 1);
    REQUIRE(cache.stats().segment_releases == 1);
    REQUIRE(dev.used == 4 * GiB);
    // with 4 GiB live, a further 2 GiB request fails loudly
    bool threw = false;
    try { cache.allocate(2 * GiB); } catch (const std::runtime_error&) { threw = true; }
    REQUIRE(threw);
    // near the limit the fallback asks for the rounded request, not a 1 GiB segment
    void* d = cache.allocate(std::size_t{600} << 20);
    REQUIRE(cache.segments().at(static_cast<char*>(d)) == (std::size_t{600} << 20));
    cache.free(c);
    cache.free(d);
    cache.release_free_segments();
    REQUIRE(dev.used == 0);
}

// The first thread that allocates owns the cache: another thread's allocate,
// free or release is refused with an exception and leaves the cache intact;
// rebind_owner() hands it over.
void test_thread_ownership() {
    FakeDevice dev{std::size_t{8} << 30};
    BlockCache cache(dev.source());
    void* a = cache.allocate(1000);
    int refused = 0;
    std::thread other([&] {
        try {
            cache.allocate(64);
        } catch (const std::logic_error&) {
            ++refused;
        }
        try {
            cache.free(a);
        } catch (const std::logic_error&) {
            ++refused;
        }
        try {
            cache.release_free_segments();
        } catch (const std::logic_error&) {
            ++refused;
        }
    });
    other.join();
    REQUIRE(refused == 3);
    REQUIRE(cache.stats().live_blocks == 1);
    std::thread taker([&] {
        cache.rebind_owner();
        cache.free(a);
        cache.release_free_segments();
    });
    taker.join();
    REQUIRE(cache.stats().live_blocks == 0 && dev.used == 0);
}

int main() {
    test_best_fit_and_coalescing();
    test_out_of_memory();
    test_random_sequences();
    test_thread_ownership();
    std::puts("block cache checks passed");
    return 0;
}
