// block_cache.hpp -- best-fit caching of device memory in segments held by the
// context.
//
// Device memory that the GPU kernel driver hands out again after a free is scrubbed on the
// way (measured on MI355X under ROCm 7.2: once roughly 64 GiB have been recycled,
// every further GiB costs about 30 ms of host time inside the allocation call).
// A multilevel hierarchy of a billion-entry graph needs ~200 GB, allocated and
// freed level by level, so handing memory back to the kernel driver between levels or
// between runs costs seconds. The cache therefore keeps every segment it takes
// until release_free_segments() (an explicit release, context teardown, or an
// allocation the device cannot serve otherwise).
//
// Requests of up to kSmallLimit bytes are served from kSmallSegment segments,
// larger ones from segments of at least kLargeSegment bytes, so long-lived small
// blocks do not pin large segments. A request takes the smallest free block that
// fits (best fit) and splits off the rest; a freed block merges with free address
// neighbours of its segment, so a fully freed segment is one block again.
//
// Host-only code: the segment source supplies and returns raw device memory.
// Not thread-safe: the first thread that allocates owns the cache, and a call from
// any other thread is refused with an exception instead of racing on the free
// lists (rebind_owner() hands the cache over, as a context's teardown on another
// thread does).
#pragma once

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>

namespace gpart_hip {

struct SegmentSource {
    std::function<void*(std::size_t)> allocate;  // nullptr when the device cannot serve it
    std::function<void(void*)> release;          // only segments without live blocks
    std::function<void()> drain;                 // waits until queued work no longer touches freed blocks
};

struct BlockCacheStats {
    std::uint64_t reserved_bytes = 0;       // held in segments
    std::uint64_t live_bytes = 0;           // in live blocks (rounded sizes)
    std::uint64_t peak_reserved_bytes = 0;
    std::uint64_t peak_live_bytes = 0;
    std::int64_t segments = 0;
    std::int64_t segment_allocations = 0;   // successful source allocations, all time
    double segment_allocation_seconds = 0;  // host seconds inside them
    std::int64_t segment_releases = 0;      // segments returned, all time
    std::int64_t out_of_memory_releases = 0;  // releases forced by a failed source allocation
    std::int64_t live_blocks = 0;
};

class BlockCache {
public:
    static constexpr std::size_t kAlign = 512;
    static constexpr std::size_t kSmallLimit = std::size_t{1} << 20;
    static constexpr std::size_t kSmallSegment = std::size_t{32} << 20;
    static constexpr std::size_t kLargeSegment = std::size_t{1} << 30;
    static constexpr std::size_t kLargeRound = std::size_t{2} << 20;
    static constexpr std::size_t kLargeMinSplit = std::size_t{1} << 20;

    explicit BlockCache(SegmentSource source) : source_(std::move(source)) {}
    BlockCache(const BlockCache&) = delete;
    BlockCache& operator=(const BlockCache&) = delete;
    ~BlockCache() {
        for (auto& kv : blocks_) delete kv.second;
    }

    static std::size_t rounded(std::size_t bytes) {
        if (bytes == 0) bytes = 1;
        return (bytes + kAlign - 1) / kAlign * kAlign;
    }

    void rebind_owner() { owner_ = std::this_thread::get_id(); }

    void* allocate(std::size_t bytes) {
        check_owner();
        const std::size_t size = rounded(bytes);
        const bool small = size <= kSmallLimit;
        auto& pool = small ? free_small_ : free_large_;
        auto it = best_fit(pool, size);
        if (it == pool.end()) {
            grow(size, small);
            it = best_fit(pool, size);
            if (it == pool.end()) throw std::logic_error("block cache: new segment does not fit the request");
        }
        Block* b = *it;
        pool.erase(it);
        const std::size_t rest = b->size - size;
        if (rest >= (small ? kAlign : kLargeMinSplit)) {
            auto* r = new Block{b->ptr + size, rest, b, b->next, b->segment, true, small};
            if (b->next) b->next->prev = r;
            b->next = r;
            b->size = size;
            blocks_.emplace(r->ptr, r);
            pool.insert(r);
        }
        b->free = false;
        stats_.live_bytes += b->size;
        stats_.live_blocks += 1;
        if (stats_.live_bytes > stats_.peak_live_bytes) stats_.peak_live_bytes = stats_.live_bytes;
        return b->ptr;
    }

    // Returns the block to the cache; the caller's stream order makes it safe for
    // the next allocation on that stream.
    void free(void* p) {
        check_owner();
        auto found = blocks_.find(static_cast<char*>(p));
        if (found == blocks_.end() || found->second->free)
            throw std::invalid_argument("block cache: free of an unknown device block");
        Block* b = found->second;
        b->free = true;
        stats_.live_bytes -= b->size;
        stats_.live_blocks -= 1;
        auto& pool = b->small ? free_small_ : free_large_;
        if (b->next && b->next->free) {
            Block* n = b->next;
            pool.erase(n);
            b->size += n->size;
            b->next = n->next;
            if (n->next) n->next->prev = b;
            blocks_.erase(n->ptr);
            delete n;
        }
        if (b->prev && b->prev->free) {
            Block* q = b->prev;
            pool.erase(q);
            q->size += b->size;
            q->next = b->next;
            if (b->next) b->next->prev = q;
            blocks_.erase(b->ptr);
            delete b;
            b = q;
        }
        pool.insert(b);
    }

    // Drains the stream and returns every segment without live blocks to the
    // source. Returns the number of segments released.
    std::int64_t release_free_segments() {
        check_owner();
        source_.drain();
        std::int64_t released = 0;
        for (auto* pool : {&free_small_, &free_large_}) {
            for (auto it = pool->begin(); it != pool->end();) {
                Block* b = *it;
                if (b->prev == nullptr && b->next == nullptr) {
                    it = pool->erase(it);
                    auto seg = segments_.find(b->segment);
                    stats_.reserved_bytes -= seg->second;
                    segments_.erase(seg);
                    blocks_.erase(b->ptr);
                    source_.release(b->ptr);
                    delete b;
                    ++released;
                } else {
                    ++it;
                }
            }
        }
        stats_.segments = static_cast<std::int64_t>(segments_.size());
        stats_.segment_releases += released;
        return released;
    }

    BlockCacheStats stats() const { return stats_; }
    std::int64_t free_blocks() const { return static_cast<std::int64_t>(free_small_.size() + free_large_.size()); }
    std::uint64_t free_bytes() const {
        std::uint64_t total = 0;
        for (const Block* b : free_small_) total += b->size;
        for (const Block* b : free_large_) total += b->size;
        return total;
    }
    // Segment bases and sizes, for checks.
    const std::map<char*, std::size_t>& segments() const { return segments_; }
    // Size of a live block, 0 if p is not one.
    std::size_t block_size(const void* p) const {
        auto it = blocks_.find(static_cast<char*>(const_cast<void*>(p)));
        return it == blocks_.end() || it->second->free ? 0 : it->second->size;
    }

private:
    void check_owner() {
        const auto me = std::this_thread::get_id();
        if (owner_ == std::thread::id()) owner_ = me;
        else if (owner_ != me) throw std::logic_error("block cache: used by a thread that does not own its context");
    }
    std::thread::id owner_{};

    struct Block {
        char* ptr;
        std::size_t size;
        Block* prev;  // address neighbours within the segment
        Block* next;
        char* segment;
        bool free;
        bool small;
    };
    struct BySize {
        bool operator()(const Block* a, const Block* b) const {
            return a->size != b->size ? a->size < b->size : a->ptr < b->ptr;
        }
    };
    using FreeSet = std::set<Block*, BySize>;

    static FreeSet::iterator best_fit(FreeSet& pool, std::size_t size) {
        Block probe{nullptr, size, nullptr, nullptr, nullptr, true, false};
        return pool.lower_bound(&probe);
    }

    void* source_allocate(std::size_t bytes) {
        const auto t0 = std::chrono::steady_clock::now();
        void* p = source_.allocate(bytes);
        stats_.segment_allocation_seconds +=
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (p) stats_.segment_allocations += 1;
        return p;
    }

    void grow(std::size_t size, bool small) {
        const std::size_t exact = small ? kSmallSegment : (size + kLargeRound - 1) / kLargeRound * kLargeRound;
        const std::size_t want = small ? kSmallSegment : (exact > kLargeSegment ? exact : kLargeSegment);
        std::size_t got = want;
        void* p = source_allocate(want);
        if (!p) {
            // Out of device memory: hand back every free segment and retry, last at
            // the request's own rounded size.
            stats_.out_of_memory_releases += 1;
            release_free_segments();
            p = source_allocate(want);
            if (!p && exact < want) {
                got = exact;
                p = source_allocate(exact);
            }
            if (!p)
                throw std::runtime_error("device out of memory: request of " + std::to_string(size) + " bytes with " +
                                         std::to_string(stats_.live_bytes) + " bytes live and " +
                                         std::to_string(stats_.reserved_bytes) + " bytes reserved");
        }
        char* base = static_cast<char*>(p);
        segments_.emplace(base, got);
        auto* b = new Block{base, got, nullptr, nullptr, base, true, small};
        blocks_.emplace(base, b);
        (small ? free_small_ : free_large_).insert(b);
        stats_.reserved_bytes += got;
        stats_.segments = static_cast<std::int64_t>(segments_.size());
        if (stats_.reserved_bytes > stats_.peak_reserved_bytes) stats_.peak_reserved_bytes = stats_.reserved_bytes;
    }

    SegmentSource source_;
    FreeSet free_small_, free_large_;
    std::unordered_map<char*, Block*> blocks_;  // every block, free or live, by address
    std::map<char*, std::size_t> segments_;
    BlockCacheStats stats_;
};

}  // namespace gpart_hip
