// internal.hpp -- host-side internals shared by the kernel library's sources.
#pragma once

#include "../../../include/gpart_hip.h"
#include "../tuned_gfx950.hpp"
#include "block_cache.hpp"
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace gpart_hip {
// Scratch of a recorded launch sequence (the refinement's round graphs): while a
// recorder is set on a context, ctx_alloc serves requests from it and ctx_free
// leaves its blocks alone, so a replayed sequence finds its buffers at the
// addresses it was captured with.
struct ScratchRecorder {
    virtual ~ScratchRecorder() = default;
    virtual void* take(gpart_ctx* ctx, std::size_t bytes) = 0;
    virtual bool owns(const void* p) const = 0;
};
}  // namespace gpart_hip

struct gpart_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Best-fit caching of device memory in segments the context keeps until an
    // explicit release (block_cache.hpp); all work of the context runs on its one
    // stream, so a freed block may serve the next allocation without a wait.
    std::unique_ptr<gpart_hip::BlockCache> cache;
    // Host seconds of the last release: waiting for the stream, freeing segments.
    double release_seconds[3] = {0, 0, 0};
    std::int64_t released_blocks = 0;
    // Small pinned staging buffer for host-out scalars.
    void* pinned = nullptr;
    // Fine-grained host memory that a kernel writes a refinement round's status
    // into: word 0 is a sequence number the host waits for by polling, the rest the
    // status. Saves the copy engine and the blocking wait of a read-back per round.
    std::uint64_t* status_host = nullptr;
    std::uint64_t* status_device = nullptr;  // the same memory, device address
    // The status sequence: the kernels that publish a status increment a device
    // word and publish its value (so a captured launch needs no new argument per
    // call); the host counts the same publications in status_seq.
    std::uint64_t status_seq = 0;
    unsigned long long* status_counter = nullptr;  // device word
    gpart_hip::ScratchRecorder* recorder = nullptr;
    // The refinement's recorded round graphs (refine.hip), dropped before the
    // cache releases memory or the context ends; round_graphs switches them.
    std::shared_ptr<void> round_plans;
    bool round_graphs = true;
    // The round graphs' scratch, apart from `cache` so that recording a round does
    // not move the addresses the partition's own buffers get (a later partition of
    // the same graph then finds its rounds recorded).
    std::unique_ptr<gpart_hip::BlockCache> plan_cache;
    // Profiling per family.
    struct Profile {
        bool enabled = false;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
        double bytes = 0.0;
        double units = 0.0;  // atomic operations, for the atomic-bound families
        std::int64_t launches = 0;
    } profile[GPART_PROFILE_FAMILIES];
    std::vector<hipEvent_t> spare_events;
    // Refinement rounds rebuild the affinity cache instead of updating it when the
    // level has at least rebuild_min_entries entries and the round's movers hold
    // more than 1/rebuild_share of them (share <= 0: never).
    std::int64_t rebuild_min_entries = gpart_tuned::kRebuildMinEntries;
    std::int64_t rebuild_share = gpart_tuned::kRebuildShare;
    // Workgroups of the rebalance proposals launch at most (each flushes its
    // bucket table into the global counters once).
    std::int64_t balance_blocks = gpart_tuned::kBalanceBlocks;
    // contraction: levels of at least this many fine vertices stash every entry's
    // coarse neighbour in the count pass (gpart_contract_set_stash)
    std::int64_t contract_stash_vertices = std::int64_t{1} << 24;
    // The rebalance's bucket-counter replicas, kept between rounds: the bucket
    // scans zero every counter they read, so a completed round leaves them zero.
    void* rebalance_counters = nullptr;
    std::size_t rebalance_counter_bytes = 0;
    bool rebalance_counters_zero = false;
};

namespace gpart_hip {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& message) : std::runtime_error(message), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

void set_last_error(const std::string& message);

[[noreturn]] void throw_hip(hipError_t error, const char* file, int line, const char* call);

#define GPART_HIP_CHECK(call)                                                   \
    do {                                                                        \
        const hipError_t error__ = (call);                                      \
        if (error__ != hipSuccess) gpart_hip::throw_hip(error__, __FILE__, __LINE__, #call); \
    } while (0)

#define GPART_LAUNCH_CHECK() GPART_HIP_CHECK(hipGetLastError())

template <typename F>
int guarded(F&& body) {
    try {
        body();
        return GPART_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code();
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return GPART_ERR_INVALID_ARGUMENT;
    } catch (const std::overflow_error& e) {
        set_last_error(e.what());
        return GPART_ERR_OVERFLOW;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return GPART_ERR_RUNTIME;
    }
}

// Resolves the calling device's copy of gpart_cut_noalloc's scratch array, once
// per device; the device must be current (refine.hip).
void cut_scratch_setup(int device);

void* ctx_alloc(gpart_ctx* ctx, std::size_t bytes);
void ctx_free(gpart_ctx* ctx, void* p);

// RAII scratch buffer from the context allocator.
template <typename T>
class Scratch {
public:
    Scratch() = default;
    Scratch(gpart_ctx* ctx, std::int64_t count) { reset(ctx, count); }
    ~Scratch() { release(); }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    void reset(gpart_ctx* ctx, std::int64_t count) {
        release();
        ctx_ = ctx;
        count_ = count;
        ptr_ = count > 0 ? static_cast<T*>(ctx_alloc(ctx, sizeof(T) * static_cast<std::size_t>(count)))
                         : nullptr;
    }
    void release() {
        if (ptr_) ctx_free(ctx_, ptr_);
        ptr_ = nullptr;
        count_ = 0;
    }
    T* get() const { return ptr_; }
    std::int64_t size() const { return count_; }
    // Hands the buffer to the caller, who frees it with ctx_free.
    T* take() {
        T* p = ptr_;
        ptr_ = nullptr;
        count_ = 0;
        return p;
    }

private:
    gpart_ctx* ctx_ = nullptr;
    T* ptr_ = nullptr;
    std::int64_t count_ = 0;
};

// Host-out readback through the context's pinned staging buffer (synchronises).
void read_back(gpart_ctx* ctx, void* host, const void* device, std::size_t bytes);
// Waits until status word 0 equals seq: polls for a short while, then synchronises
// the stream (which also reports a failed kernel) and checks again.
void wait_status(gpart_ctx* ctx, std::uint64_t seq);
constexpr int kStatusWords = 512;  // 4 KiB of status memory per context

template <typename T>
T read_scalar(gpart_ctx* ctx, const T* device) {
    T value{};
    read_back(ctx, &value, device, sizeof(T));
    return value;
}

void fill_bytes(gpart_ctx* ctx, void* dst, int value, std::size_t bytes);
// Stream-ordered device-to-device copy by kernel.
void copy_device(gpart_ctx* ctx, void* dst, const void* src, std::size_t bytes);

// Profiling hooks around a family's main kernel launches.
void profile_begin(gpart_ctx* ctx, int family);
void profile_end(gpart_ctx* ctx, int family, double algorithmic_bytes);
// Adds bytes to the family's last launch once its size is known on the host.
void profile_add_bytes(gpart_ctx* ctx, int family, double algorithmic_bytes);
// Adds atomic operations to the family's count.
void profile_add_units(gpart_ctx* ctx, int family, double units);

// ---- primitives (primitives.hip) ----
// Stable LSD radix sort over key bits [begin_bit, end_bit); alternate buffers
// provided by the caller; returns true when the result lies in the alternate
// buffers. count may be read from device memory (d_count) when non-null, with n
// the upper bound used for grid sizing.
template <typename V>
bool radix_sort_pairs(gpart_ctx* ctx, std::uint64_t* keys, std::uint64_t* keys_alt, V* values,
                      V* values_alt, std::int64_t n, int begin_bit, int end_bit);
bool radix_sort_keys(gpart_ctx* ctx, std::uint64_t* keys, std::uint64_t* keys_alt, std::int64_t n,
                     int begin_bit, int end_bit);

// Exclusive / inclusive scans (sum) of n values; out may alias in.
template <typename T>
void exclusive_scan(gpart_ctx* ctx, const T* in, T* out, std::int64_t n, T* total_device = nullptr);
template <typename T>
void inclusive_scan(gpart_ctx* ctx, const T* in, T* out, std::int64_t n);

// Stable compaction of i in [0, n) with flags[i] != 0 into out; count written to
// d_count (device).
void compact_indices(gpart_ctx* ctx, const std::int32_t* flags, std::int64_t n, std::int32_t* out,
                     std::int64_t* d_count);

inline int bit_width_u64(std::uint64_t v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

}  // namespace gpart_hip
