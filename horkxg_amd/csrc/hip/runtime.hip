// runtime.hip -- context, caching allocator, copies, fills and event timing.
//
// Replaces the reference's CUDA_CHECK / ScratchPoolSession / GpuEventTimer
// (include/check.hpp:8-18, include/detail/scratch_pool.hpp:15-106,
// src/coarsen.cu:86-110) with a per-context stream and a stream-ordered caching
// allocator: all work of a context runs on its one stream, so a freed block may
// be handed to the next allocation without a device synchronisation.
#include "internal.hpp"

#include <chrono>
#include <cstring>
#include <sstream>

namespace gpart_hip {
namespace {
thread_local std::string last_error_message;

std::size_t round_size(std::size_t bytes) {
    if (bytes <= 256) return 256;
    // 1/8 power-of-two steps bound internal waste to 12.5%.
    std::size_t p = 256;
    while (p < bytes) p <<= 1;
    const std::size_t step = p / 16;
    return (bytes + step - 1) / step * step;
}

__global__ void fill_u32_kernel(std::uint32_t* dst, std::uint32_t value, std::int64_t words) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t j = i; j < words; j += stride) dst[j] = value;
}

__global__ void fill_u64_kernel(std::uint64_t* dst, std::uint64_t value, std::int64_t words) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t j = i; j < words; j += stride) dst[j] = value;
}

// 16-byte vector copies when both ends are 16-byte aligned, bytes otherwise.
__global__ void copy_vec_kernel(uint4* dst, const uint4* src, std::int64_t vectors) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t j = i; j < vectors; j += stride) dst[j] = src[j];
}

__global__ void copy_bytes_kernel(std::uint8_t* dst, const std::uint8_t* src, std::int64_t bytes) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t j = i; j < bytes; j += stride) dst[j] = src[j];
}

__global__ void fill_u8_kernel(std::uint8_t* dst, std::uint8_t value, std::int64_t bytes) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < bytes) dst[i] = value;
}

// Copies bytes into status words 1.. (host-visible), then stores seq in word 0.
__global__ void __launch_bounds__(kWave) status_copy_kernel(const std::uint8_t* src, std::int64_t bytes,
                                                            unsigned long long* status, unsigned long long* counter) {
    auto* dst = reinterpret_cast<std::uint8_t*>(status + 1);
    if ((reinterpret_cast<std::uintptr_t>(src) & 7) == 0) {
        const auto* s8 = reinterpret_cast<const unsigned long long*>(src);
        const std::int64_t words = bytes / 8;
        for (std::int64_t i = threadIdx.x; i < words; i += kWave) status[1 + i] = s8[i];
        for (std::int64_t i = words * 8 + threadIdx.x; i < bytes; i += kWave) dst[i] = src[i];
    } else {
        for (std::int64_t i = threadIdx.x; i < bytes; i += kWave) dst[i] = src[i];
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long seq = *counter + 1;
        *counter = seq;
        __hip_atomic_store(status, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ void sequence_kernel(std::int32_t* out, std::int64_t n) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<std::int32_t>(i);
}

template <typename W>
__global__ void widen_kernel(const W* in, std::uint64_t* out, std::int64_t n) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<std::uint64_t>(in[i]);
}
}  // namespace

void set_last_error(const std::string& message) { last_error_message = message; }

void throw_hip(hipError_t error, const char* file, int line, const char* call) {
    std::ostringstream s;
    s << "HIP failure at " << file << ':' << line << " for " << call << ": "
      << hipGetErrorString(error);
    throw Error(GPART_ERR_RUNTIME, s.str());
}

// Returns every cached segment without live blocks to the device: waits for the
// stream first, since freed blocks may still be read by queued work.
void release_blocks(gpart_ctx* ctx) {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->round_plans.reset();  // their blocks become free and releasable
    const auto t1 = clock::now();
    ctx->released_blocks = ctx->cache->release_free_segments();
    if (ctx->plan_cache) ctx->released_blocks += ctx->plan_cache->release_free_segments();
    const auto t2 = clock::now();
    ctx->release_seconds[0] = 0.0;
    ctx->release_seconds[1] = std::chrono::duration<double>(t1 - t0).count();
    ctx->release_seconds[2] = std::chrono::duration<double>(t2 - t1).count();
}

void* ctx_alloc(gpart_ctx* ctx, std::size_t bytes) {
    if (ctx->recorder) return ctx->recorder->take(ctx, bytes);
    try {
        return ctx->cache->allocate(bytes);
    } catch (const std::runtime_error& e) {
        if (dynamic_cast<const Error*>(&e)) throw;
        throw Error(GPART_ERR_RUNTIME, e.what());
    }
}

void ctx_free(gpart_ctx* ctx, void* p) {
    if (!p) return;
    if (ctx->recorder && ctx->recorder->owns(p)) return;
    try {
        ctx->cache->free(p);
    } catch (const std::invalid_argument& e) {
        throw Error(GPART_ERR_INVALID_ARGUMENT, e.what());
    }
}

void wait_status(gpart_ctx* ctx, std::uint64_t seq) {
    using clock = std::chrono::steady_clock;
    const auto start = clock::now();
    for (unsigned spin = 0;; ++spin) {
        if (__atomic_load_n(ctx->status_host, __ATOMIC_ACQUIRE) == seq) return;
        // poll for up to 20 ms (rounds of the large levels run for hundreds of us; a
        // blocking wait on the stream wakes up later than the poll sees the word)
        if ((spin & 255) == 255 && clock::now() - start > std::chrono::milliseconds(20)) break;
        __builtin_ia32_pause();
    }
    GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (__atomic_load_n(ctx->status_host, __ATOMIC_ACQUIRE) != seq)
        throw Error(GPART_ERR_RUNTIME, "round status did not reach the host");
}

// Small read-backs: a one-wave kernel copies the bytes into the fine-grained
// status words and stores the sequence number after them; the host polls for it
// (no copy engine, no blocking wait).
void read_back(gpart_ctx* ctx, void* host, const void* device, std::size_t bytes) {
    if (bytes == 0) return;
    if (bytes <= static_cast<std::size_t>(kStatusWords - 1) * 8) {
        const std::uint64_t seq = ++ctx->status_seq;
        hipLaunchKernelGGL(status_copy_kernel, dim3(1), dim3(kWave), 0, ctx->stream,
                           static_cast<const std::uint8_t*>(device), static_cast<std::int64_t>(bytes),
                           reinterpret_cast<unsigned long long*>(ctx->status_device), ctx->status_counter);
        GPART_LAUNCH_CHECK();
        wait_status(ctx, seq);
        std::memcpy(host, ctx->status_host + 1, bytes);
    } else {
        GPART_HIP_CHECK(hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, ctx->stream));
        GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
}

// Device-to-device copy as a kernel on the context stream: the runtime's own
// device-to-device copies of multi-GB arrays measured ~8 GB/s here.
void copy_device(gpart_ctx* ctx, void* dst, const void* src, std::size_t bytes) {
    if (bytes == 0) return;
    const auto d = reinterpret_cast<std::uintptr_t>(dst), s = reinterpret_cast<std::uintptr_t>(src);
    const std::size_t head = (d % 16 == 0 && s % 16 == 0) ? bytes / 16 * 16 : 0;
    if (head) {
        const auto vectors = static_cast<std::int64_t>(head / 16);
        const int blocks = static_cast<int>(std::min<std::int64_t>((vectors + 255) / 256, 16384));
        hipLaunchKernelGGL(copy_vec_kernel, dim3(blocks), dim3(256), 0, ctx->stream, static_cast<uint4*>(dst),
                           static_cast<const uint4*>(src), vectors);
    }
    if (bytes > head) {
        const auto rest = static_cast<std::int64_t>(bytes - head);
        const int blocks = static_cast<int>(std::min<std::int64_t>((rest + 255) / 256, 16384));
        hipLaunchKernelGGL(copy_bytes_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<std::uint8_t*>(dst) + head, static_cast<const std::uint8_t*>(src) + head,
                           rest);
    }
    GPART_LAUNCH_CHECK();
}

void fill_bytes(gpart_ctx* ctx, void* dst, int value, std::size_t bytes) {
    if (bytes == 0) return;
    const auto byte = static_cast<std::uint8_t>(value);
    if (reinterpret_cast<std::uintptr_t>(dst) % 4 == 0 && bytes % 4 == 0) {
        const std::uint32_t word = 0x01010101U * byte;
        const auto words = static_cast<std::int64_t>(bytes / 4);
        const int blocks = static_cast<int>(std::min<std::int64_t>((words + 255) / 256, 8192));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<std::uint32_t*>(dst), word, words);
    } else {
        const auto n = static_cast<std::int64_t>(bytes);
        hipLaunchKernelGGL(fill_u8_kernel, dim3(ceil_div_int(n, 256)), dim3(256), 0, ctx->stream,
                           static_cast<std::uint8_t*>(dst), byte, n);
    }
    GPART_LAUNCH_CHECK();
}

void profile_begin(gpart_ctx* ctx, int family) {
    auto& p = ctx->profile[family];
    if (!p.enabled) return;
    hipEvent_t a, b;
    if (ctx->spare_events.size() >= 2) {
        a = ctx->spare_events.back();
        ctx->spare_events.pop_back();
        b = ctx->spare_events.back();
        ctx->spare_events.pop_back();
    } else {
        GPART_HIP_CHECK(hipEventCreate(&a));
        GPART_HIP_CHECK(hipEventCreate(&b));
    }
    GPART_HIP_CHECK(hipEventRecord(a, ctx->stream));
    p.events.emplace_back(a, b);
}

void profile_end(gpart_ctx* ctx, int family, double algorithmic_bytes) {
    auto& p = ctx->profile[family];
    if (!p.enabled) return;
    GPART_HIP_CHECK(hipEventRecord(p.events.back().second, ctx->stream));
    p.bytes += algorithmic_bytes;
    p.launches += 1;
}

void profile_add_bytes(gpart_ctx* ctx, int family, double algorithmic_bytes) {
    auto& p = ctx->profile[family];
    if (p.enabled) p.bytes += algorithmic_bytes;
}

void profile_add_units(gpart_ctx* ctx, int family, double units) {
    auto& p = ctx->profile[family];
    if (p.enabled) p.units += units;
}

}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

const char* gpart_last_error(void) { return gpart_hip::last_error_message.c_str(); }

void gpart_hip_set_error(const char* message) { gpart_hip::set_last_error(message ? message : ""); }

int gpart_device_count(int* count) {
    return guarded([&] { GPART_HIP_CHECK(hipGetDeviceCount(count)); });
}

int gpart_ctx_create(int device, gpart_ctx** out) {
    return guarded([&] {
        auto ctx = new gpart_ctx;
        try {
            ctx->device = device;
            GPART_HIP_CHECK(hipSetDevice(device));
            // Default priority: no stream priorities anywhere in this library.
            GPART_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            cut_scratch_setup(device);
            GPART_HIP_CHECK(hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault));
            GPART_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ctx->status_host), kStatusWords * 8,
                                          hipHostMallocCoherent | hipHostMallocMapped));
            GPART_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->status_device), ctx->status_host, 0));
            std::memset(ctx->status_host, 0, kStatusWords * 8);
            GPART_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ctx->status_counter), sizeof(unsigned long long)));
            GPART_HIP_CHECK(hipMemset(ctx->status_counter, 0, sizeof(unsigned long long)));
            SegmentSource source;
            source.allocate = [ctx](std::size_t bytes) -> void* {
                void* p = nullptr;
                const hipError_t e = hipMalloc(&p, bytes);
                if (e == hipErrorOutOfMemory) {
                    (void)hipGetLastError();
                    return nullptr;
                }
                GPART_HIP_CHECK(e);
                return p;
            };
            source.release = [](void* p) { GPART_HIP_CHECK(hipFree(p)); };
            source.drain = [ctx] { GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream)); };
            ctx->cache = std::make_unique<BlockCache>(std::move(source));
        } catch (...) {
            if (ctx->pinned) (void)hipHostFree(ctx->pinned);
            if (ctx->status_host) (void)hipHostFree(ctx->status_host);
            if (ctx->status_counter) (void)hipFree(ctx->status_counter);
            if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
            delete ctx;
            throw;
        }
        *out = ctx;
    });
}

int gpart_ctx_destroy(gpart_ctx* ctx) {
    if (!ctx) return GPART_OK;
    return guarded([&] {
        GPART_HIP_CHECK(hipSetDevice(ctx->device));
        GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        // teardown may run on another thread (a thread that ended leaves its
        // context to the next one that creates a context)
        ctx->cache->rebind_owner();
        if (ctx->plan_cache) ctx->plan_cache->rebind_owner();
        ctx->round_plans.reset();
        // Every segment goes, including those of blocks the caller never freed.
        for (const auto& kv : ctx->cache->segments()) GPART_HIP_CHECK(hipFree(kv.first));
        ctx->cache.reset();
        if (ctx->plan_cache) {
            for (const auto& kv : ctx->plan_cache->segments()) GPART_HIP_CHECK(hipFree(kv.first));
            ctx->plan_cache.reset();
        }
        for (auto& p : ctx->profile)
            for (auto& e : p.events) {
                GPART_HIP_CHECK(hipEventDestroy(e.first));
                GPART_HIP_CHECK(hipEventDestroy(e.second));
            }
        for (auto e : ctx->spare_events) GPART_HIP_CHECK(hipEventDestroy(e));
        if (ctx->pinned) GPART_HIP_CHECK(hipHostFree(ctx->pinned));
        if (ctx->status_host) GPART_HIP_CHECK(hipHostFree(ctx->status_host));
        if (ctx->status_counter) GPART_HIP_CHECK(hipFree(ctx->status_counter));
        GPART_HIP_CHECK(hipStreamDestroy(ctx->stream));
        delete ctx;
    });
}

int gpart_cache_stats(gpart_ctx* ctx, int64_t* cached_blocks, uint64_t* cached_bytes, uint64_t* live_bytes,
                      int64_t* released_blocks, double* release_seconds) {
    return guarded([&] {
        const auto st = ctx->cache->stats();
        *cached_blocks = ctx->cache->free_blocks();
        *cached_bytes = ctx->cache->free_bytes();
        *live_bytes = st.live_bytes;
        *released_blocks = ctx->released_blocks;
        for (int i = 0; i < 3; ++i) release_seconds[i] = ctx->release_seconds[i];
    });
}

int gpart_cache_info(gpart_ctx* ctx, gpart_cache_info_t* out) {
    return guarded([&] {
        if (!out) throw Error(GPART_ERR_INVALID_ARGUMENT, "null cache info");
        const auto st = ctx->cache->stats();
        out->reserved_bytes = st.reserved_bytes;
        out->live_bytes = st.live_bytes;
        out->free_bytes = ctx->cache->free_bytes();
        out->peak_reserved_bytes = st.peak_reserved_bytes;
        out->peak_live_bytes = st.peak_live_bytes;
        if (ctx->plan_cache) {  // the round graphs' scratch (their peaks add as an upper bound)
            const auto ps = ctx->plan_cache->stats();
            out->reserved_bytes += ps.reserved_bytes;
            out->live_bytes += ps.live_bytes;
            out->free_bytes += ctx->plan_cache->free_bytes();
            out->peak_reserved_bytes += ps.peak_reserved_bytes;
            out->peak_live_bytes += ps.peak_live_bytes;
        }
        out->segments = st.segments;
        out->live_blocks = st.live_blocks;
        out->free_blocks = ctx->cache->free_blocks();
        out->segment_allocations = st.segment_allocations;
        out->segment_allocation_seconds = st.segment_allocation_seconds;
        out->segment_releases = st.segment_releases;
        out->out_of_memory_releases = st.out_of_memory_releases;
    });
}

// Waits for the context's stream by a status word written after the queued work
// (polled), falling back to the blocking stream wait, which also reports a failed
// kernel, when the word does not come.
int gpart_ctx_synchronize(gpart_ctx* ctx) {
    return guarded([&] {
        const std::uint64_t seq = ++ctx->status_seq;
        hipLaunchKernelGGL(status_copy_kernel, dim3(1), dim3(kWave), 0, ctx->stream,
                           static_cast<const std::uint8_t*>(nullptr), std::int64_t{0},
                           reinterpret_cast<unsigned long long*>(ctx->status_device), ctx->status_counter);
        GPART_LAUNCH_CHECK();
        wait_status(ctx, seq);
    });
}

void* gpart_ctx_stream(gpart_ctx* ctx) { return ctx->stream; }

int gpart_device_alloc(gpart_ctx* ctx, size_t bytes, void** out) {
    return guarded([&] { *out = bytes ? ctx_alloc(ctx, bytes) : nullptr; });
}

int gpart_device_free(gpart_ctx* ctx, void* ptr) {
    return guarded([&] { ctx_free(ctx, ptr); });
}

int gpart_release_cache(gpart_ctx* ctx) {
    return guarded([&] {
        release_blocks(ctx);
    });
}

int gpart_memcpy_htod(gpart_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (!bytes) return;
        GPART_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int gpart_memcpy_dtoh(gpart_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (!bytes) return;
        read_back(ctx, dst, src, bytes);
    });
}

int gpart_memcpy_dtod(gpart_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        if (!bytes) return;
        copy_device(ctx, dst, src, bytes);
    });
}

int gpart_memset(gpart_ctx* ctx, void* dst, int value, size_t bytes) {
    return guarded([&] { fill_bytes(ctx, dst, value, bytes); });
}

int gpart_fill_pattern(gpart_ctx* ctx, void* dst, uint64_t pattern, int pattern_bytes, int64_t count) {
    return guarded([&] {
        if (count <= 0) return;
        const int blocks = static_cast<int>(std::min<std::int64_t>((count + 255) / 256, 8192));
        if (pattern_bytes == 8) {
            hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                               static_cast<std::uint64_t*>(dst), pattern, count);
        } else if (pattern_bytes == 4) {
            hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                               static_cast<std::uint32_t*>(dst), static_cast<std::uint32_t>(pattern), count);
        } else {
            throw Error(GPART_ERR_INVALID_ARGUMENT, "fill pattern must be 4 or 8 bytes");
        }
        GPART_LAUNCH_CHECK();
    });
}

int gpart_sequence_i32(gpart_ctx* ctx, int32_t* out, int64_t n) {
    return guarded([&] {
        if (n <= 0) return;
        hipLaunchKernelGGL(sequence_kernel, dim3(ceil_div_int(n, 256)), dim3(256), 0, ctx->stream, out, n);
        GPART_LAUNCH_CHECK();
    });
}

int gpart_widen_weights(gpart_ctx* ctx, const void* weights, int width, int64_t n, uint64_t* out) {
    return guarded([&] {
        if (n <= 0) return;
        if (width == GPART_WIDTH_STANDARD)
            hipLaunchKernelGGL(widen_kernel<std::uint32_t>, dim3(ceil_div_int(n, 256)), dim3(256), 0,
                               ctx->stream, static_cast<const std::uint32_t*>(weights), out, n);
        else
            hipLaunchKernelGGL(widen_kernel<std::uint64_t>, dim3(ceil_div_int(n, 256)), dim3(256), 0,
                               ctx->stream, static_cast<const std::uint64_t*>(weights), out, n);
        GPART_LAUNCH_CHECK();
    });
}

int gpart_profile_enable(gpart_ctx* ctx, int family, int enable) {
    return guarded([&] {
        if (family < 0 || family >= GPART_PROFILE_FAMILIES)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "unknown profile family");
        ctx->profile[family].enabled = enable != 0;
    });
}

int gpart_profile_read(gpart_ctx* ctx, int family, double* total_ms, int64_t* launches,
                       double* bytes) {
    return guarded([&] {
        if (family < 0 || family >= GPART_PROFILE_FAMILIES)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "unknown profile family");
        GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        auto& p = ctx->profile[family];
        double ms = 0.0;
        for (auto& e : p.events) {
            float t = 0.0f;
            GPART_HIP_CHECK(hipEventElapsedTime(&t, e.first, e.second));
            ms += t;
        }
        *total_ms = ms;
        *launches = p.launches;
        *bytes = p.bytes;
    });
}

int gpart_profile_read_units(gpart_ctx* ctx, int family, double* units) {
    return guarded([&] {
        if (family < 0 || family >= GPART_PROFILE_FAMILIES)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "unknown profile family");
        *units = ctx->profile[family].units;
    });
}

int gpart_profile_reset(gpart_ctx* ctx) {
    return guarded([&] {
        GPART_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (auto& p : ctx->profile) {
            for (auto& e : p.events) {
                ctx->spare_events.push_back(e.first);
                ctx->spare_events.push_back(e.second);
            }
            p.events.clear();
            p.bytes = 0.0;
            p.units = 0.0;
            p.launches = 0;
        }
    });
}

}  // extern "C"
