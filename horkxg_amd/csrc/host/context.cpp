// context.cpp -- the calling thread's kernel-library context.
//
// The reference runs every kernel on the legacy default stream of the current
// CUDA device. Here each host thread owns one gpart_ctx (one stream, one caching
// allocator), created on first use on device 0 or bound explicitly.
#include "device_vector.hpp"
#include "gpart.h"
#include "log.hpp"

#include <atomic>
#include <iostream>
#include <mutex>
#include <streambuf>
#include <thread>
#include <vector>

namespace gpart::detail {
namespace {
// No HIP call runs from a thread_local destructor: at thread exit the order of
// TLS destructors across libraries is unspecified, so the HIP runtime's own
// per-thread state may already be gone. The loading thread's context is left to
// process teardown; a context owned by another thread that ends is handed to
// `orphans` and destroyed by the next thread that creates a context.
const std::thread::id loading_thread = std::this_thread::get_id();
std::mutex orphan_mutex;
std::vector<gpart_ctx*> orphans;

struct ThreadContext {
    gpart_ctx* ctx = nullptr;
    bool owned = false;
    ~ThreadContext() {
        if (owned && ctx && std::this_thread::get_id() != loading_thread) {
            std::lock_guard<std::mutex> lock(orphan_mutex);
            orphans.push_back(ctx);
        }
    }
};
thread_local ThreadContext current;

void destroy_orphans() {
    std::vector<gpart_ctx*> dead;
    {
        std::lock_guard<std::mutex> lock(orphan_mutex);
        dead.swap(orphans);
    }
    for (auto* ctx : dead) gpart_ctx_destroy(ctx);
}

class NullBuffer : public std::streambuf {
protected:
    int overflow(int c) override { return c == traits_type::eof() ? 0 : c; }
    std::streamsize xsputn(const char*, std::streamsize n) override { return n; }
};
std::atomic<bool> log_enabled{true};

// One discarding stream per thread: an ostream's formatting state is not safe to
// share between threads (std::cout alone is synchronised by the standard).
std::ostream& null_stream() {
    thread_local NullBuffer buffer;
    thread_local std::ostream stream(&buffer);
    return stream;
}
}  // namespace

std::ostream& log() { return log_enabled.load(std::memory_order_relaxed) ? std::cout : null_stream(); }

gpart_ctx* context() {
    if (!current.ctx) {
        destroy_orphans();
        GPART_CHECK(gpart_ctx_create(0, &current.ctx));
        current.owned = true;
    }
    return current.ctx;
}

void bind_context(gpart_ctx* ctx) {
    if (current.owned && current.ctx && current.ctx != ctx) gpart_ctx_destroy(current.ctx);
    current.ctx = ctx;
    current.owned = false;
}

void release_context_cache() {
    if (current.ctx) GPART_CHECK(gpart_release_cache(current.ctx));
}

}  // namespace gpart::detail

extern "C" void gpart_set_log(int enabled) { gpart::detail::log_enabled.store(enabled != 0); }
