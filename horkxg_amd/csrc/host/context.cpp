// context.cpp -- the calling thread's kernel-library context.
//
// The reference runs every kernel on the legacy default stream of the current
// CUDA device. Here each host thread owns one gpart_ctx (one stream, one caching
// allocator), created on first use on device 0 or bound explicitly.
#include "device_vector.hpp"

namespace gpart::detail {
namespace {
struct ThreadContext {
    gpart_ctx* ctx = nullptr;
    bool owned = false;
    ~ThreadContext() {
        if (owned && ctx) gpart_ctx_destroy(ctx);
    }
};
thread_local ThreadContext current;
}  // namespace

gpart_ctx* context() {
    if (!current.ctx) {
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
