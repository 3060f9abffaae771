// device_vector.hpp -- owning device array over the kernel library's allocator.
//
// Stands where the reference uses thrust::device_vector (include/graph.hpp:37-48,
// include/hierarchy.hpp:30-38). Storage comes from the thread's gpart_ctx, whose
// stream-ordered caching allocator replaces the reference's ScratchPoolSession
// (include/detail/scratch_pool.hpp:15-106).
#pragma once

#include "check.hpp"
#include "gpart_hip.h"

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <utility>
#include <vector>

namespace gpart {

namespace detail {
// The calling thread's context (device 0 unless bound otherwise).
gpart_ctx* context();
// Bind / unbind an externally created context for the calling thread.
void bind_context(gpart_ctx* ctx);
// Synchronise and return all cached device blocks (ScratchPoolSession::finish).
void release_context_cache();
}  // namespace detail

template <typename T>
class DeviceVector {
public:
    using value_type = T;

    DeviceVector() = default;
    explicit DeviceVector(std::size_t n) { allocate(n); }
    DeviceVector(std::size_t n, const T& value) {
        allocate(n);
        fill(value);
    }
    DeviceVector(const std::vector<T>& host) { assign(host); }
    DeviceVector(const DeviceVector& other) { copy_from(other); }
    DeviceVector(DeviceVector&& other) noexcept
        : ctx_(other.ctx_), data_(other.data_), borrowed_(other.borrowed_), size_(other.size_) {
        other.ctx_ = nullptr;
        other.data_ = nullptr;
        other.size_ = 0;
        other.borrowed_ = false;
    }
    DeviceVector& operator=(const DeviceVector& other) {
        if (this != &other) copy_from(other);
        return *this;
    }
    DeviceVector& operator=(DeviceVector&& other) noexcept {
        if (this != &other) {
            release();
            ctx_ = other.ctx_;
            data_ = other.data_;
            size_ = other.size_;
            borrowed_ = other.borrowed_;
            other.ctx_ = nullptr;
            other.data_ = nullptr;
            other.size_ = 0;
            other.borrowed_ = false;
        }
        return *this;
    }
    DeviceVector& operator=(const std::vector<T>& host) {
        assign(host);
        return *this;
    }
    ~DeviceVector() { release(); }

    // Take ownership of a block allocated from the thread's context.
    static DeviceVector adopt(T* pointer, std::size_t n) {
        DeviceVector v;
        v.ctx_ = detail::context();
        v.data_ = pointer;
        v.size_ = n;
        return v;
    }

    // A read-only view of another vector's memory, never freed by this object; the
    // owner must outlive it. Copies of a view own their memory.
    static DeviceVector borrow(const DeviceVector& owner) {
        DeviceVector v;
        v.ctx_ = owner.ctx_;
        v.data_ = const_cast<T*>(owner.data_);
        v.size_ = owner.size_;
        v.borrowed_ = true;
        return v;
    }

    std::size_t size() const { return size_; }
    bool empty() const { return size_ == 0; }
    T* data() { return data_; }
    const T* data() const { return data_; }

    // Resize without preserving contents (device arrays here are rebuilt, not grown).
    void resize_discard(std::size_t n) {
        if (n == size_ && !borrowed_) return;  // a view is replaced, never written
        release();
        allocate(n);
    }
    void clear() { release(); }

    void assign(const std::vector<T>& host) {
        resize_discard(host.size());
        if (!host.empty())
            GPART_CHECK(gpart_memcpy_htod(detail::context(), data_, host.data(), host.size() * sizeof(T)));
    }
    void fill(const T& value);
    std::vector<T> to_host() const {
        std::vector<T> host(size_);
        if (size_) GPART_CHECK(gpart_memcpy_dtoh(detail::context(), host.data(), data_, size_ * sizeof(T)));
        return host;
    }
    T element(std::size_t i) const {
        T value{};
        GPART_CHECK(gpart_memcpy_dtoh(detail::context(), &value, data_ + i, sizeof(T)));
        return value;
    }
    void swap(DeviceVector& other) noexcept {
        std::swap(ctx_, other.ctx_);
        std::swap(data_, other.data_);
        std::swap(size_, other.size_);
        std::swap(borrowed_, other.borrowed_);
    }

private:
    void allocate(std::size_t n) {
        size_ = n;
        if (!n) return;
        ctx_ = detail::context();
        void* p = nullptr;
        GPART_CHECK(gpart_device_alloc(ctx_, n * sizeof(T), &p));
        data_ = static_cast<T*>(p);
    }
    void release() {
        if (data_ && !borrowed_) gpart_device_free(ctx_, data_);
        data_ = nullptr;
        size_ = 0;
        borrowed_ = false;
    }
    void copy_from(const DeviceVector& other) {
        resize_discard(other.size_);
        if (size_)
            GPART_CHECK(gpart_memcpy_dtod(detail::context(), data_, other.data_, size_ * sizeof(T)));
    }

    gpart_ctx* ctx_ = nullptr;
    T* data_ = nullptr;
    bool borrowed_ = false;
    std::size_t size_ = 0;
};

template <typename T>
void DeviceVector<T>::fill(const T& value) {
    if (!size_) return;
    if (borrowed_) throw std::logic_error("fill of a borrowed device view");
    bool uniform_bytes = true;
    const auto* bytes = reinterpret_cast<const unsigned char*>(&value);
    for (std::size_t i = 1; i < sizeof(T); ++i) uniform_bytes = uniform_bytes && bytes[i] == bytes[0];
    if (uniform_bytes) {
        GPART_CHECK(gpart_memset(detail::context(), data_, bytes[0], size_ * sizeof(T)));
    } else if (sizeof(T) == 4 || sizeof(T) == 8) {
        std::uint64_t pattern = 0;
        std::memcpy(&pattern, &value, sizeof(T));
        GPART_CHECK(gpart_fill_pattern(detail::context(), data_, pattern, static_cast<int>(sizeof(T)),
                                       static_cast<std::int64_t>(size_)));
    } else {
        std::vector<T> host(size_, value);
        GPART_CHECK(gpart_memcpy_htod(detail::context(), data_, host.data(), size_ * sizeof(T)));
    }
}

}  // namespace gpart
