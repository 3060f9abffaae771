// load.hip -- int64 CSR input narrowed on the device.
//
// The reference loads indptr / indices files into host vectors, narrows them on
// the host (offsets to the width's offset type, neighbours to int32, unit
// weights) and copies the result to the GPU (src/graph.cpp:95-116, 160-195). Here
// the int64 arrays stream to the device in chunks and a kernel narrows each chunk
// in place of the host loop; edge and vertex weights are filled on the device, so
// only the int64 input crosses PCIe. Checks are flags raised by the kernels (the
// caller maps them to the reference's messages); nothing is read back per chunk.
//
// From host memory the chunks are plain pageable copies. From files, two pinned
// staging buffers alternate: the host reads chunk i + 1 from the file while the
// copy and the kernel of chunk i run.
#include "internal.hpp"

#include <algorithm>
#include <cstdio>
#include <limits>
#include <string>

namespace gpart_hip {
namespace {

constexpr int kBlock = 256;
constexpr std::int64_t kChunk = std::int64_t{1} << 23;  // int64 entries per transfer: 64 MiB

inline unsigned grid_for(std::int64_t threads) { return static_cast<unsigned>((threads + kBlock - 1) / kBlock); }

// status bits (gpart_hip.h)
constexpr int kNotMonotone = 1, kOffsetTooLarge = 2, kBadNeighbor = 4, kBadEndpoints = 8;

// Offsets [first, first + count) of n + 1: monotone and non-negative (previous is
// the entry before the chunk, or 0), within O, and 0 / m at the ends.
template <typename O>
__global__ void narrow_offsets_kernel(const std::int64_t* src, std::int64_t count, std::int64_t first,
                                      std::int64_t previous, std::int64_t n, std::int64_t m, O* dst, int* status) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const std::int64_t v = src[i];
    const std::int64_t before = i > 0 ? src[i - 1] : previous;
    int bits = 0;
    if (v < 0 || v < before) bits |= kNotMonotone;
    if (v > static_cast<std::int64_t>(std::numeric_limits<O>::max())) bits |= kOffsetTooLarge;
    const std::int64_t index = first + i;
    if ((index == 0 && v != 0) || (index == n && v != m)) bits |= kBadEndpoints;
    if (bits) atomicOr(status, bits);
    dst[index] = static_cast<O>(v);
}

__global__ void narrow_neighbors_kernel(const std::int64_t* src, std::int64_t count, std::int64_t n,
                                        std::int32_t* dst, int* status) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const std::int64_t u = src[i];
    if (u < 0 || u >= n) atomicOr(status, kBadNeighbor);
    dst[i] = static_cast<std::int32_t>(u);
}

template <typename T>
__global__ void fill_ones_kernel(T* dst, std::int64_t count) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t j = i; j < count; j += stride) dst[j] = T{1};
}

// Source of int64 chunks: host memory or a file.
struct Chunks {
    const std::int64_t* memory = nullptr;
    std::FILE* file = nullptr;
    std::string path;
};

// The device arrays of one CSR, freed unless released to the caller.
template <typename O, typename W>
struct Arrays {
    gpart_ctx* ctx;
    O* offsets = nullptr;
    std::int32_t* neighbors = nullptr;
    W* edge_weights = nullptr;
    W* vertex_weights = nullptr;
    ~Arrays() {
        for (void* p : {static_cast<void*>(offsets), static_cast<void*>(neighbors), static_cast<void*>(edge_weights),
                        static_cast<void*>(vertex_weights)})
            if (p) ctx_free(ctx, p);
    }
};

// Streams `count` int64 values from src through the device stage, calling
// launch(stage, offset, take) after each chunk's copy. Pinned double buffering
// for files; plain copies from host memory.
template <typename Launch>
void stream_int64(gpart_ctx* ctx, const Chunks& src, std::int64_t count, std::int64_t* const stage[2],
                  std::int64_t* const pinned[2], hipEvent_t done[2], Launch launch) {
    for (std::int64_t off = 0, c = 0; off < count; off += kChunk, ++c) {
        const std::int64_t take = std::min(kChunk, count - off);
        const int b = static_cast<int>(c & 1);
        const std::size_t bytes = static_cast<std::size_t>(take) * 8;
        if (src.memory) {
            GPART_HIP_CHECK(hipMemcpyAsync(stage[b], src.memory + off, bytes, hipMemcpyHostToDevice, ctx->stream));
        } else {
            // the pinned buffer is free once the copy issued two chunks ago is done
            GPART_HIP_CHECK(hipEventSynchronize(done[b]));
            if (std::fread(pinned[b], 8, static_cast<std::size_t>(take), src.file) != static_cast<std::size_t>(take))
                throw Error(GPART_ERR_RUNTIME, "cannot read " + src.path);
            GPART_HIP_CHECK(hipMemcpyAsync(stage[b], pinned[b], bytes, hipMemcpyHostToDevice, ctx->stream));
            GPART_HIP_CHECK(hipEventRecord(done[b], ctx->stream));
        }
        launch(stage[b], off, take);
        GPART_LAUNCH_CHECK();
    }
}

template <typename O, typename W>
void csr_from_int64(gpart_ctx* ctx, std::int64_t n, std::int64_t m, const Chunks& indptr, const Chunks& indices,
                    gpart_device_csr* out, std::int32_t* status) {
    Arrays<O, W> a{ctx};
    a.offsets = static_cast<O*>(ctx_alloc(ctx, sizeof(O) * static_cast<std::size_t>(n + 1)));
    a.neighbors = static_cast<std::int32_t*>(ctx_alloc(ctx, 4 * static_cast<std::size_t>(std::max<std::int64_t>(m, 1))));
    a.edge_weights = static_cast<W*>(ctx_alloc(ctx, sizeof(W) * static_cast<std::size_t>(std::max<std::int64_t>(m, 1))));
    a.vertex_weights = static_cast<W*>(ctx_alloc(ctx, sizeof(W) * static_cast<std::size_t>(std::max<std::int64_t>(n, 1))));
    Scratch<std::int64_t> stage(ctx, 2 * kChunk);
    Scratch<int> flags(ctx, 1);
    fill_bytes(ctx, flags.get(), 0, sizeof(int));
    std::int64_t* const stages[2] = {stage.get(), stage.get() + kChunk};
    const bool files = indptr.file != nullptr;
    std::int64_t* pinned[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        for (int b = 0; b < 2; ++b) {
            if (done[b]) (void)hipEventDestroy(done[b]);
            if (pinned[b]) (void)hipHostFree(pinned[b]);
        }
    };
    try {
        if (files)
            for (int b = 0; b < 2; ++b) {
                GPART_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&pinned[b]), kChunk * 8, hipHostMallocDefault));
                GPART_HIP_CHECK(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
                GPART_HIP_CHECK(hipEventRecord(done[b], ctx->stream));
            }
        // offsets: the entry before each chunk is read back from the host side
        std::int64_t previous = 0;
        stream_int64(ctx, indptr, n + 1, stages, pinned, done,
                     [&](const std::int64_t* s, std::int64_t off, std::int64_t take) {
                         hipLaunchKernelGGL(narrow_offsets_kernel<O>, dim3(grid_for(take)), dim3(kBlock), 0,
                                            ctx->stream, s, take, off, previous, n, m, a.offsets, flags.get());
                         // the chunk's last value, for the next chunk's monotone check
                         if (off + take < n + 1) {
                             if (indptr.memory) previous = indptr.memory[off + take - 1];
                             else previous = pinned[(off / kChunk) & 1][take - 1];
                         }
                     });
        stream_int64(ctx, indices, m, stages, pinned, done,
                     [&](const std::int64_t* s, std::int64_t off, std::int64_t take) {
                         hipLaunchKernelGGL(narrow_neighbors_kernel, dim3(grid_for(take)), dim3(kBlock), 0,
                                            ctx->stream, s, take, n, a.neighbors + off, flags.get());
                     });
        const unsigned fill_blocks = 4096;
        if (m > 0)
            hipLaunchKernelGGL(fill_ones_kernel<W>, dim3(fill_blocks), dim3(kBlock), 0, ctx->stream, a.edge_weights,
                               m);
        if (n > 0)
            hipLaunchKernelGGL(fill_ones_kernel<W>, dim3(fill_blocks), dim3(kBlock), 0, ctx->stream,
                               a.vertex_weights, n);
        GPART_LAUNCH_CHECK();
        *status = read_scalar(ctx, flags.get());  // synchronises: staging is idle afterwards
    } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        cleanup();
        throw;
    }
    cleanup();
    if (*status) return;  // Arrays frees everything
    out->n = n;
    out->m = m;
    out->offsets = a.offsets;
    out->neighbors = a.neighbors;
    out->edge_weights = a.edge_weights;
    out->vertex_weights = a.vertex_weights;
    a.offsets = nullptr;
    a.neighbors = nullptr;
    a.edge_weights = nullptr;
    a.vertex_weights = nullptr;
}

template <typename F>
void by_width(int width, F&& f) {
    if (width == GPART_WIDTH_STANDARD) f(std::int32_t{}, std::uint32_t{});
    else if (width == GPART_WIDTH_BIG) f(std::int64_t{}, std::uint64_t{});
    else throw Error(GPART_ERR_INVALID_ARGUMENT, "unknown graph width");
}

}  // namespace
}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

int gpart_csr_from_int64(gpart_ctx* ctx, int width, int64_t n, const int64_t* indptr, int64_t m,
                         const int64_t* indices, gpart_device_csr* out, int32_t* status) {
    return guarded([&] {
        if (!out || !status || n < 0 || m < 0 || !indptr || (m > 0 && !indices))
            throw Error(GPART_ERR_INVALID_ARGUMENT, "invalid int64 CSR arguments");
        *status = 0;
        Chunks p, i;
        p.memory = indptr;
        i.memory = indices;
        by_width(width, [&](auto o, auto w) { csr_from_int64<decltype(o), decltype(w)>(ctx, n, m, p, i, out, status); });
    });
}

int gpart_csr_from_int64_files(gpart_ctx* ctx, int width, const char* indptr_path, const char* indices_path,
                               int64_t n, int64_t m, gpart_device_csr* out, int32_t* status) {
    return guarded([&] {
        if (!out || !status || n < 0 || m < 0 || !indptr_path || !indices_path)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "invalid int64 CSR arguments");
        *status = 0;
        Chunks p, i;
        p.path = indptr_path;
        i.path = indices_path;
        p.file = std::fopen(indptr_path, "rb");
        if (!p.file) throw Error(GPART_ERR_RUNTIME, "cannot open " + p.path);
        i.file = std::fopen(indices_path, "rb");
        if (!i.file) {
            std::fclose(p.file);
            throw Error(GPART_ERR_RUNTIME, "cannot open " + i.path);
        }
        try {
            by_width(width,
                     [&](auto o, auto w) { csr_from_int64<decltype(o), decltype(w)>(ctx, n, m, p, i, out, status); });
        } catch (...) {
            std::fclose(p.file);
            std::fclose(i.file);
            throw;
        }
        std::fclose(p.file);
        std::fclose(i.file);
    });
}

}  // extern "C"
