// spectral_device.hpp -- device pieces of the weighted recursive spectral
// bisection shared by the per-depth launches (initial.hip) and the persistent
// task queue (spectral_queue.hip): the reference's FP64 summation trees
// (src/initial_partition_gpu.cu:30-79), the normalisation, the (x, id) key, and
// the sort, weight prefix, feasible split and label assignment of one task
// (:143-153, :230-248).
#pragma once

#include "internal.hpp"

#include <cstdint>

namespace gpart_hip {
namespace {

constexpr int kHalf = 32;
constexpr int kMaxBatchedTask = 4096;
constexpr int kBatchThreads = 1024;

struct Moments {
    double mass, sum, squares;
};

__device__ __forceinline__ double tree32(double v) {
#pragma unroll
    for (int s = 16; s; s >>= 1) v += __shfl_down(v, s, kHalf);
    return v;
}

__device__ __forceinline__ double z(double v) { return v + 0.0; }

// Lane 0 of a 32-lane warp_sum over lanes 0..7 holding b[0..7], others zero.
__device__ __forceinline__ double tree8(const double* b) {
    double c0 = z(z(b[0])) + z(z(b[4])), c1 = z(z(b[1])) + z(z(b[5]));
    double c2 = z(z(b[2])) + z(z(b[6])), c3 = z(z(b[3])) + z(z(b[7]));
    return (c0 + c2) + (c1 + c3);
}

// One 8-vertex block partial: each vertex moment sits on lane 0 of its warp, so
// warp_sum leaves it as z(m); the eight warp totals then use tree8.
__device__ __forceinline__ Moments block_partial(const Moments* vm, int first, int count) {
    double a[8], b[8], c[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        const int i = first + w;
        const Moments m = i < count ? vm[i] : Moments{0, 0, 0};
        a[w] = z(z(z(z(z(m.mass)))));
        b[w] = z(z(z(z(z(m.sum)))));
        c[w] = z(z(z(z(z(m.squares)))));
    }
    return Moments{tree8(a), tree8(b), tree8(c)};
}

struct Norm {
    double mean, scale;
};

__device__ __forceinline__ Norm normalization_of(const Moments& m) {
    const double mean = m.sum / m.mass;
    const double variance = fmax(0.0, fma(-mean, mean, m.squares / m.mass));
    return Norm{mean, variance > 1e-30 ? 1.0 / sqrt(variance) : 1.0};
}

// finish_moments over `count` partials, executed by threads 0..255 of the block
// (other threads only pass the barriers). Returns the normalisation to all.
__device__ Norm finish_block(const Moments* partials, int count, double* shared8, Norm* shared_norm) {
    const int t = threadIdx.x;
    double ms = 0.0, ss = 0.0, qs = 0.0;
    if (t < 256) {
        for (int i = t; i < count; i += 256) {
            ms += partials[i].mass;
            ss += partials[i].sum;
            qs += partials[i].squares;
        }
    }
    // 32-lane trees inside threads 0..255 (half-waves 0..7)
    if (t < 256) {
        ms = tree32(ms);
        ss = tree32(ss);
        qs = tree32(qs);
        if ((t & 31) == 0) {
            shared8[(t >> 5) * 3 + 0] = ms;
            shared8[(t >> 5) * 3 + 1] = ss;
            shared8[(t >> 5) * 3 + 2] = qs;
        }
    }
    __syncthreads();
    if (t == 0) {
        double a[8], b[8], c[8];
        for (int w = 0; w < 8; ++w) {
            a[w] = shared8[w * 3 + 0];
            b[w] = shared8[w * 3 + 1];
            c[w] = shared8[w * 3 + 2];
        }
        *shared_norm = normalization_of(Moments{tree8(a), tree8(b), tree8(c)});
    }
    __syncthreads();
    return *shared_norm;
}

// Algorithmic HBM bytes of one diffusion step over tasks of `vertices` vertices
// whose rows hold `entries` CSR entries: their CSR rows, a label and an x gather
// per entry, and per vertex its id, degree, x and y (spectral_step, :109-136).
inline double step_bytes(double vertices, int tasks, double entries, std::size_t offset_bytes,
                         std::size_t weight_bytes, std::size_t value_bytes) {
    return (vertices + tasks) * static_cast<double>(offset_bytes) +
           entries * static_cast<double>(4 + weight_bytes) + entries * static_cast<double>(4 + value_bytes) +
           vertices * static_cast<double>(4 + 8 + 2 * value_bytes);
}

// Order-preserving integer image of a double; -0.0 and +0.0 map together so that
// (x, id) ties match the reference comparator (x[a] == x[b] for signed zeros).
__device__ __forceinline__ std::uint64_t ordered_bits(double v) {
    if (v == 0.0) v = 0.0;
    const std::uint64_t b = static_cast<std::uint64_t>(__double_as_longlong(v));
    return (b >> 63) ? ~b : (b | (1ULL << 63));
}

__device__ __forceinline__ std::uint64_t ordered_bits(float v) {
    if (v == 0.0f) v = 0.0f;
    const std::uint32_t b = static_cast<std::uint32_t>(__float_as_uint(v));
    return (b >> 31) ? static_cast<std::uint32_t>(~b) : (b | 0x80000000U);
}

template <typename T>
struct VecTraits;
template <>
struct VecTraits<double> {
    __device__ static double fmadd(double w, double x, double s) { return fma(w, x, s); }
};
template <>
struct VecTraits<float> {
    __device__ static float fmadd(float w, float x, float s) { return fmaf(w, x, s); }
};

struct BatchShared {
    std::uint64_t keys[kMaxBatchedTask];
    int ids[kMaxBatchedTask];
    double shared8[24];
    Norm norm;
    int boundary;
    int status;
};

// floor(weight * lp / parts) exactly (the reference evaluates it in long double,
// which is exact for weight < 2^58).
__device__ __forceinline__ std::uint64_t split_target(std::uint64_t weight, int lp, int parts) {
    const std::uint64_t q = weight / static_cast<std::uint64_t>(parts);
    const std::uint64_t r = weight % static_cast<std::uint64_t>(parts);
    return q * static_cast<std::uint64_t>(lp) + (r * static_cast<std::uint64_t>(lp)) / static_cast<std::uint64_t>(parts);
}

template <typename T>
__device__ __forceinline__ T normalized(const T* x, int v, const Norm& norm) {
    return static_cast<T>((static_cast<double>(x[v]) - norm.mean) * norm.scale);
}

// The final (x, id) sort (:143-148, :230), the inclusive weight prefix and
// feasible split (:231-248) and the label assignment (:149-153) of one task of at
// most kMaxBatchedTask vertices, by one workgroup of kBatchThreads threads.
// value(i, v) is the final normalised coordinate of the task's i-th vertex v =
// ids[i]; thread 0 calls emit(child) for each child of more than one part, right
// child first. Returns the split status to every thread: 0, 1 (no feasible split)
// or 2 (a one-part side above capacity); on a failure no label is written.
template <typename T, typename Value, typename Emit>
__device__ int split_task(const gpart_bisect_task& task, int* ids, int* labels, const std::uint64_t* vertex_weights,
                          std::uint64_t capacity, double* coords, BatchShared& sh, Value value, Emit emit) {
    const int count = task.count;
    // (x, id) sort: bitonic over the next power of two, padding keys sort last
    int size = 1;
    while (size < count) size <<= 1;
    for (int i = threadIdx.x; i < size; i += kBatchThreads) {
        if (i < count) {
            const int v = ids[i];
            const T xv = value(i, v);
            sh.keys[i] = ordered_bits(xv);
            sh.ids[i] = v;
            if (coords) coords[v] = static_cast<double>(xv);
        } else {
            sh.keys[i] = ~0ULL;
            sh.ids[i] = 0x7fffffff;
        }
    }
    __syncthreads();
    for (int k = 2; k <= size; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < size; i += kBatchThreads) {
                const int p = i ^ j;
                if (p > i) {
                    const bool up = (i & k) == 0;
                    const std::uint64_t ki = sh.keys[i], kp = sh.keys[p];
                    const int ii = sh.ids[i], ip = sh.ids[p];
                    const bool greater = ki > kp || (ki == kp && ii > ip);
                    if (greater == up) {
                        sh.keys[i] = kp;
                        sh.keys[p] = ki;
                        sh.ids[i] = ip;
                        sh.ids[p] = ii;
                    }
                }
            }
            __syncthreads();
        }
    }
    // inclusive weight prefix of the sorted segment, in place of the keys
    for (int i = threadIdx.x; i < count; i += kBatchThreads) {
        ids[i] = sh.ids[i];
        sh.keys[i] = vertex_weights[sh.ids[i]];
    }
    __syncthreads();
    for (int offset = 1; offset < count; offset <<= 1) {
        std::uint64_t add[kMaxBatchedTask / kBatchThreads] = {};
        int slot = 0;
        for (int i = threadIdx.x; i < count; i += kBatchThreads, ++slot)
            add[slot] = i >= offset ? sh.keys[i - offset] : 0;
        __syncthreads();
        slot = 0;
        for (int i = threadIdx.x; i < count; i += kBatchThreads, ++slot) sh.keys[i] += add[slot];
        __syncthreads();
    }
    // split selection (src/initial_partition_gpu.cu:233-248)
    if (threadIdx.x == 0) {
        const int lp = task.parts / 2, rp = task.parts - lp;
        const std::uint64_t W_ = task.weight;
        const std::uint64_t target = split_target(W_, lp, task.parts);
        const std::uint64_t left_limit = capacity > W_ / lp ? W_ : capacity * lp;
        const std::uint64_t right_limit = capacity > W_ / rp ? W_ : capacity * rp;
        const std::uint64_t min_left = W_ > right_limit ? W_ - right_limit : 0;
        int lo = 0, hi = count;  // lower_bound(prefix, target)
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (sh.keys[mid] < target) lo = mid + 1;
            else hi = mid;
        }
        int boundary = lo + 1;
        boundary = max(lp, min(count - rp, boundary));
        std::uint64_t left_weight = sh.keys[boundary - 1];
        if (left_weight > left_limit && boundary > lp) {
            --boundary;
            left_weight = sh.keys[boundary - 1];
        }
        if (left_weight < min_left && boundary < count - rp) {
            ++boundary;
            left_weight = sh.keys[boundary - 1];
        }
        int st = 0;
        if (left_weight > left_limit || left_weight < min_left) st = 1;
        const std::uint64_t right_weight = W_ - left_weight;
        if (!st) {
            if (lp == 1 && left_weight > capacity) st = 2;
            if (rp == 1 && right_weight > capacity) st = 2;
        }
        sh.boundary = boundary;
        sh.status = st;
        if (!st) {
            if (rp > 1)
                emit(gpart_bisect_task{task.begin + boundary, count - boundary, task.first_part + lp, rp,
                                       right_weight});
            if (lp > 1) emit(gpart_bisect_task{task.begin, boundary, task.first_part, lp, left_weight});
        }
    }
    __syncthreads();
    const int status = sh.status;
    if (status) return status;
    const int boundary = sh.boundary;
    const int lp = task.parts / 2;
    for (int i = threadIdx.x; i < count; i += kBatchThreads)
        labels[sh.ids[i]] = i < boundary ? task.first_part : task.first_part + lp;
    return 0;
}

}  // namespace
}  // namespace gpart_hip
