// initial.hip -- weighted recursive spectral bisection for gfx950
// (src/initial_partition_gpu.cu of the reference).
//
// Both forms keep the reference's FP64 summation order so that results are
// reproducible and identical between forms:
//   * a vertex is owned by 32 lanes (half a wave64); lane l accumulates entries
//     e = begin+l, begin+l+32, ... in increasing e, then a 32-lane shuffle-down
//     tree (16, 8, 4, 2, 1) combines them;
//   * moments are combined per 8 consecutive vertices with the reference's
//     block tree ((z0+z4)+(z2+z6))+((z1+z5)+(z3+z7)), z = m + 0.0;
//   * the normalisation sums those partials with 256 strided accumulators and
//     the same trees.
// Multiply-adds are fused only where written with fma() (the SpMV accumulation
// and the variance); the library is built with -ffp-contract=off.
//
// Plain form: the reference's launch structure, one task at a time (start, then
// per step: SpMV, finish_moments, normalize).
// Batched form: all tasks of one recursion depth share every launch. A depth is
// 66 launches whatever its task count: start, 64 steps (each also finishing the
// previous step's moments and normalising on read), and one 1024-thread
// workgroup per task for the (x, id) sort (LDS bitonic), the weight prefix, the
// feasible split and the label assignment, emitting the child tasks.
#include "internal.hpp"
#include "scan.hpp"
#include "spectral_device.hpp"

#include <algorithm>
#include <cstring>
#include <cmath>

namespace gpart_hip {
namespace {

// ------------------------------------------------------------------ plain form

template <typename O, typename W>
__global__ void __launch_bounds__(256) spectral_start_kernel(int count, const int* ids, const int* labels, int group,
                                                             const O* offsets, const int* neighbors, const W* weights,
                                                             double* degree, double* x, unsigned seed,
                                                             Moments* partials, unsigned long long* entries) {
    __shared__ Moments vm[8];
    const int lane = threadIdx.x & 31, hw = threadIdx.x >> 5;
    const int i = blockIdx.x * 8 + hw;
    double d = 0;
    if (i < count) {
        const int v = ids[i];
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + lane; e < offsets[v + 1]; e += 32)
            if (labels[neighbors[e]] == group) d += static_cast<double>(weights[e]);
    }
    d = tree32(d);
    if (lane == 0) {
        Moments m{0, 0, 0};
        if (i < count) {
            const int v = ids[i];
            // CSR entries each diffusion step scans (for the family's algorithmic bytes)
            atomicAdd(entries, static_cast<unsigned long long>(offsets[v + 1] - offsets[v]));
            const double a = static_cast<double>(mix32(static_cast<unsigned>(v) ^ seed)) / 4294967296.0 - 0.5;
            degree[v] = d;
            x[v] = a;
            const double mass = d > 0 ? d : 1.0;
            m = Moments{mass, mass * a, mass * a * a};
        }
        vm[hw] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = block_partial(vm, 0, 8);
}

template <typename O, typename W, typename T>
__global__ void __launch_bounds__(256) spectral_step_kernel(int count, const int* ids, const int* labels, int group,
                                                            const O* offsets, const int* neighbors, const W* weights,
                                                            const double* degree, const T* x, T* y,
                                                            Moments* partials) {
    __shared__ Moments vm[8];
    const int lane = threadIdx.x & 31, hw = threadIdx.x >> 5;
    const int i = blockIdx.x * 8 + hw;
    double sum = 0;
    if (i < count) {
        const int v = ids[i];
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + lane; e < offsets[v + 1]; e += 32) {
            const int u = neighbors[e];
            if (labels[u] == group) sum = fma(static_cast<double>(weights[e]), static_cast<double>(x[u]), sum);
        }
    }
    sum = tree32(sum);
    if (lane == 0) {
        Moments m{0, 0, 0};
        if (i < count) {
            const int v = ids[i];
            const double d = degree[v];
            const double a = d > 0 ? 0.5 * (static_cast<double>(x[v]) + sum / d) : static_cast<double>(x[v]);
            y[v] = static_cast<T>(a);
            const double mass = d > 0 ? d : 1.0;
            m = Moments{mass, mass * a, mass * a * a};
        }
        vm[hw] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = block_partial(vm, 0, 8);
}

__global__ void __launch_bounds__(256) finish_moments_kernel(int count, const Moments* partials, Norm* norm) {
    __shared__ double shared8[24];
    __shared__ Norm shared_norm;
    const Norm r = finish_block(partials, count, shared8, &shared_norm);
    if (threadIdx.x == 0) *norm = r;
}

__global__ void normalize_kernel(int count, const int* ids, double* x, const Norm* norm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const int v = ids[i];
        x[v] = (x[v] - norm->mean) * norm->scale;
    }
}

__global__ void sort_keys_kernel(int count, const int* ids, const double* x, std::uint64_t* keys, std::uint64_t* vals,
                                 bool by_id) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int v = ids[i];
    keys[i] = by_id ? static_cast<std::uint64_t>(v) : ordered_bits(x[v]);
    vals[i] = static_cast<std::uint64_t>(v);
}

__global__ void copy_ids_kernel(int count, const std::uint64_t* vals, int* ids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) ids[i] = static_cast<int>(vals[i]);
}

struct OrderWeightLoad {
    const int* ids;
    const std::uint64_t* weights;
    __device__ __forceinline__ std::uint64_t operator()(std::int64_t i) const { return weights[ids[i]]; }
};

__global__ void assign_split_kernel(int count, const int* ids, int boundary, int left, int right, int* labels) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) labels[ids[i]] = i < boundary ? left : right;
}

// ------------------------------------------------------------------ batched form

// All tasks of one depth share each launch: the grid is the concatenation of the
// tasks' 8-vertex blocks (the reference's own mapping), so every task, small or
// large, is spread over the whole GPU. One step is one launch: each workgroup
// first recomputes its task's normalisation from the previous step's block
// partials (finish_block, the reference's fixed order) and normalises x values as
// it reads them, (x - mean) * scale, the same expression the reference stores.
constexpr int kMaxDepthTasks = 32;

struct DepthTasks {
    int count;
    int first_block[kMaxDepthTasks + 1];  // prefix of ceil(task.count / 8)
    gpart_bisect_task task[kMaxDepthTasks];
};

__device__ __forceinline__ int task_of_block(const DepthTasks& dt, int block) {
    int t = 0;
    while (t + 1 < dt.count && dt.first_block[t + 1] <= block) ++t;
    return t;
}

// spectral_start (:81-107) for every task of the depth.
template <typename O, typename W, typename T>
__global__ void __launch_bounds__(256) depth_start_kernel(DepthTasks dt, const int* order, const int* labels,
                                                          const O* offsets, const int* neighbors, const W* weights,
                                                          unsigned seed, double* degree, T* x, Moments* partials,
                                                          unsigned long long* entries) {
    __shared__ Moments vm[8];
    const int t = task_of_block(dt, blockIdx.x);
    const gpart_bisect_task task = dt.task[t];
    const int* ids = order + task.begin;
    const int group = task.first_part;
    const unsigned tseed = seed ^ static_cast<unsigned>(task.first_part);
    const int lane = threadIdx.x & 31, hw = threadIdx.x >> 5;
    const int i = (blockIdx.x - dt.first_block[t]) * 8 + hw;
    double d = 0;
    if (i < task.count) {
        const int v = ids[i];
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + lane; e < offsets[v + 1]; e += 32)
            if (labels[neighbors[e]] == group) d += static_cast<double>(weights[e]);
    }
    d = tree32(d);
    if (lane == 0) {
        Moments m{0, 0, 0};
        if (i < task.count) {
            const int v = ids[i];
            atomicAdd(entries, static_cast<unsigned long long>(offsets[v + 1] - offsets[v]));
            const double a = static_cast<double>(mix32(static_cast<unsigned>(v) ^ tseed)) / 4294967296.0 - 0.5;
            degree[v] = d;
            x[v] = static_cast<T>(a);
            const double av = static_cast<double>(static_cast<T>(a));
            const double mass = d > 0 ? d : 1.0;
            m = Moments{mass, mass * av, mass * av * av};
        }
        vm[hw] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = block_partial(vm, 0, 8);
}

// finish_moments (:65-79) of the previous step, normalize (:138-142) on read, and
// spectral_step (:109-136), for every task of the depth.
template <typename O, typename W, typename T>
__global__ void __launch_bounds__(256) depth_step_kernel(DepthTasks dt, const int* order, const int* labels,
                                                         const O* offsets, const int* neighbors, const W* weights,
                                                         const double* degree, const T* x, T* y,
                                                         const Moments* previous, Moments* partials) {
    __shared__ Moments vm[8];
    __shared__ double shared8[24];
    __shared__ Norm shared_norm;
    const int t = task_of_block(dt, blockIdx.x);
    const gpart_bisect_task task = dt.task[t];
    const int first = dt.first_block[t];
    const Norm norm = finish_block(previous + first, dt.first_block[t + 1] - first, shared8, &shared_norm);
    const int* ids = order + task.begin;
    const int group = task.first_part;
    const int lane = threadIdx.x & 31, hw = threadIdx.x >> 5;
    const int i = (blockIdx.x - first) * 8 + hw;
    T sum = 0;
    if (i < task.count) {
        const int v = ids[i];
        for (std::int64_t e = static_cast<std::int64_t>(offsets[v]) + lane; e < offsets[v + 1]; e += 32) {
            const int u = neighbors[e];
            if (labels[u] == group) sum = VecTraits<T>::fmadd(static_cast<T>(weights[e]), normalized(x, u, norm), sum);
        }
    }
    double sd;
    if constexpr (sizeof(T) == 8) {
        sd = tree32(static_cast<double>(sum));
    } else {
        float f = sum;
#pragma unroll
        for (int s2 = 16; s2; s2 >>= 1) f += __shfl_down(f, s2, kHalf);
        sd = static_cast<double>(f);
    }
    if (lane == 0) {
        Moments m{0, 0, 0};
        if (i < task.count) {
            const int v = ids[i];
            const double d = degree[v];
            const T xv = normalized(x, v, norm);
            double a;
            if constexpr (sizeof(T) == 8) {
                a = d > 0 ? 0.5 * (static_cast<double>(xv) + sd / d) : static_cast<double>(xv);
            } else {
                const float df = static_cast<float>(d);
                const float af = d > 0 ? 0.5f * (xv + static_cast<float>(sd) / df) : xv;
                a = static_cast<double>(af);
            }
            y[v] = static_cast<T>(a);
            const double mass = d > 0 ? d : 1.0;
            m = Moments{mass, mass * a, mass * a * a};
        }
        vm[hw] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = block_partial(vm, 0, 8);
}

// The last normalisation, the (x, id) sort (:143-148, :230), the weight prefix and
// feasible split (:231-248) and the label assignment (:149-153): one workgroup per
// task, emitting the child tasks of the next depth.
template <typename T>
__global__ void __launch_bounds__(kBatchThreads) depth_finish_kernel(DepthTasks dt, int* labels, int* order,
                                                                     const std::uint64_t* vertex_weights,
                                                                     std::uint64_t capacity, const T* x,
                                                                     const Moments* partials, double* coords,
                                                                     gpart_bisect_task* next, int* next_count,
                                                                     int* status) {
    __shared__ BatchShared sh;
    const int t = blockIdx.x;
    const gpart_bisect_task task = dt.task[t];
    const int first = dt.first_block[t];
    const Norm norm = finish_block(partials + first, dt.first_block[t + 1] - first, sh.shared8, &sh.norm);
    const int st = split_task<T>(
        task, order + task.begin, labels, vertex_weights, capacity, coords, sh,
        [&](int, int v) { return normalized(x, v, norm); },
        [&](const gpart_bisect_task& child) { next[atomicAdd(next_count, 1)] = child; });
    if (st && threadIdx.x == 0) atomicMax(status, st);
}

template <typename O, typename W>
void plain_task(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, int32_t* order,
                const std::uint64_t* vertex_weights, const gpart_bisect_task* task, uint32_t seed, double* x,
                double* y, double* degree, std::uint64_t* prefix) {
    const int count = task->count;
    if (count <= 0) return;
    int* ids = order + task->begin;
    const auto* off = static_cast<const O*>(g->offsets);
    const auto* w = static_cast<const W*>(g->edge_weights);
    const int blocks = (count + 255) / 256;
    const int spectral_blocks = (count + 7) / 8;
    Scratch<Moments> partials(ctx, spectral_blocks);
    Scratch<Norm> norm(ctx, 1);
    Scratch<unsigned long long> entries(ctx, 1);
    fill_bytes(ctx, entries.get(), 0, 8);
    profile_begin(ctx, GPART_PROFILE_SPECTRAL);
    hipLaunchKernelGGL((spectral_start_kernel<O, W>), dim3(spectral_blocks), dim3(256), 0, ctx->stream, count, ids,
                       labels, task->first_part, off, g->neighbors, w, degree, x,
                       seed ^ static_cast<unsigned>(task->first_part), partials.get(), entries.get());
    hipLaunchKernelGGL(finish_moments_kernel, dim3(1), dim3(256), 0, ctx->stream, spectral_blocks, partials.get(),
                       norm.get());
    GPART_LAUNCH_CHECK();
    for (int iteration = 0; iteration <= 64; ++iteration) {
        hipLaunchKernelGGL(normalize_kernel, dim3(blocks), dim3(256), 0, ctx->stream, count, ids, x, norm.get());
        if (iteration == 64) break;
        hipLaunchKernelGGL((spectral_step_kernel<O, W, double>), dim3(spectral_blocks), dim3(256), 0, ctx->stream,
                           count, ids, labels, task->first_part, off, g->neighbors, w, degree, x, y, partials.get());
        hipLaunchKernelGGL(finish_moments_kernel, dim3(1), dim3(256), 0, ctx->stream, spectral_blocks, partials.get(),
                           norm.get());
        GPART_LAUNCH_CHECK();
        std::swap(x, y);
    }
    profile_end(ctx, GPART_PROFILE_SPECTRAL, 0.0);
    if (ctx->profile[GPART_PROFILE_SPECTRAL].enabled)
        profile_add_bytes(ctx, GPART_PROFILE_SPECTRAL,
                          64.0 * step_bytes(count, 1, static_cast<double>(read_scalar(ctx, entries.get())), sizeof(O),
                                            sizeof(W), sizeof(double)));
    // (x, id) total order: stable sort by id, then stable sort by x
    Scratch<std::uint64_t> keys(ctx, count), keys_alt(ctx, count), vals(ctx, count), vals_alt(ctx, count);
    hipLaunchKernelGGL(sort_keys_kernel, dim3(blocks), dim3(256), 0, ctx->stream, count, ids, x, keys.get(), vals.get(),
                       true);
    GPART_LAUNCH_CHECK();
    std::uint64_t* cur_v = vals.get();
    if (radix_sort_pairs(ctx, keys.get(), keys_alt.get(), vals.get(), vals_alt.get(), count, 0, 31))
        cur_v = vals_alt.get();
    hipLaunchKernelGGL(copy_ids_kernel, dim3(blocks), dim3(256), 0, ctx->stream, count, cur_v, ids);
    hipLaunchKernelGGL(sort_keys_kernel, dim3(blocks), dim3(256), 0, ctx->stream, count, ids, x, keys.get(), vals.get(),
                       false);
    GPART_LAUNCH_CHECK();
    cur_v = vals.get();
    if (radix_sort_pairs(ctx, keys.get(), keys_alt.get(), vals.get(), vals_alt.get(), count, 0, 64))
        cur_v = vals_alt.get();
    hipLaunchKernelGGL(copy_ids_kernel, dim3(blocks), dim3(256), 0, ctx->stream, count, cur_v, ids);
    GPART_LAUNCH_CHECK();
    device_scan<std::uint64_t, SumOp<std::uint64_t>>(ctx, OrderWeightLoad{ids, vertex_weights},
                                                      InclusiveStore<std::uint64_t>{prefix}, count,
                                                      SumOp<std::uint64_t>{});
    // x now holds the final coordinates (the caller's x pointer after an even swap count)
}

}  // namespace
}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

int gpart_spectral_task(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, int32_t* order,
                        const uint64_t* vertex_weights, const gpart_bisect_task* task, uint32_t seed, double* x,
                        double* y, double* degree, uint64_t* prefix) {
    return guarded([&] {
        if (g->width == GPART_WIDTH_STANDARD)
            plain_task<std::int32_t, std::uint32_t>(ctx, g, labels, order, vertex_weights, task, seed, x, y, degree,
                                                    prefix);
        else
            plain_task<std::int64_t, std::uint64_t>(ctx, g, labels, order, vertex_weights, task, seed, x, y, degree,
                                                    prefix);
    });
}

int gpart_spectral_assign(gpart_ctx* ctx, const int32_t* order, const gpart_bisect_task* task, int32_t boundary,
                          int32_t left_parts, int32_t* labels) {
    return guarded([&] {
        const int count = task->count;
        if (count <= 0) return;
        hipLaunchKernelGGL(assign_split_kernel, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, count,
                           order + task->begin, boundary, task->first_part, task->first_part + left_parts, labels);
        GPART_LAUNCH_CHECK();
    });
}

int64_t gpart_spectral_depth_max_task(void) { return kMaxBatchedTask; }

int gpart_spectral_depth(gpart_ctx* ctx, const gpart_csr_view* g, int32_t* labels, int32_t* order,
                         const uint64_t* vertex_weights, const gpart_bisect_task* tasks, int32_t task_count,
                         uint32_t seed, uint64_t capacity, int precision, double* coords, gpart_bisect_task* next,
                         int32_t* next_count, int32_t* status) {
    return guarded([&] {
        *next_count = 0;
        *status = 0;
        if (task_count <= 0) return;
        if (task_count > kMaxDepthTasks) throw Error(GPART_ERR_INVALID_ARGUMENT, "too many tasks in one depth");
        DepthTasks dt{};
        dt.count = task_count;
        dt.first_block[0] = 0;
        for (int t = 0; t < task_count; ++t) {
            if (tasks[t].count < 1 || tasks[t].count > kMaxBatchedTask)
                throw Error(GPART_ERR_INVALID_ARGUMENT, "batched bisection task size out of range");
            dt.task[t] = tasks[t];
            dt.first_block[t + 1] = dt.first_block[t] + (tasks[t].count + 7) / 8;
        }
        const int blocks = dt.first_block[task_count];
        const std::int64_t n = g->n;
        Scratch<double> degree(ctx, n), xa(ctx, n), xb(ctx, n);
        Scratch<Moments> pa(ctx, blocks), pb(ctx, blocks);
        // next_count and status (two ints), then the scanned CSR entries
        Scratch<unsigned long long> counters(ctx, 2);
        fill_bytes(ctx, counters.get(), 0, 16);
        int* next_and_status = reinterpret_cast<int*>(counters.get());
        unsigned long long* entries = counters.get() + 1;
        int vertices = 0;
        for (int t = 0; t < task_count; ++t) vertices += tasks[t].count;
        profile_begin(ctx, GPART_PROFILE_SPECTRAL);
        auto launch = [&](auto o_tag, auto w_tag, auto t_tag) {
            using O = decltype(o_tag);
            using W = decltype(w_tag);
            using T = decltype(t_tag);
            const auto* off = static_cast<const O*>(g->offsets);
            const auto* w = static_cast<const W*>(g->edge_weights);
            T* x = reinterpret_cast<T*>(xa.get());
            T* y = reinterpret_cast<T*>(xb.get());
            Moments* p = pa.get();
            Moments* q = pb.get();
            hipLaunchKernelGGL((depth_start_kernel<O, W, T>), dim3(blocks), dim3(256), 0, ctx->stream, dt, order,
                               labels, off, g->neighbors, w, seed, degree.get(), x, p, entries);
            for (int step = 0; step < 64; ++step) {
                hipLaunchKernelGGL((depth_step_kernel<O, W, T>), dim3(blocks), dim3(256), 0, ctx->stream, dt, order,
                                   labels, off, g->neighbors, w, degree.get(), x, y, p, q);
                std::swap(x, y);
                std::swap(p, q);
            }
            hipLaunchKernelGGL((depth_finish_kernel<T>), dim3(task_count), dim3(kBatchThreads), 0, ctx->stream, dt,
                               labels, order, vertex_weights, capacity, x, p, coords, next, next_and_status,
                               next_and_status + 1);
            GPART_LAUNCH_CHECK();
        };
        const bool fp32 = precision == GPART_SPECTRAL_FP32;
        if (g->width == GPART_WIDTH_STANDARD) {
            if (fp32) launch(std::int32_t{}, std::uint32_t{}, float{});
            else launch(std::int32_t{}, std::uint32_t{}, double{});
        } else {
            if (fp32) launch(std::int64_t{}, std::uint64_t{}, float{});
            else launch(std::int64_t{}, std::uint64_t{}, double{});
        }
        profile_end(ctx, GPART_PROFILE_SPECTRAL, 0.0);
        unsigned long long h[2];
        read_back(ctx, h, counters.get(), sizeof(h));
        int ns[2];
        std::memcpy(ns, h, sizeof(ns));
        *next_count = ns[0];
        *status = ns[1];
        const std::size_t o_size = g->width == GPART_WIDTH_STANDARD ? 4 : 8;
        profile_add_bytes(ctx, GPART_PROFILE_SPECTRAL,
                          64.0 * step_bytes(vertices, task_count, static_cast<double>(h[1]), o_size, o_size,
                                            fp32 ? sizeof(float) : sizeof(double)));
    });
}

}  // extern "C"
