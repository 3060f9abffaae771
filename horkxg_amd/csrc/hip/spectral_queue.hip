// spectral_queue.hip -- the batched weighted recursive spectral bisection as one
// persistent launch over a device work queue. It replaces the reference's host
// task loop (src/initial_partition_gpu.cu:198-255: a stack of tasks, about 195
// launches per task) for every task of at most kMaxBatchedTask vertices, of any
// recursion depth and of any number of graphs.
//
// Queue. Every bisection task is a record (task, graph, team size, barrier
// words). A task of c vertices is served by a team of P workgroups of 1024
// threads, P = ceil(ceil(c / 8) / own_blocks), and occupies P consecutive slots
// of one slot array; a slot word names the task and the team member. Workgroups
// take slot indices from one counter in increasing order and wait until their
// slot is written. Parents write their children's records and slots only after
// the children's labels and order segments are final, and the count of
// unfinished tasks goes up for the children before it goes down for the parent,
// so it reads zero only when no task will ever be written again: a workgroup
// waiting on an unwritten slot then exits.
//
// Progress. Teams wait for each other at barriers, so a team needs all its
// members running at once. Slot indices are handed out in order and a team's
// slots are consecutive, so the earliest team with a member missing always gets
// the next index; the workgroups holding later indices are either waiting for
// their own (later) slot or run complete teams that finish. With at least P
// workgroups of the launch running at once the queue drains. Every wait is
// bounded: after kQueueSpinLimit polls a waiter raises the abort word, which all
// waiters watch, and the host reports the abort as an error.
//
// Numerics. Each member owns a contiguous run of the task's 8-vertex blocks and
// computes exactly what the per-depth launches compute for them (initial.hip):
// the same 32-lane row sums in entry order, the same block partials, and the
// normalisation from all of the task's partials in the same order. Results are
// bit-identical to the per-depth form in both precisions. What differs is where
// values live: the task's normalised vector is staged in LDS by position, and
// each row entry's neighbour is resolved once per task to a position in the task
// (or -1 for a neighbour of another task), so a step reads two arrays per entry
// and gathers from LDS.
#include "internal.hpp"
#include "spectral_device.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

namespace gpart_hip {
namespace {

constexpr int kQueueMaxOwnBlocks = 128;                      // blocks per member at most
constexpr int kQueueOwn = kQueueMaxOwnBlocks * 8;            // vertices per member at most
constexpr int kQueueRows = 4;                                // rows in flight per half-wave
constexpr int kQueueHalfWaves = kBatchThreads / kHalf;       // 32
constexpr unsigned long long kQueueSpinLimit = 1ULL << 22;  // about 4 s of polling per wait
constexpr unsigned long long kSlotValid = 1ULL << 63;
constexpr int kMemberBits = 20;

static_assert(kBatchThreads == 1024, "one member is 32 half-waves");

// defaults from the autotuner's choice
std::atomic<int> queue_own_blocks{gpart_tuned::kQueueOwnBlocks};  // the team-size rule's blocks per member
std::atomic<int> queue_max_task{gpart_tuned::kQueueSingleMaxTask > 0 ? gpart_tuned::kQueueSingleMaxTask
                                                                     : kMaxBatchedTask};
std::atomic<int> queue_enabled{gpart_tuned::kQueueSingleMaxTask > 0 ? 1 : 0};  // single-graph driver uses it

struct QueueGraph {
    const void* offsets;
    const int* neighbors;
    const void* edge_weights;
    int* labels;
    int* order;
    const std::uint64_t* vertex_weights;
    double* coords;
    void* y[2];            // per position of the graph's tasks, T each
    Moments* partials[2];  // per position (a task's blocks start at its begin)
    int* pos;              // vertex -> position in its task
    int* local;            // row entry -> position of a same-task neighbour, or -1
    std::uint64_t capacity;
    unsigned seed;
    int pad;
};

struct QueueTask {
    gpart_bisect_task task;
    int graph;
    int team;
    unsigned long long arrive;
    unsigned long long release;
};

struct QueueState {
    unsigned long long head;         // next slot index to hand out
    unsigned long long tail;         // slots reserved so far
    unsigned long long tasks;        // task records written so far
    unsigned long long outstanding;  // tasks written and not finished
    unsigned long long entries;      // CSR entries of every served task (algorithmic bytes)
    unsigned long long served;       // tasks finished
    unsigned long long vertices;     // vertices of the finished tasks
    unsigned int abort;
    unsigned int overflow;
};

struct QueueArgs {
    const QueueGraph* graphs;
    QueueTask* tasks;
    unsigned long long* slots;
    QueueState* state;
    int* status;  // per graph: 0, 1 (no feasible split) or 2 (leaf above capacity)
    long long slot_capacity;
    long long task_capacity;
    int own_blocks;
};

template <typename T>
struct QueueStep {
    T x[kMaxBatchedTask];  // the task's normalised vector, by position
    Moments vm[kQueueOwn];
    long long row_begin[kQueueOwn];
    long long row_end[kQueueOwn];
    double degree[kQueueOwn];
};

template <typename T>
struct QueueShared {
    union {
        QueueStep<T> step;
        BatchShared finish;
    } u;
    double shared8[24];
    Norm norm;
    gpart_bisect_task children[2];
    int child_count;
    unsigned long long word;
    unsigned long long entries;
    int done;
    int barrier_ok;
};

// Polls read relaxed (an agent-scope acquire invalidates the caches, which would
// evict the data of every task running on the XCD at each poll); the waiter takes
// one acquire fence once the wait is over.
__device__ __forceinline__ unsigned long long poll_agent(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool aborted(QueueState* state) {
    return __hip_atomic_load(&state->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}
__device__ __forceinline__ void raise_abort(QueueState* state) {
    __hip_atomic_store(&state->abort, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int team_size(int count, int own_blocks) {
    const int blocks = (count + 7) / 8;
    return max(1, (blocks + own_blocks - 1) / own_blocks);
}

// The team's barrier number `epoch`: every member writes back its stores and
// arrives; the last arrival releases the epoch. Returns false on abort.
template <typename T>
__device__ bool team_barrier(QueueTask* rec, int team, unsigned long long epoch, QueueState* state,
                             QueueShared<T>& sh) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned long long before =
            __hip_atomic_fetch_add(&rec->arrive, 1ULL, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (before + 1 == epoch * static_cast<unsigned long long>(team)) {
            __hip_atomic_store(&rec->release, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned long long spins = 0;
            while (poll_agent(&rec->release) < epoch) {
                if (++spins > kQueueSpinLimit || aborted(state)) {
                    raise_abort(state);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        sh.barrier_ok = ok;
    }
    __syncthreads();
    return sh.barrier_ok != 0;
}

// Writes a child's record and its team's slots (thread 0 only). The caller has
// made the child's labels and order segment visible before.
__device__ void push_task(const QueueArgs& a, int graph, const gpart_bisect_task& child) {
    const int team = team_size(child.count, a.own_blocks);
    const unsigned long long id = atomicAdd(&a.state->tasks, 1ULL);
    const unsigned long long base = atomicAdd(&a.state->tail, static_cast<unsigned long long>(team));
    if (id >= static_cast<unsigned long long>(a.task_capacity) ||
        base + team > static_cast<unsigned long long>(a.slot_capacity)) {
        __hip_atomic_store(&a.state->overflow, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        raise_abort(a.state);
        return;
    }
    QueueTask* rec = a.tasks + id;
    rec->task = child;
    rec->graph = graph;
    rec->team = team;
    rec->arrive = 0;
    rec->release = 0;
    atomicAdd(&a.state->outstanding, 1ULL);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    for (int r = 0; r < team; ++r)
        __hip_atomic_store(&a.slots[base + r], kSlotValid | (id << kMemberBits) | static_cast<unsigned long long>(r),
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Row sums of up to kQueueRows own vertices per half-wave: lane l takes entries
// begin + l, begin + l + 32, ... of each row in increasing order (the reference's
// order), the rows' loads interleaved.
template <typename W, typename T>
__device__ __forceinline__ void row_sums(const QueueStep<T>& s, const int* local, const W* weights, int j0, int own,
                                         int lane, T (&sum)[kQueueRows]) {
    long long e[kQueueRows], end[kQueueRows];
#pragma unroll
    for (int r = 0; r < kQueueRows; ++r) {
        const int j = j0 + r * kQueueHalfWaves;
        sum[r] = 0;
        e[r] = j < own ? s.row_begin[j] + lane : 0;
        end[r] = j < own ? s.row_end[j] : 0;
    }
    for (;;) {
        bool any = false;
#pragma unroll
        for (int r = 0; r < kQueueRows; ++r) any = any || e[r] < end[r];
        if (!any) break;
        int l[kQueueRows];
        W w[kQueueRows];
#pragma unroll
        for (int r = 0; r < kQueueRows; ++r) {
            l[r] = -1;
            w[r] = 0;
            if (e[r] < end[r]) {
                l[r] = local[e[r]];
                w[r] = weights[e[r]];
            }
        }
#pragma unroll
        for (int r = 0; r < kQueueRows; ++r) {
            if (e[r] < end[r]) {
                if (l[r] >= 0) sum[r] = VecTraits<T>::fmadd(static_cast<T>(w[r]), s.x[l[r]], sum[r]);
                e[r] += kHalf;
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ double half_wave_total(T v) {
    if constexpr (sizeof(T) == 8) {
        return tree32(static_cast<double>(v));
    } else {
        float f = v;
#pragma unroll
        for (int s2 = 16; s2; s2 >>= 1) f += __shfl_down(f, s2, kHalf);
        return static_cast<double>(f);
    }
}

// One member's share of one task: start, 64 steps, and (member 0) the split.
template <typename O, typename W, typename T>
__device__ void serve(const QueueArgs& a, unsigned long long id, int member, QueueShared<T>& sh) {
    QueueTask* rec = a.tasks + id;
    const gpart_bisect_task task = rec->task;
    const int gi = rec->graph, team = rec->team;
    const QueueGraph& g = a.graphs[gi];
    const auto* offsets = static_cast<const O*>(g.offsets);
    const auto* weights = static_cast<const W*>(g.edge_weights);
    int* ids = g.order + task.begin;
    const int count = task.count, group = task.first_part;
    const int nb = (count + 7) / 8;
    const int b0 = static_cast<int>(static_cast<long long>(member) * nb / team);
    const int b1 = static_cast<int>(static_cast<long long>(member + 1) * nb / team);
    const int i0 = 8 * b0;
    const int own = min(count, 8 * b1) - i0;  // own vertices
    const int padded = 8 * (b1 - b0);         // own block slots
    const int lane = threadIdx.x & (kHalf - 1), hw = threadIdx.x / kHalf;
    T* y[2] = {static_cast<T*>(g.y[0]) + task.begin, static_cast<T*>(g.y[1]) + task.begin};
    Moments* p[2] = {g.partials[0] + task.begin, g.partials[1] + task.begin};
    QueueStep<T>& s = sh.u.step;
    unsigned long long epoch = 0;
    auto barrier = [&]() -> bool {
        if (team == 1) {
            __syncthreads();  // one workgroup: its own global stores are visible to it
            return true;
        }
        return team_barrier(rec, team, ++epoch, a.state, sh);
    };

    // positions of the task's vertices
    for (int j = threadIdx.x; j < own; j += kBatchThreads) g.pos[ids[i0 + j]] = i0 + j;
    if (!barrier()) return;

    // spectral_start (:81-107): degree to same-task neighbours, local neighbour
    // positions, the hashed start vector and its block moments
    const unsigned tseed = g.seed ^ static_cast<unsigned>(group);
    for (int j = hw; j < padded; j += kQueueHalfWaves) {
        double d = 0;
        if (j < own) {
            const int v = ids[i0 + j];
            const long long rb = static_cast<long long>(offsets[v]), re = static_cast<long long>(offsets[v + 1]);
            for (long long e = rb + lane; e < re; e += kHalf) {
                const int u = g.neighbors[e];
                const bool same = g.labels[u] == group;
                if (same) d += static_cast<double>(weights[e]);
                g.local[e] = same ? g.pos[u] : -1;
            }
            if (lane == 0) {
                s.row_begin[j] = rb;
                s.row_end[j] = re;
                atomicAdd(&sh.entries, static_cast<unsigned long long>(re - rb));
            }
        }
        d = tree32(d);
        if (lane == 0) {
            Moments m{0, 0, 0};
            if (j < own) {
                const int v = ids[i0 + j];
                const double start = static_cast<double>(mix32(static_cast<unsigned>(v) ^ tseed)) / 4294967296.0 - 0.5;
                s.degree[j] = d;
                const T xt = static_cast<T>(start);
                y[0][i0 + j] = xt;
                const double av = static_cast<double>(xt);
                const double mass = d > 0 ? d : 1.0;
                m = Moments{mass, mass * av, mass * av * av};
            }
            s.vm[j] = m;
        }
    }
    __syncthreads();
    for (int b = b0 + static_cast<int>(threadIdx.x); b < b1; b += kBatchThreads)
        p[0][b] = block_partial(s.vm, 8 * (b - b0), padded);
    if (!barrier()) return;

    // 64 steps of finish_moments (:65-79), normalize (:138-142) and
    // spectral_step (:109-136)
    int cur = 0;
    for (int step = 0; step < 64; ++step) {
        const int nxt = cur ^ 1;
        const Norm norm = finish_block(p[cur], nb, sh.shared8, &sh.norm);
        for (int i = threadIdx.x; i < count; i += kBatchThreads) s.x[i] = normalized(y[cur], i, norm);
        __syncthreads();
        for (int j0 = hw; j0 < padded; j0 += kQueueHalfWaves * kQueueRows) {
            T sum[kQueueRows];
            row_sums<W, T>(s, g.local, weights, j0, own, lane, sum);
#pragma unroll
            for (int r = 0; r < kQueueRows; ++r) {
                const int j = j0 + r * kQueueHalfWaves;
                const double sd = half_wave_total(sum[r]);
                if (lane == 0 && j < padded) {
                    Moments m{0, 0, 0};
                    if (j < own) {
                        const double d = s.degree[j];
                        const T xv = s.x[i0 + j];
                        double av;
                        if constexpr (sizeof(T) == 8) {
                            av = d > 0 ? 0.5 * (static_cast<double>(xv) + sd / d) : static_cast<double>(xv);
                        } else {
                            const float df = static_cast<float>(d);
                            const float af = d > 0 ? 0.5f * (xv + static_cast<float>(sd) / df) : xv;
                            av = static_cast<double>(af);
                        }
                        y[nxt][i0 + j] = static_cast<T>(av);
                        const double mass = d > 0 ? d : 1.0;
                        m = Moments{mass, mass * av, mass * av * av};
                    }
                    s.vm[j] = m;
                }
            }
        }
        __syncthreads();
        for (int b = b0 + static_cast<int>(threadIdx.x); b < b1; b += kBatchThreads)
            p[nxt][b] = block_partial(s.vm, 8 * (b - b0), padded);
        if (!barrier()) return;
        cur = nxt;
    }
    if (member != 0) return;

    // the last normalisation, the split and the labels (:143-153, :230-248)
    const Norm norm = finish_block(p[cur], nb, sh.shared8, &sh.norm);
    if (threadIdx.x == 0) sh.child_count = 0;
    __syncthreads();
    const int st = split_task<T>(
        task, ids, g.labels, g.vertex_weights, g.capacity, g.coords, sh.u.finish,
        [&](int i, int) { return normalized(y[cur], i, norm); },
        [&](const gpart_bisect_task& child) { sh.children[sh.child_count++] = child; });
    // every wave's label and order stores complete before thread 0 writes them
    // back and publishes the children
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (st) {
            atomicMax(a.status + gi, st);
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            for (int c = 0; c < sh.child_count; ++c) push_task(a, gi, sh.children[c]);
        }
        atomicAdd(&a.state->served, 1ULL);
        atomicAdd(&a.state->vertices, static_cast<unsigned long long>(count));
        __hip_atomic_fetch_add(&a.state->outstanding, ~0ULL, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename O, typename W, typename T>
__global__ void __launch_bounds__(kBatchThreads) spectral_queue_kernel(QueueArgs a) {
    __shared__ QueueShared<T> sh;
    if (threadIdx.x == 0) sh.entries = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            int done = 0;
            unsigned long long word = 0;
            const unsigned long long slot = atomicAdd(&a.state->head, 1ULL);
            if (slot >= static_cast<unsigned long long>(a.slot_capacity) || aborted(a.state)) {
                done = 1;
            } else {
                unsigned long long spins = 0;
                for (;;) {
                    word = poll_agent(&a.slots[slot]);
                    if (word) break;
                    if (poll_agent(&a.state->outstanding) == 0 || aborted(a.state)) {
                        done = 1;
                        break;
                    }
                    if (++spins > kQueueSpinLimit) {
                        raise_abort(a.state);
                        done = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            sh.done = done;
            sh.word = word;
        }
        __syncthreads();
        if (sh.done) break;
        const unsigned long long word = sh.word;
        const unsigned long long id = (word & ~kSlotValid) >> kMemberBits;
        const int member = static_cast<int>(word & ((1ULL << kMemberBits) - 1));
        serve<O, W, T>(a, id, member, sh);
        __syncthreads();
    }
    __syncthreads();
    if (threadIdx.x == 0 && sh.entries) atomicAdd(&a.state->entries, sh.entries);
}

int queue_grid(const void* kernel, int device) {
    int per_cu = 0, cus = 0;
    GPART_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBatchThreads, 0));
    GPART_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    return std::max(1, per_cu) * std::max(1, cus);
}

}  // namespace
}  // namespace gpart_hip

using namespace gpart_hip;

extern "C" {

int gpart_spectral_queue_set_tuning(int32_t own_blocks, int32_t max_task) {
    return guarded([&] {
        if (own_blocks < 8 || own_blocks > kQueueMaxOwnBlocks)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "blocks per queue member out of range");
        if (max_task < 1 || max_task > kMaxBatchedTask)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "largest queue task out of range");
        queue_own_blocks = own_blocks;
        queue_max_task = max_task;
    });
}

void gpart_spectral_queue_tuning(int32_t* own_blocks, int32_t* max_task) {
    if (own_blocks) *own_blocks = queue_own_blocks.load();
    if (max_task) *max_task = queue_max_task.load();
}

int gpart_spectral_set_queue(int32_t enabled) {
    return guarded([&] { queue_enabled = enabled != 0 ? 1 : 0; });
}

int32_t gpart_spectral_queue_enabled(void) { return queue_enabled.load(); }

int gpart_spectral_queue(gpart_ctx* ctx, gpart_queue_graph* graphs, int32_t graph_count,
                         const gpart_bisect_task* roots, const int32_t* root_graph, int32_t root_count,
                         int precision, gpart_queue_report* report) {
    return guarded([&] {
        if (report) *report = gpart_queue_report{};
        if (graph_count < 0 || root_count < 0) throw Error(GPART_ERR_INVALID_ARGUMENT, "negative queue sizes");
        if (precision != GPART_SPECTRAL_FP64 && precision != GPART_SPECTRAL_FP32)
            throw Error(GPART_ERR_INVALID_ARGUMENT, "unknown spectral precision");
        for (int i = 0; i < graph_count; ++i) graphs[i].status = 0;
        if (root_count == 0) return;
        if (graph_count == 0) throw Error(GPART_ERR_INVALID_ARGUMENT, "queue roots without graphs");
        const int width = graphs[0].graph.width;
        std::int64_t total_n = 0, total_m = 0;
        for (int i = 0; i < graph_count; ++i) {
            const auto& q = graphs[i];
            if (q.graph.width != width) throw Error(GPART_ERR_INVALID_ARGUMENT, "queue graphs of mixed width");
            if (q.graph.n < 1 || q.graph.n > (std::int64_t{1} << 30))
                throw Error(GPART_ERR_INVALID_ARGUMENT, "queue graph size out of range");
            if (!q.labels || !q.order || !q.vertex_weights)
                throw Error(GPART_ERR_INVALID_ARGUMENT, "queue graph without labels, order or weights");
            total_n += q.graph.n;
            total_m += q.graph.m;
        }
        const int own_blocks = queue_own_blocks.load();
        long long task_capacity = 0, slot_capacity = 0, max_team = 0, concurrent = 0;
        for (int r = 0; r < root_count; ++r) {
            const auto& t = roots[r];
            const int gi = root_graph[r];
            if (gi < 0 || gi >= graph_count) throw Error(GPART_ERR_INVALID_ARGUMENT, "queue root of no graph");
            if (t.count < 1 || t.count > kMaxBatchedTask || t.parts < 2 || t.parts > 32 || t.begin < 0 ||
                static_cast<std::int64_t>(t.begin) + t.count > graphs[gi].graph.n || t.count < t.parts)
                throw Error(GPART_ERR_INVALID_ARGUMENT, "queue root task out of range");
            // a task of p parts has at most p - 1 tasks in its subtree (itself included)
            task_capacity += t.parts - 1;
            // members serving the subtree at once: at most one team per part pair
            // plus one member per own_blocks blocks of its vertices
            concurrent += t.parts / 2 + (t.count + 8 * own_blocks - 1) / (8 * own_blocks);
        }
        max_team = (kMaxBatchedTask / 8 + own_blocks - 1) / own_blocks;
        slot_capacity = task_capacity * max_team;

        // scratch: per position two vectors and two partial arrays, the position
        // map, and one local neighbour per row entry
        const std::size_t value_bytes = precision == GPART_SPECTRAL_FP32 ? sizeof(float) : sizeof(double);
        Scratch<unsigned char> values(ctx, static_cast<std::int64_t>(2 * total_n * value_bytes));
        Scratch<Moments> partials(ctx, 2 * total_n);
        Scratch<int> pos(ctx, total_n), local(ctx, std::max<std::int64_t>(1, total_m));
        std::vector<QueueGraph> qg(static_cast<std::size_t>(graph_count));
        std::int64_t n_at = 0, m_at = 0;
        for (int i = 0; i < graph_count; ++i) {
            const auto& q = graphs[i];
            QueueGraph& d = qg[static_cast<std::size_t>(i)];
            d.offsets = q.graph.offsets;
            d.neighbors = q.graph.neighbors;
            d.edge_weights = q.graph.edge_weights;
            d.labels = q.labels;
            d.order = q.order;
            d.vertex_weights = q.vertex_weights;
            d.coords = q.coords;
            d.y[0] = values.get() + n_at * value_bytes;
            d.y[1] = values.get() + (total_n + n_at) * value_bytes;
            d.partials[0] = partials.get() + n_at;
            d.partials[1] = partials.get() + total_n + n_at;
            d.pos = pos.get() + n_at;
            d.local = local.get() + m_at;
            d.capacity = q.capacity;
            d.seed = q.seed;
            n_at += q.graph.n;
            m_at += q.graph.m;
        }
        // one upload: graphs, root records, root slots, state, per-graph status
        std::vector<QueueTask> records(static_cast<std::size_t>(root_count));
        std::vector<unsigned long long> root_slots;
        for (int r = 0; r < root_count; ++r) {
            QueueTask& rec = records[static_cast<std::size_t>(r)];
            rec.task = roots[r];
            rec.graph = root_graph[r];
            const int blocks = (roots[r].count + 7) / 8;
            rec.team = std::max(1, (blocks + own_blocks - 1) / own_blocks);
            rec.arrive = 0;
            rec.release = 0;
            for (int m = 0; m < rec.team; ++m)
                root_slots.push_back(kSlotValid | (static_cast<unsigned long long>(r) << kMemberBits) |
                                     static_cast<unsigned long long>(m));
        }
        QueueState state{};
        state.tail = root_slots.size();
        state.tasks = static_cast<unsigned long long>(root_count);
        state.outstanding = static_cast<unsigned long long>(root_count);
        auto align = [](std::size_t v) { return (v + 255) / 256 * 256; };
        const std::size_t graphs_at = 0;
        const std::size_t tasks_at = align(graphs_at + qg.size() * sizeof(QueueGraph));
        const std::size_t slots_at = align(tasks_at + static_cast<std::size_t>(task_capacity) * sizeof(QueueTask));
        const std::size_t state_at = align(slots_at + static_cast<std::size_t>(slot_capacity) * 8);
        const std::size_t status_at = align(state_at + sizeof(QueueState));
        const std::size_t total = align(status_at + static_cast<std::size_t>(graph_count) * sizeof(int));
        std::vector<unsigned char> host(total, 0);
        std::memcpy(host.data() + graphs_at, qg.data(), qg.size() * sizeof(QueueGraph));
        std::memcpy(host.data() + tasks_at, records.data(), records.size() * sizeof(QueueTask));
        std::memcpy(host.data() + slots_at, root_slots.data(), root_slots.size() * 8);
        std::memcpy(host.data() + state_at, &state, sizeof(state));
        Scratch<unsigned char> meta(ctx, static_cast<std::int64_t>(total));
        GPART_HIP_CHECK(hipMemcpyAsync(meta.get(), host.data(), total, hipMemcpyHostToDevice, ctx->stream));
        QueueArgs args{};
        args.graphs = reinterpret_cast<const QueueGraph*>(meta.get() + graphs_at);
        args.tasks = reinterpret_cast<QueueTask*>(meta.get() + tasks_at);
        args.slots = reinterpret_cast<unsigned long long*>(meta.get() + slots_at);
        args.state = reinterpret_cast<QueueState*>(meta.get() + state_at);
        args.status = reinterpret_cast<int*>(meta.get() + status_at);
        args.slot_capacity = slot_capacity;
        args.task_capacity = task_capacity;
        args.own_blocks = own_blocks;

        auto launch = [&](auto o_tag, auto w_tag, auto t_tag) {
            using O = decltype(o_tag);
            using W = decltype(w_tag);
            using T = decltype(t_tag);
            const void* kernel = reinterpret_cast<const void*>(&spectral_queue_kernel<O, W, T>);
            // every workgroup of the grid can be resident at once; more than the
            // queue's slots would only poll
            const long long grid = std::max<long long>(
                max_team, std::min<long long>({queue_grid(kernel, ctx->device), slot_capacity, concurrent}));
            hipLaunchKernelGGL((spectral_queue_kernel<O, W, T>), dim3(static_cast<unsigned>(grid)),
                               dim3(kBatchThreads), 0, ctx->stream, args);
            GPART_LAUNCH_CHECK();
            if (report) report->workgroups = static_cast<int32_t>(grid);
        };
        profile_begin(ctx, GPART_PROFILE_SPECTRAL);
        const bool fp32 = precision == GPART_SPECTRAL_FP32;
        if (width == GPART_WIDTH_STANDARD) {
            if (fp32) launch(std::int32_t{}, std::uint32_t{}, float{});
            else launch(std::int32_t{}, std::uint32_t{}, double{});
        } else {
            if (fp32) launch(std::int64_t{}, std::uint64_t{}, float{});
            else launch(std::int64_t{}, std::uint64_t{}, double{});
        }
        profile_end(ctx, GPART_PROFILE_SPECTRAL, 0.0);
        QueueState done{};
        read_back(ctx, &done, args.state, sizeof(done));
        std::vector<int> status(static_cast<std::size_t>(graph_count));
        read_back(ctx, status.data(), args.status, status.size() * sizeof(int));
        if (done.overflow) throw Error(GPART_ERR_RUNTIME, "spectral queue capacity exceeded");
        if (done.abort) throw Error(GPART_ERR_RUNTIME, "spectral queue wait exceeded its spin limit");
        for (int i = 0; i < graph_count; ++i) graphs[i].status = status[static_cast<std::size_t>(i)];
        if (report) {
            report->tasks = static_cast<int64_t>(done.served);
            report->slots = static_cast<int64_t>(done.tail);
            report->entries = static_cast<int64_t>(done.entries);
            report->vertices = static_cast<int64_t>(done.vertices);
        }
        // the reference's bytes of 64 diffusion steps over every served task
        const std::size_t o_size = width == GPART_WIDTH_STANDARD ? 4 : 8;
        profile_add_bytes(ctx, GPART_PROFILE_SPECTRAL,
                          64.0 * step_bytes(static_cast<double>(done.vertices), static_cast<int>(done.served),
                                            static_cast<double>(done.entries), o_size, o_size, value_bytes));
    });
}

}  // extern "C"
