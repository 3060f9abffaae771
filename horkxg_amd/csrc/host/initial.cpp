// initial.cpp -- host driver of the recursive spectral bisection
// (reference: src/initial_partition_gpu.cu:170-276).
//
// PlainFp64 keeps the reference's host task stack and per-task launch sequence
// (split chosen on the host from the device prefix, as the reference does).
// The batched modes run every task of one recursion depth in one launch; tasks
// at one depth touch disjoint vertices and label values, so the processing
// order does not change the result (src/initial_partition_gpu.cu:198-255).
#include "log.hpp"
#include "initial_partition.hpp"
#include "partition_capacity.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <vector>

namespace gpart {
namespace {

using U64 = std::uint64_t;

// Split selection of src/initial_partition_gpu.cu:233-248 from a host copy of
// the sorted segment's inclusive weight prefix. Returns the boundary.
int choose_split(const std::vector<U64>& prefix, const gpart_bisect_task& task, U64 capacity, U64& left_weight) {
    const int count = task.count;
    const int left_parts = task.parts / 2, right_parts = task.parts - left_parts;
    const U64 target = static_cast<U64>(static_cast<long double>(task.weight) * left_parts / task.parts);
    const U64 left_limit = capacity > task.weight / left_parts ? task.weight : capacity * left_parts;
    const U64 right_limit = capacity > task.weight / right_parts ? task.weight : capacity * right_parts;
    const U64 min_left = task.weight > right_limit ? task.weight - right_limit : 0;
    int boundary = static_cast<int>(std::lower_bound(prefix.begin(), prefix.begin() + count, target) - prefix.begin()) + 1;
    boundary = std::max(left_parts, std::min(count - right_parts, boundary));
    left_weight = prefix[boundary - 1];
    if (left_weight > left_limit && boundary > left_parts) {
        --boundary;
        left_weight = prefix[boundary - 1];
    }
    if (left_weight < min_left && boundary < count - right_parts) {
        ++boundary;
        left_weight = prefix[boundary - 1];
    }
    if (left_weight > left_limit || left_weight < min_left)
        throw std::runtime_error("GPU spectral split cannot satisfy capacity");
    return boundary;
}

template <typename Types>
void plain_task(const DeviceWeightedGraph<Types>& graph, DeviceVector<std::int32_t>& labels,
                DeviceVector<std::int32_t>& order, const DeviceVector<U64>& vertex_weights, const gpart_bisect_task& task,
                std::uint32_t seed, U64 capacity, DeviceVector<double>& x, DeviceVector<double>& y,
                DeviceVector<double>& degree, DeviceVector<U64>& prefix, std::vector<gpart_bisect_task>& children) {
    auto* ctx = detail::context();
    const auto view = graph.view();
    GPART_CHECK(gpart_spectral_task(ctx, &view, labels.data(), order.data(), vertex_weights.data(), &task, seed,
                                    x.data(), y.data(), degree.data(), prefix.data()));
    std::vector<U64> host(static_cast<std::size_t>(task.count));
    GPART_CHECK(gpart_memcpy_dtoh(ctx, host.data(), prefix.data(), host.size() * sizeof(U64)));
    U64 left_weight = 0;
    const int boundary = choose_split(host, task, capacity, left_weight);
    const int left_parts = task.parts / 2, right_parts = task.parts - left_parts;
    GPART_CHECK(gpart_spectral_assign(ctx, order.data(), &task, boundary, left_parts, labels.data()));
    children.push_back({task.begin + boundary, task.count - boundary, task.first_part + left_parts, right_parts,
                        task.weight - left_weight});
    children.push_back({task.begin, boundary, task.first_part, left_parts, left_weight});
}

// One graph's recursive bisection up to the persistent launch: validation, the
// device arrays, and every task the queue does not take (queue_max: the largest
// queue task, 0 for none). The tasks left for the queue are in `roots`.
template <typename Types>
struct Bisection {
    const DeviceWeightedGraph<Types>* graph = nullptr;
    int n = 0, parts = 0;
    SpectralMode mode = SpectralMode::BatchedFp64;
    std::uint32_t seed = 0;
    U64 total = 0, capacity = 0;
    DeviceVector<std::int32_t> labels, order;
    DeviceVector<U64> vertex_weights;
    DeviceVector<double> x, y, degree;
    DeviceVector<U64> prefix;
    std::vector<gpart_bisect_task> roots;
    std::chrono::steady_clock::time_point start;
};

template <typename Types>
void bisect_outside_queue(Bisection<Types>& s, const DeviceWeightedGraph<Types>& graph, int parts, double imbalance,
                          std::uint32_t seed, SpectralMode mode, int queue_max) {
    s.start = std::chrono::steady_clock::now();
    auto* ctx = detail::context();
    const auto n64 = graph.vertices();
    if (n64 <= 0 || n64 > std::numeric_limits<int>::max() - 255 || parts < 2 || parts > n64 || parts > 32 ||
        !std::isfinite(imbalance) || imbalance < 1.0)
        throw std::invalid_argument("invalid GPU initial partition input");
    const int n = static_cast<int>(n64);
    s.graph = &graph;
    s.n = n;
    s.parts = parts;
    s.mode = mode;
    s.seed = seed;
    const auto view = graph.view();
    U64 total = 0, max_weight = 0, min_weight = 0, edge_total = 0, edge_limited = 0;
    GPART_CHECK(gpart_vertex_weight_stats(ctx, &view, &total, &max_weight, &min_weight));
    GPART_CHECK(gpart_edge_weight_sum(ctx, &view, &edge_total, &edge_limited));
    if (!total || total == ~U64{0} || edge_total == ~U64{0})
        throw std::overflow_error("GPU initializer weight accumulation overflow");
    const U64 capacity = partition_capacity(total, parts, imbalance);
    if (max_weight > capacity) throw std::invalid_argument("a vertex exceeds partition capacity");
    s.total = total;
    s.capacity = capacity;

    auto& labels = s.labels;
    auto& order = s.order;
    auto& vertex_weights = s.vertex_weights;
    labels = DeviceVector<std::int32_t>(static_cast<std::size_t>(n), 0);
    order = DeviceVector<std::int32_t>(static_cast<std::size_t>(n));
    GPART_CHECK(gpart_sequence_i32(ctx, order.data(), n));
    vertex_weights = DeviceVector<U64>(static_cast<std::size_t>(n));
    GPART_CHECK(gpart_widen_weights(ctx, graph.vertex_weights.data(), Types::kWidth, n, vertex_weights.data()));
    s.x = DeviceVector<double>(static_cast<std::size_t>(n));
    s.y = DeviceVector<double>(static_cast<std::size_t>(n));
    s.degree = DeviceVector<double>(static_cast<std::size_t>(n));
    s.prefix = DeviceVector<U64>(static_cast<std::size_t>(n));
    auto& x = s.x;
    auto& y = s.y;
    auto& degree = s.degree;
    auto& prefix = s.prefix;

    if (mode == SpectralMode::PlainFp64) {
        std::vector<gpart_bisect_task> tasks{{0, n, 0, parts, total}};
        while (!tasks.empty()) {
            const auto task = tasks.back();
            tasks.pop_back();
            if (task.parts == 1) {
                if (task.weight > capacity) throw std::runtime_error("GPU initializer infeasible balance");
                continue;
            }
            plain_task(graph, labels, order, vertex_weights, task, seed, capacity, x, y, degree, prefix, tasks);
        }
        return;
    }
    const int precision = mode == SpectralMode::BatchedFp32 ? GPART_SPECTRAL_FP32 : GPART_SPECTRAL_FP64;
    const auto max_task = gpart_spectral_depth_max_task();
    std::vector<gpart_bisect_task> depth{{0, n, 0, parts, total}};
    DeviceVector<gpart_bisect_task> device_next(64);
    while (!depth.empty()) {
        std::vector<gpart_bisect_task> batched, next;
        for (const auto& t : depth) {
            if (t.parts == 1) {
                if (t.weight > capacity) throw std::runtime_error("GPU initializer infeasible balance");
            } else if (t.count > max_task) {
                std::vector<gpart_bisect_task> children;
                plain_task(graph, labels, order, vertex_weights, t, seed, capacity, x, y, degree, prefix, children);
                next.insert(next.end(), children.begin(), children.end());
            } else if (t.count <= queue_max) {
                s.roots.push_back(t);
            } else {
                batched.push_back(t);
            }
        }
        if (!batched.empty()) {
            std::int32_t produced = 0, status = 0;
            GPART_CHECK(gpart_spectral_depth(ctx, &view, labels.data(), order.data(), vertex_weights.data(),
                                             batched.data(), static_cast<std::int32_t>(batched.size()), seed,
                                             capacity, precision, x.data(), device_next.data(), &produced, &status));
            if (status == 1) throw std::runtime_error("GPU spectral split cannot satisfy capacity");
            if (status == 2) throw std::runtime_error("GPU initializer infeasible balance");
            std::vector<gpart_bisect_task> produced_tasks(static_cast<std::size_t>(produced));
            if (produced)
                GPART_CHECK(gpart_memcpy_dtoh(ctx, produced_tasks.data(), device_next.data(),
                                              produced_tasks.size() * sizeof(gpart_bisect_task)));
            next.insert(next.end(), produced_tasks.begin(), produced_tasks.end());
        }
        depth.swap(next);
    }
}

// One persistent launch over the roots of every graph of the group.
template <typename Types>
void run_queue(std::vector<Bisection<Types>>& group) {
    std::vector<gpart_queue_graph> graphs;
    std::vector<gpart_bisect_task> roots;
    std::vector<std::int32_t> root_graph;
    std::vector<std::size_t> member;
    for (std::size_t i = 0; i < group.size(); ++i) {
        auto& s = group[i];
        if (s.roots.empty()) continue;
        gpart_queue_graph q{};
        q.graph = s.graph->view();
        q.labels = s.labels.data();
        q.order = s.order.data();
        q.vertex_weights = s.vertex_weights.data();
        q.coords = s.x.data();
        q.capacity = s.capacity;
        q.seed = s.seed;
        for (const auto& t : s.roots) {
            roots.push_back(t);
            root_graph.push_back(static_cast<std::int32_t>(graphs.size()));
        }
        graphs.push_back(q);
        member.push_back(i);
    }
    if (roots.empty()) return;
    const int precision = group[member[0]].mode == SpectralMode::BatchedFp32 ? GPART_SPECTRAL_FP32 : GPART_SPECTRAL_FP64;
    for (std::size_t j = 0; j < member.size(); ++j)
        if ((group[member[j]].mode == SpectralMode::BatchedFp32 ? GPART_SPECTRAL_FP32 : GPART_SPECTRAL_FP64) != precision)
            throw std::invalid_argument("one queue launch runs one precision");
    GPART_CHECK(gpart_spectral_queue(detail::context(), graphs.data(), static_cast<std::int32_t>(graphs.size()),
                                     roots.data(), root_graph.data(), static_cast<std::int32_t>(roots.size()), precision,
                                     nullptr));
    for (const auto& q : graphs) {
        if (q.status == 1) throw std::runtime_error("GPU spectral split cannot satisfy capacity");
        if (q.status == 2) throw std::runtime_error("GPU initializer infeasible balance");
    }
}

// Validation of the loads, the cut and the report lines.
template <typename Types>
DeviceVector<std::int32_t> finish_bisection(Bisection<Types>& s, DeviceVector<double>* coords) {
    auto* ctx = detail::context();
    const auto view = s.graph->view();
    // plain tasks leave their final coordinates in x; batched tasks write them there
    if (coords) *coords = s.x;
    DeviceVector<U64> loads(static_cast<std::size_t>(s.parts));
    GPART_CHECK(gpart_part_weights(ctx, &view, s.labels.data(), s.parts, loads.data()));
    const auto host_loads = loads.to_host();
    if (*std::max_element(host_loads.begin(), host_loads.end()) > s.capacity)
        throw std::runtime_error("GPU initializer balance validation failed");
    U64 cut = 0;
    GPART_CHECK(gpart_cut(ctx, &view, s.labels.data(), &cut));
    static const char* names[] = {"gpu_spectral_plain_fp64", "gpu_spectral_batched_fp64", "gpu_spectral_batched_fp32"};
    gpart::detail::log() << "initial_partition_backend=" << names[static_cast<int>(s.mode)] << '\n'
                         << "coarsest_vertices=" << s.n << '\n'
                         << "coarsest_edge_entries=" << s.graph->edges() << '\n'
                         << "initial_partition_parts=" << s.parts << '\n'
                         << "initial_partition_edgecut=" << cut << '\n'
                         << "initial_partition_seconds="
                         << std::chrono::duration<double>(std::chrono::steady_clock::now() - s.start).count() << '\n';
    return std::move(s.labels);
}

}  // namespace

template <typename Types>
DeviceVector<typename Types::VertexT> initial_partition_gpu(const DeviceWeightedGraph<Types>& graph, int parts,
                                                            double imbalance, std::uint32_t seed, SpectralMode mode,
                                                            DeviceVector<double>* coords) {
    // Tasks of at most queue_max vertices, of any depth, are roots of one
    // persistent launch; larger batched tasks run depth by depth.
    std::int32_t own_blocks = 0, queue_max = 0;
    gpart_spectral_queue_tuning(&own_blocks, &queue_max);
    if (!gpart_spectral_queue_enabled()) queue_max = 0;
    std::vector<Bisection<Types>> one(1);
    bisect_outside_queue(one[0], graph, parts, imbalance, seed, mode, queue_max);
    run_queue(one);
    return finish_bisection(one[0], coords);
}

template <typename Types>
std::vector<DeviceVector<typename Types::VertexT>> initial_partition_group(
    const std::vector<const DeviceWeightedGraph<Types>*>& graphs, const std::vector<int>& parts,
    const std::vector<double>& imbalance, const std::vector<std::uint32_t>& seeds, SpectralMode mode) {
    if (parts.size() != graphs.size() || imbalance.size() != graphs.size() || seeds.size() != graphs.size())
        throw std::invalid_argument("initial partition group arrays of different lengths");
    const int queue_max = mode == SpectralMode::PlainFp64 ? 0 : static_cast<int>(gpart_spectral_depth_max_task());
    std::vector<Bisection<Types>> group(graphs.size());
    for (std::size_t i = 0; i < graphs.size(); ++i)
        bisect_outside_queue(group[i], *graphs[i], parts[i], imbalance[i], seeds[i], mode, queue_max);
    run_queue(group);
    std::vector<DeviceVector<typename Types::VertexT>> out;
    out.reserve(graphs.size());
    for (auto& s : group) out.push_back(finish_bisection(s, nullptr));
    return out;
}

template <typename Types>
SpectralComparison compare_spectral_precision(const DeviceWeightedGraph<Types>& graph, int parts, double imbalance,
                                              std::uint32_t seed) {
    auto* ctx = detail::context();
    const auto n64 = graph.vertices();
    if (n64 <= 0 || n64 > std::numeric_limits<int>::max() - 255 || parts < 2 || parts > n64 || parts > 32 ||
        !std::isfinite(imbalance) || imbalance < 1.0)
        throw std::invalid_argument("invalid GPU initial partition input");
    const int n = static_cast<int>(n64);
    const auto view = graph.view();
    U64 total = 0, max_weight = 0, min_weight = 0;
    GPART_CHECK(gpart_vertex_weight_stats(ctx, &view, &total, &max_weight, &min_weight));
    const U64 capacity = partition_capacity(total, parts, imbalance);
    DeviceVector<std::int32_t> labels(static_cast<std::size_t>(n), 0), order(static_cast<std::size_t>(n));
    GPART_CHECK(gpart_sequence_i32(ctx, order.data(), n));
    DeviceVector<U64> vertex_weights(static_cast<std::size_t>(n));
    GPART_CHECK(gpart_widen_weights(ctx, graph.vertex_weights.data(), Types::kWidth, n, vertex_weights.data()));
    DeviceVector<double> x(static_cast<std::size_t>(n)), y(static_cast<std::size_t>(n)),
        degree(static_cast<std::size_t>(n)), x32(static_cast<std::size_t>(n));
    DeviceVector<U64> prefix(static_cast<std::size_t>(n));
    DeviceVector<gpart_bisect_task> next64(64), next32(64);
    const auto max_task = gpart_spectral_depth_max_task();

    auto run_depth = [&](DeviceVector<std::int32_t>& lab, DeviceVector<std::int32_t>& ord, int precision,
                         DeviceVector<double>& coords, DeviceVector<gpart_bisect_task>& next,
                         const std::vector<gpart_bisect_task>& tasks) {
        std::int32_t produced = 0, status = 0;
        GPART_CHECK(gpart_spectral_depth(ctx, &view, lab.data(), ord.data(), vertex_weights.data(), tasks.data(),
                                         static_cast<std::int32_t>(tasks.size()), seed, capacity, precision,
                                         coords.data(), next.data(), &produced, &status));
        if (status == 1) throw std::runtime_error("GPU spectral split cannot satisfy capacity");
        if (status == 2) throw std::runtime_error("GPU initializer infeasible balance");
        std::vector<gpart_bisect_task> out(static_cast<std::size_t>(produced));
        if (produced)
            GPART_CHECK(gpart_memcpy_dtoh(ctx, out.data(), next.data(), out.size() * sizeof(gpart_bisect_task)));
        return out;
    };

    SpectralComparison result;
    std::vector<gpart_bisect_task> depth{{0, n, 0, parts, total}};
    for (int d = 0; !depth.empty(); ++d) {
        std::vector<gpart_bisect_task> batched, next;
        for (const auto& t : depth) {
            if (t.parts == 1) continue;
            if (t.count > max_task) {
                std::vector<gpart_bisect_task> children;
                plain_task(graph, labels, order, vertex_weights, t, seed, capacity, x, y, degree, prefix, children);
                next.insert(next.end(), children.begin(), children.end());
            } else {
                batched.push_back(t);
            }
        }
        if (!batched.empty()) {
            DeviceVector<std::int32_t> labels32(labels), order32(order);
            const auto children32 = run_depth(labels32, order32, GPART_SPECTRAL_FP32, x32, next32, batched);
            const auto children64 = run_depth(labels, order, GPART_SPECTRAL_FP64, x, next64, batched);
            const auto h64 = x.to_host(), h32 = x32.to_host();
            const auto l64 = labels.to_host(), l32 = labels32.to_host();
            const auto ord = order.to_host();
            SpectralDepthError e;
            e.depth = d;
            e.tasks = static_cast<int>(batched.size());
            double squares = 0.0;
            for (const auto& t : batched) {
                for (int i = t.begin; i < t.begin + t.count; ++i) {
                    const int v = ord[static_cast<std::size_t>(i)];
                    const double diff = std::fabs(h32[v] - h64[v]);
                    e.max_abs_error = std::max(e.max_abs_error, diff);
                    squares += diff * diff;
                    e.labels_differ += l32[v] != l64[v];
                }
                e.vertices += t.count;
                // the left child starts at the task's segment and keeps its first part
                auto left = [&](const std::vector<gpart_bisect_task>& kids) {
                    for (const auto& c : kids)
                        if (c.begin == t.begin && c.first_part == t.first_part) return c.count;
                    return -1;
                };
                e.splits_differ += left(children32) != left(children64);
            }
            e.rms_error = e.vertices ? std::sqrt(squares / static_cast<double>(e.vertices)) : 0.0;
            result.depths.push_back(e);
            next.insert(next.end(), children64.begin(), children64.end());
        }
        depth.swap(next);
    }
    auto measure = [&](SpectralMode mode, U64& cut, U64& max_load) {
        auto lab = initial_partition_gpu<Types>(graph, parts, imbalance, seed, mode);
        GPART_CHECK(gpart_cut(ctx, &view, lab.data(), &cut));
        DeviceVector<U64> loads(static_cast<std::size_t>(parts));
        GPART_CHECK(gpart_part_weights(ctx, &view, lab.data(), parts, loads.data()));
        const auto h = loads.to_host();
        max_load = *std::max_element(h.begin(), h.end());
        return lab.to_host();
    };
    const auto final64 = measure(SpectralMode::BatchedFp64, result.cut_fp64, result.max_load_fp64);
    const auto final32 = measure(SpectralMode::BatchedFp32, result.cut_fp32, result.max_load_fp32);
    for (std::size_t v = 0; v < final64.size(); ++v) result.labels_differ += final64[v] != final32[v];
    return result;
}

template SpectralComparison compare_spectral_precision<StandardTypes>(const DeviceWeightedGraph<StandardTypes>&, int,
                                                                      double, std::uint32_t);
template SpectralComparison compare_spectral_precision<BigTypes>(const DeviceWeightedGraph<BigTypes>&, int, double,
                                                                 std::uint32_t);

template std::vector<DeviceVector<std::int32_t>> initial_partition_group<StandardTypes>(
    const std::vector<const DeviceWeightedGraph<StandardTypes>*>&, const std::vector<int>&, const std::vector<double>&,
    const std::vector<std::uint32_t>&, SpectralMode);
template std::vector<DeviceVector<std::int32_t>> initial_partition_group<BigTypes>(
    const std::vector<const DeviceWeightedGraph<BigTypes>*>&, const std::vector<int>&, const std::vector<double>&,
    const std::vector<std::uint32_t>&, SpectralMode);

template DeviceVector<std::int32_t> initial_partition_gpu<StandardTypes>(const DeviceWeightedGraph<StandardTypes>&, int,
                                                                         double, std::uint32_t, SpectralMode,
                                                                         DeviceVector<double>*);
template DeviceVector<std::int32_t> initial_partition_gpu<BigTypes>(const DeviceWeightedGraph<BigTypes>&, int, double,
                                                                    std::uint32_t, SpectralMode, DeviceVector<double>*);

}  // namespace gpart
