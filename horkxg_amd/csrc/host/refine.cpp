// refine.cpp -- host scheduling of incremental bucketed refinement
// (reference: src/refine_incremental.cuh:235-361 run_incremental_refinement,
// src/refine.cu:162-660 contexts, gain-range validation, refine_partition and
// refine_hierarchy_impl).
#include "log.hpp"
#include "refine.hpp"

#include "partition_capacity.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace gpart {
namespace {

using clock_type = std::chrono::steady_clock;
double since(clock_type::time_point t) { return std::chrono::duration<double>(clock_type::now() - t).count(); }

void check_options(const RefineOptions& o) {
    if (o.parts < 2 || o.parts > 32 || o.max_rounds < 0 || o.max_rounds > 64 || o.stale_rounds < 1 ||
        o.stale_rounds > 64 || !std::isfinite(o.refinement_growth_ratio) || o.refinement_growth_ratio < 1.0 ||
        !std::isfinite(o.imbalance_ratio) || o.imbalance_ratio < 1.0)
        throw std::invalid_argument("invalid refinement options");
}

// validate_refinement_gain_range (src/refine.cu:374-383)
template <typename Types>
void validate_gain_range(const DeviceWeightedGraph<Types>& g) {
    constexpr auto limit = static_cast<std::uint64_t>(std::numeric_limits<std::int64_t>::max()) / 2;
    constexpr auto max_weight = static_cast<std::uint64_t>(std::numeric_limits<typename Types::WeightT>::max());
    if (static_cast<std::uint64_t>(g.edges()) <= limit / max_weight) return;
    const auto view = g.view();
    std::uint64_t checked = 0, limited = 0;
    GPART_CHECK(gpart_edge_weight_sum(detail::context(), &view, &checked, &limited));
    if (limited > limit) throw std::overflow_error("refinement signed gain range exceeded");
}

// host overload (src/refine.cu:346-365)
template <typename Types>
void validate_gain_range_host(const WeightedGraph<Types>& g) {
    constexpr auto limit = static_cast<std::uint64_t>(std::numeric_limits<std::int64_t>::max()) / 2;
    constexpr auto max_weight = static_cast<std::uint64_t>(std::numeric_limits<typename Types::WeightT>::max());
    if (g.edge_weights.size() <= limit / max_weight) return;
    std::uint64_t directed = 0;
    for (const auto w : g.edge_weights) {
        const auto x = static_cast<std::uint64_t>(w);
        if (x > limit || directed > limit - x) throw std::overflow_error("refinement signed gain range exceeded");
        directed += x;
    }
}

template <typename Types>
struct Level {
    DeviceWeightedGraph<Types> graph;
    DeviceVector<std::int32_t> partition;
    DeviceVector<std::uint64_t> part_weights;
    std::vector<std::uint64_t> host_part_weights;
    std::uint64_t capacity = 0;

    std::uint64_t cut() const {
        const auto view = graph.view();
        std::uint64_t c = 0;
        GPART_CHECK(gpart_cut(detail::context(), &view, partition.data(), &c));
        return c;
    }
    std::uint64_t maximum() const { return *std::max_element(host_part_weights.begin(), host_part_weights.end()); }
};

// run_incremental_refinement (src/refine_incremental.cuh:235-361)
template <typename Types>
// known_cut: the cut of the level's labels when the caller knows it exactly (a
// projection keeps the coarse level's cut), saving a scan of the CSR.
RefineStats run_refinement(Level<Types>& level, const RefineOptions& options, int rounds,
                           const std::uint64_t* known_cut = nullptr) {
    const auto start = clock_type::now();
    auto* ctx = detail::context();
    validate_gain_range(level.graph);
    const auto n = level.graph.vertices();
    const int k = options.parts;
    const auto view = level.graph.view();
    const auto N = static_cast<std::size_t>(n);

    DeviceVector<std::int32_t> low(N), medium(N), high(N);
    gpart_degree_classes classes{low.data(), medium.data(), high.data(), 0, 0, 0, 0, 0, 0};
    GPART_CHECK(gpart_degree_classes_build(ctx, &view, &classes));
    // Cache entries are 4 bytes when the level's total edge weight, a bound on
    // every row sum, fits 32 bits; otherwise 8.
    std::uint64_t maximum_edge = 0;
    GPART_CHECK(gpart_max_edge_weight(ctx, &view, &maximum_edge));
    int cache_bytes = 8;
    const auto m = static_cast<std::uint64_t>(level.graph.edges());
    if (m == 0 || maximum_edge <= 0xffffffffULL / m) {
        cache_bytes = 4;
    } else {
        std::uint64_t checked = 0, limited = 0;
        GPART_CHECK(gpart_edge_weight_sum(ctx, &view, &checked, &limited));
        if (checked <= 0xffffffffULL) cache_bytes = 4;
    }
    DeviceVector<std::uint8_t> cache(N * static_cast<std::size_t>(k) * static_cast<std::size_t>(cache_bytes));
    DeviceVector<std::int32_t> targets(N), locks(N, 0), flags(N, 0);
    DeviceVector<std::int64_t> gains(N);
    GPART_CHECK(gpart_refine_build_cache(ctx, &view, &classes, level.partition.data(), k, cache.data(), cache_bytes));

    RefineStats stats;
    stats.initial_cut = known_cut ? *known_cut : level.cut();
    const double filter = maximum_edge <= 1 ? 0.25 : 0.75;
    auto finish = [&](std::uint64_t best_cut) {
        level.host_part_weights = level.part_weights.to_host();
        stats.final_cut = level.cut();
        stats.final_max_part_weight = level.maximum();
        if (stats.final_cut != best_cut || stats.final_cut > stats.initial_cut || level.maximum() > level.capacity)
            throw std::runtime_error("incremental refinement postcondition failed");
        stats.seconds = since(start);
        return stats;
    };
    // Levels within the engine's tuning run the whole loop below in one launch.
    int grid = 0;
    if (!options.strict_verify)
        GPART_CHECK(gpart_refine_engine_grid(ctx, &view, &classes, cache_bytes, &grid));
    if (grid > 0) {
        gpart_refine_level_stats e{};
        GPART_CHECK(gpart_refine_level(ctx, &view, &classes, cache.data(), cache_bytes, level.partition.data(), k,
                                       level.part_weights.data(), level.capacity, filter, rounds,
                                       options.stale_rounds, stats.initial_cut, targets.data(), gains.data(),
                                       flags.data(), locks.data(), grid, &e));
        stats.rounds = e.rounds;
        stats.proposals = e.proposals;
        stats.accepted = e.accepted;
        stats.admission_iterations = e.admission_iterations;
        stats.admission_fallbacks = e.admission_fallbacks;
        stats.mover_entries = e.mover_entries;
        return finish(e.best_cut);
    }
    auto current_cut = stats.initial_cut, best_cut = current_cut;
    DeviceVector<std::int32_t> best_partition(level.partition);
    DeviceVector<std::uint64_t> best_part_weights(level.part_weights);
    auto checkpoint = [&]() {
        GPART_CHECK(gpart_memcpy_dtod(ctx, best_partition.data(), level.partition.data(), N * sizeof(std::int32_t)));
        GPART_CHECK(gpart_memcpy_dtod(ctx, best_part_weights.data(), level.part_weights.data(), k * sizeof(std::uint64_t)));
    };
    // Book-keeping of an applied round (the reference's apply, :280-301).
    auto absorb = [&](const gpart_refine_status& status) {
        const std::int64_t delta = status.cut_delta;
        if (status.rebuilt) {
            current_cut = status.cut_after;
        } else if (delta >= 0) {
            if (static_cast<std::uint64_t>(delta) > current_cut) throw std::runtime_error("incremental cut underflow");
            current_cut -= static_cast<std::uint64_t>(delta);
        } else {
            current_cut += static_cast<std::uint64_t>(-delta);
        }
        stats.accepted += static_cast<std::uint64_t>(status.movers);
        stats.mover_entries += static_cast<std::uint64_t>(status.mover_entries);
        stats.cache_rebuilds += status.rebuilt;
        level.host_part_weights.assign(status.part_weights, status.part_weights + k);
        if (options.strict_verify) {
            if (level.cut() != current_cut) throw std::runtime_error("incremental cut differs from CSR scan");
            DeviceVector<std::uint8_t> check(cache.size());
            GPART_CHECK(gpart_refine_build_cache(ctx, &view, &classes, level.partition.data(), k, check.data(),
                                                 cache_bytes));
            if (check.to_host() != cache.to_host())
                throw std::runtime_error("incremental affinity cache differs from CSR rebuild");
        }
    };
    auto rebalance = [&]() {
        for (int iteration = 0; iteration < 2 * k + 4 && level.maximum() > level.capacity; ++iteration) {
            ++stats.admission_iterations;
            gpart_refine_status status{};
            GPART_CHECK(gpart_refine_rebalance_round(ctx, &view, &classes, cache.data(), cache_bytes,
                                                     level.partition.data(), k,
                                                     level.part_weights.data(), level.capacity, targets.data(),
                                                     flags.data(), &status));
            if (!status.movers) return false;
            absorb(status);
        }
        return level.maximum() <= level.capacity;
    };
    int stale = 0, empty = 0;
    for (int phase = 0; phase < rounds && stale < options.stale_rounds; ++phase) {
        stats.rounds = phase + 1;
        gpart_refine_status status{};
        GPART_CHECK(gpart_refine_lp_round(ctx, &view, &classes, cache.data(), cache_bytes, level.partition.data(),
                                          k, filter,
                                          locks.data(), targets.data(), gains.data(), flags.data(),
                                          level.part_weights.data(), &status));
        stats.proposals += static_cast<std::uint64_t>(status.movers);
        if (!status.movers) {
            if (++empty >= 2) break;
            ++stale;
            continue;
        }
        empty = 0;
        absorb(status);
        if (!rebalance()) {
            ++stats.admission_fallbacks;
            break;
        }
        ++stale;
        if (current_cut < best_cut) {
            if (static_cast<long double>(current_cut) < 0.999L * best_cut) stale = 0;
            best_cut = current_cut;
            checkpoint();
        }
    }
    GPART_CHECK(gpart_memcpy_dtod(ctx, level.partition.data(), best_partition.data(), N * sizeof(std::int32_t)));
    GPART_CHECK(gpart_memcpy_dtod(ctx, level.part_weights.data(), best_part_weights.data(), k * sizeof(std::uint64_t)));
    return finish(best_cut);
}

template <typename Types>
std::vector<std::uint64_t> measure(const WeightedGraph<Types>& graph, const std::vector<typename Types::VertexT>& labels,
                                   int parts, std::uint64_t& total) {
    if (labels.size() != static_cast<std::size_t>(graph.vertices()))
        throw std::invalid_argument("refinement partition has the wrong length");
    std::vector<std::uint64_t> w(static_cast<std::size_t>(parts), 0);
    total = 0;
    for (std::size_t v = 0; v < labels.size(); ++v) {
        const auto p = labels[v];
        if (p < 0 || p >= parts) throw std::invalid_argument("refinement partition has an invalid part id");
        const auto x = static_cast<std::uint64_t>(graph.vertex_weights[v]);
        if (w[p] > std::numeric_limits<std::uint64_t>::max() - x ||
            total > std::numeric_limits<std::uint64_t>::max() - x)
            throw std::overflow_error("refinement vertex weight sum overflow");
        w[p] += x;
        total += x;
    }
    return w;
}

template <typename Types, bool Resident, typename H, typename P>
DeviceUncoarsenResult<Types> refine_hierarchy_impl(H& hierarchy, P&& coarsest_partition, const RefineOptions& options) {
    check_options(options);
    auto* ctx = detail::context();
    if (hierarchy.levels.empty() || hierarchy.fine_to_coarse.size() + 1 != hierarchy.levels.size())
        throw std::invalid_argument("resident refinement requires a complete hierarchy");
    const int k = options.parts;
    Level<Types> current;
    std::uint64_t total_weight = 0;
    if constexpr (Resident) {
        for (const auto& l : hierarchy.levels) validate_gain_range(l);
        auto& coarsest = hierarchy.levels.back();
        if (coarsest_partition.size() != static_cast<std::size_t>(coarsest.vertices()))
            throw std::invalid_argument("invalid coarsest device partition");
        const auto labels = coarsest_partition.to_host();
        for (const auto p : labels)
            if (p < 0 || p >= k) throw std::invalid_argument("invalid coarsest device partition");
        current.graph = std::move(coarsest);
        current.partition = std::move(coarsest_partition);
        current.part_weights.resize_discard(static_cast<std::size_t>(k));
        const auto view = current.graph.view();
        GPART_CHECK(gpart_part_weights(ctx, &view, current.partition.data(), k, current.part_weights.data()));
        current.host_part_weights = current.part_weights.to_host();
        for (const auto w : current.host_part_weights) {
            if (total_weight > std::numeric_limits<std::uint64_t>::max() - w)
                throw std::overflow_error("refinement vertex weight sum overflow");
            total_weight += w;
        }
    } else {
        const auto& coarsest = hierarchy.levels.back();
        current.host_part_weights = measure(coarsest, coarsest_partition, k, total_weight);
        current.graph = make_device_weighted(coarsest);
        current.partition.assign(coarsest_partition);
        current.part_weights.assign(current.host_part_weights);
        for (const auto& l : hierarchy.levels) validate_gain_range_host(l);
    }
    current.capacity = partition_capacity(total_weight, k, options.imbalance_ratio);
    if (current.maximum() > current.capacity) throw std::invalid_argument("initial refinement partition is imbalanced");

    DeviceUncoarsenResult<Types> result;
    const auto stats = run_refinement(current, options, options.max_rounds);
    std::uint64_t last_cut = stats.final_cut;
    auto last_refined = current.graph.vertices();
    gpart::detail::log() << "refine_coarsest vertices=" << current.graph.vertices() << " initial_cut=" << stats.initial_cut
              << " final_cut=" << stats.final_cut << " rounds=" << stats.rounds << " seconds=" << stats.seconds << '\n';
    for (std::size_t coarse_level = hierarchy.levels.size() - 1; coarse_level > 0; --coarse_level) {
        const auto fine_level = coarse_level - 1;
        RefineLevelResult report;
        report.level = fine_level;
        Level<Types> next;
        DeviceVector<std::int32_t> map;
        const auto h2d_start = clock_type::now();
        if constexpr (Resident) {
            next.graph = std::move(hierarchy.levels[fine_level]);
            map = std::move(hierarchy.fine_to_coarse[fine_level]);
        } else {
            next.graph = make_device_weighted(hierarchy.levels[fine_level]);
            map.assign(hierarchy.fine_to_coarse[fine_level]);
            report.timings.graph_h2d_seconds = since(h2d_start);
        }
        report.vertices = next.graph.vertices();
        report.edge_entries = next.graph.edges();
        if (map.size() != static_cast<std::size_t>(report.vertices))
            throw std::invalid_argument("fine-to-coarse map has wrong length");
        const auto projection_start = clock_type::now();
        next.partition.resize_discard(static_cast<std::size_t>(report.vertices));
        GPART_CHECK(gpart_project(ctx, report.vertices, map.data(), current.partition.data(), next.partition.data()));
        GPART_CHECK(gpart_ctx_synchronize(ctx));
        report.timings.projection_gpu_seconds = since(projection_start);
        next.capacity = current.capacity;
        next.host_part_weights = current.host_part_weights;
        if (options.strict_verify) {
            const auto t = clock_type::now();
            next.part_weights.resize_discard(static_cast<std::size_t>(k));
            const auto view = next.graph.view();
            GPART_CHECK(gpart_part_weights(ctx, &view, next.partition.data(), k, next.part_weights.data()));
            next.host_part_weights = next.part_weights.to_host();
            report.projection_cut = next.cut();
            if (current.cut() != report.projection_cut || current.host_part_weights != next.host_part_weights)
                throw std::runtime_error("GPU projection failed cut or part-weight preservation");
            report.timings.verification_seconds = since(t);
        } else {
            // projection preserves part weights
            next.part_weights = std::move(current.part_weights);
        }
        report.projection_max_part_weight = next.maximum();
        const auto projected_total =
            std::accumulate(next.host_part_weights.begin(), next.host_part_weights.end(), std::uint64_t{0});
        report.projection_imbalance = projected_total == 0 ? 0.0
                                                           : static_cast<double>(report.projection_max_part_weight) *
                                                                 k / static_cast<double>(projected_total);
        const bool refine_this = fine_level == 0 || static_cast<double>(report.vertices) >=
                                                        options.refinement_growth_ratio *
                                                            static_cast<double>(last_refined);
        if (refine_this) {
            // the projection keeps the cut (the strict checks above confirm it)
            report.refinement = run_refinement(next, options, options.max_rounds, &last_cut);
            last_refined = report.vertices;
        } else {
            report.refinement.initial_cut = report.refinement.final_cut = last_cut;
            report.refinement.final_max_part_weight = report.projection_max_part_weight;
        }
        if (!options.strict_verify) report.projection_cut = report.refinement.initial_cut;
        report.timings.refinement_seconds = report.refinement.seconds;
        last_cut = report.refinement.final_cut;
        current = std::move(next);
        result.levels.push_back(report);
    }
    result.final_cut = last_cut;
    result.final_part_weights = current.host_part_weights;
    {
        const auto view = current.graph.view();
        std::uint64_t checked = 0, limited = 0;
        GPART_CHECK(gpart_edge_weight_sum(ctx, &view, &checked, &limited));
        if (checked & 1ULL) throw std::runtime_error("symmetric graph has odd directed edge weight");
        result.total_edge_weight = checked / 2;
    }
    if constexpr (Resident) {
        result.device_partition = std::move(current.partition);
    } else {
        const auto t = clock_type::now();
        result.partition = current.partition.to_host();
        result.final_d2h_seconds = since(t);
    }
    return result;
}

}  // namespace

template <typename Types>
void refine_partition(const WeightedGraph<Types>& graph, std::vector<typename Types::VertexT>& partition,
                      const RefineOptions& options, RefineStats* output_stats) {
    // RefineDeviceContext checks (src/refine.cu:259-289)
    if (options.parts < 2 || options.parts > 32 || options.max_rounds < 0 || options.max_rounds > 64 ||
        options.stale_rounds < 1 || options.stale_rounds > 64 || !std::isfinite(options.refinement_growth_ratio) ||
        options.refinement_growth_ratio < 1.0 || !std::isfinite(options.imbalance_ratio) ||
        options.imbalance_ratio < 1.0 || graph.vertices() <= 0)
        throw std::invalid_argument("invalid refinement options");
    if (graph.offsets.size() != graph.vertex_weights.size() + 1 || graph.offsets.empty() || graph.offsets.front() != 0 ||
        static_cast<std::int64_t>(graph.offsets.back()) != graph.edges() ||
        graph.neighbors.size() != graph.edge_weights.size())
        throw std::invalid_argument("invalid weighted CSR for refinement");
    Level<Types> level;
    std::uint64_t total = 0;
    level.host_part_weights = measure(graph, partition, options.parts, total);
    level.capacity = partition_capacity(total, options.parts, options.imbalance_ratio);
    if (level.maximum() > level.capacity) throw std::invalid_argument("initial refinement partition is imbalanced");
    level.graph = make_device_weighted(graph);
    level.partition.assign(partition);
    level.part_weights.assign(level.host_part_weights);
    const auto stats = run_refinement(level, options, options.max_rounds);
    partition = level.partition.to_host();
    if (output_stats) *output_stats = stats;
}

template <typename Types>
DeviceUncoarsenResult<Types> refine_hierarchy_device(const Hierarchy<Types>& hierarchy,
                                                     const std::vector<typename Types::VertexT>& partition,
                                                     const RefineOptions& options) {
    return refine_hierarchy_impl<Types, false>(hierarchy, partition, options);
}

template <typename Types>
DeviceUncoarsenResult<Types> refine_hierarchy_device(DeviceHierarchy<Types>& hierarchy,
                                                     DeviceVector<typename Types::VertexT>&& partition,
                                                     const RefineOptions& options) {
    return refine_hierarchy_impl<Types, true>(hierarchy, std::move(partition), options);
}

#define GPART_INSTANTIATE_REFINE(T)                                                                              \
    template void refine_partition<T>(const WeightedGraph<T>&, std::vector<typename T::VertexT>&,                \
                                      const RefineOptions&, RefineStats*);                                      \
    template DeviceUncoarsenResult<T> refine_hierarchy_device<T>(                                               \
        const Hierarchy<T>&, const std::vector<typename T::VertexT>&, const RefineOptions&);                    \
    template DeviceUncoarsenResult<T> refine_hierarchy_device<T>(DeviceHierarchy<T>&,                           \
                                                                 DeviceVector<typename T::VertexT>&&,           \
                                                                 const RefineOptions&);
GPART_INSTANTIATE_REFINE(StandardTypes)
GPART_INSTANTIATE_REFINE(BigTypes)

}  // namespace gpart
