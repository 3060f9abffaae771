// coarsen.cpp -- host scheduling of SCLP coarsening over the kernel library's C ABI
// (reference: src/coarsen.cu:937-1659 aggregate_typed, :1674-1831 contract,
// :1835-1902 validate_coarsening_step, :1906-2174 coarsen_impl).
//
// The level loop, the LP-round loop and every stop rule are those of the
// reference; each step below is one kernel-family call (gpart_hip.h).
#include "log.hpp"
#include "coarsen.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace gpart {

template <typename Types>
struct SclpWorkspace<Types>::Impl {
    bool record_bound_known = false;
    bool sorted_unique_rows = false;
    // Vertex-weight statistics of the next graph to aggregate when it is the last
    // contraction's output: the total is the finer level's (contraction keeps it),
    // the maximum is the largest cluster weight, and every weight is a sum of
    // positive weights, so the graph needs no reduction pass for them.
    bool weight_stats_known = false;
    std::uint64_t known_total_weight = 0;
    std::uint64_t known_maximum_weight = 0;
    std::uint64_t maximum_row_weight = 0;
    DeviceVector<std::uint64_t> row_bounds;
    DeviceVector<std::uint64_t> next_row_bounds;
    // degree classes of the graph last aggregated (reused by its contraction)
    DeviceVector<std::int32_t> low, medium, high;
    gpart_degree_classes classes{};
    const void* classes_for = nullptr;
    std::int64_t classes_n = -1;

    const gpart_degree_classes& classes_of(const DeviceWeightedGraph<Types>& g) {
        const auto n = g.vertices();
        if (classes_for != g.offsets.data() || classes_n != n) {
            low.resize_discard(static_cast<std::size_t>(std::max<std::int64_t>(n, 1)));
            medium.resize_discard(static_cast<std::size_t>(std::max<std::int64_t>(n, 1)));
            high.resize_discard(static_cast<std::size_t>(std::max<std::int64_t>(n, 1)));
            classes = gpart_degree_classes{low.data(), medium.data(), high.data(), 0, 0, 0, 0, 0, 0};
            const auto view = g.view();
            GPART_CHECK(gpart_degree_classes_build(detail::context(), &view, &classes));
            classes_for = g.offsets.data();
            classes_n = n;
        }
        return classes;
    }
    void forget_classes() {
        classes_for = nullptr;
        classes_n = -1;
    }
};

template <typename Types>
SclpWorkspace<Types>::SclpWorkspace() : impl(std::make_unique<Impl>()) {}
template <typename Types>
SclpWorkspace<Types>::~SclpWorkspace() = default;
template <typename Types>
SclpWorkspace<Types>::SclpWorkspace(SclpWorkspace&&) noexcept = default;
template <typename Types>
SclpWorkspace<Types>& SclpWorkspace<Types>::operator=(SclpWorkspace&&) noexcept = default;

// Edge-sized scratch is drawn per contraction from the context's caching
// allocator, which keeps it for the next level; nothing to reserve up front.
template <typename Types>
void SclpWorkspace<Types>::reserve_contract_buffers(std::int64_t max_edges) {
    if (max_edges < 0) throw std::invalid_argument("negative contraction buffer size");
}

namespace {

using clock_type = std::chrono::steady_clock;
double since(clock_type::time_point t) { return std::chrono::duration<double>(clock_type::now() - t).count(); }

gpart_sclp_round make_round(std::uint32_t seed, int level, int round, bool filter, std::uint32_t threshold,
                            bool singleton, std::uint64_t value_bound) {
    gpart_sclp_round r;
    r.role_salt = seed ^ (static_cast<std::uint32_t>(level) * 0x9e3779b9U) ^
                  (static_cast<std::uint32_t>(round) * 0x85ebca6bU) ^ 0x27d4eb2dU;
    r.tie_salt = seed ^ (static_cast<std::uint32_t>(level) * 0xc2b2ae35U) ^
                 (static_cast<std::uint32_t>(round) * 0x165667b1U) ^ 0xd3a2646cU;
    r.mover_threshold = threshold;
    r.filter_roles = filter ? 1 : 0;
    r.singleton = singleton ? 1 : 0;
    r.value_bound = value_bound;
    return r;
}

template <typename Types>
std::uint64_t cluster_cut(const DeviceWeightedGraph<Types>& g, const DeviceVector<std::int32_t>& clusters) {
    const auto view = g.view();
    std::uint64_t cut = 0;
    GPART_CHECK(gpart_cut(detail::context(), &view, clusters.data(), &cut));
    return cut;
}

}  // namespace

template <typename Types>
DeviceAggregateResult<Types> aggregate(const DeviceWeightedGraph<Types>& graph, SclpWorkspace<Types>& workspace,
                                       int parts, std::uint32_t seed, int level, SclpStats& stats, bool diagnostics,
                                       int cluster_size_factor, bool pack_isolated) {
    if (parts < 2 || parts > 32 || cluster_size_factor < 2)
        throw std::invalid_argument("invalid SCLP cluster capacity options");
    auto& ws = *workspace.impl;
    auto* ctx = detail::context();
    const auto start = clock_type::now();
    const auto n = graph.vertices();
    const auto m = graph.edges();
    if (n <= 0 || n > std::numeric_limits<std::int32_t>::max())
        throw std::runtime_error("SCLP coarsening received an unsupported graph size");
    constexpr int maximum_rounds = 4;
    constexpr double two_hop_threshold = 0.60;
    const auto view = graph.view();
    std::uint64_t total_weight = 0, maximum_vertex = 0, minimum_vertex = 0;
    if (ws.weight_stats_known) {
        total_weight = ws.known_total_weight;
        maximum_vertex = ws.known_maximum_weight;
        minimum_vertex = 1;  // a lower bound: sums of positive weights
        ws.weight_stats_known = false;
    } else {
        GPART_CHECK(gpart_vertex_weight_stats(ctx, &view, &total_weight, &maximum_vertex, &minimum_vertex));
    }
    ws.known_total_weight = total_weight;
    if (minimum_vertex == 0)
        throw std::invalid_argument("SCLP coarsening requires positive vertex weights");
    const auto denominator = static_cast<std::uint64_t>(cluster_size_factor) * static_cast<std::uint64_t>(parts);
    const auto average_cap = total_weight / denominator + (total_weight % denominator != 0 ? 1 : 0);
    stats.capacity = std::max(maximum_vertex, average_cap);

    // gains are bounded by row weights; aggregate() called on its own measures them
    std::uint64_t gain_bound = ws.record_bound_known ? ws.maximum_row_weight : 0;
    if (!ws.record_bound_known) {
        DeviceVector<std::uint64_t> bounds(static_cast<std::size_t>(n));
        int unsorted = 0;
        GPART_CHECK(gpart_record_bounds(ctx, &view, bounds.data(), &gain_bound, &unsorted));
    }

    DeviceVector<std::int32_t> clusters(static_cast<std::size_t>(n));
    GPART_CHECK(gpart_sequence_i32(ctx, clusters.data(), n));
    DeviceVector<std::uint64_t> cluster_weights(static_cast<std::size_t>(n));
    GPART_CHECK(gpart_widen_weights(ctx, graph.vertex_weights.data(), Types::kWidth, n, cluster_weights.data()));
    DeviceVector<std::uint64_t> vertex_weights64(cluster_weights);
    ws.forget_classes();
    const auto& classes = ws.classes_of(graph);

    DeviceVector<std::int32_t> proposals(static_cast<std::size_t>(n));
    DeviceVector<std::uint64_t> gains(static_cast<std::size_t>(n));
    DeviceVector<std::uint32_t> ties(static_cast<std::size_t>(n));

    std::int64_t cluster_count = n;
    const std::int64_t desired_clusters = (n + 1) / 2;
    for (int round = 0; round < maximum_rounds; ++round) {
        if (static_cast<double>(cluster_count) <= 0.60 * n) break;
        const auto removable = std::max<std::int64_t>(0, cluster_count - desired_clusters);
        const double mover_fraction =
            std::min(0.50, static_cast<double>(removable) / static_cast<double>(cluster_count));
        const auto threshold = static_cast<std::uint32_t>(mover_fraction * 4294967296.0);
        const auto affinity_start = clock_type::now();
        const auto r =
            make_round(seed, level, round, true, threshold, round == 0 && ws.sorted_unique_rows, gain_bound);
        GPART_CHECK(gpart_sclp_proposals(ctx, &view, clusters.data(), &classes, &r, proposals.data(), gains.data(),
                                         ties.data()));
        stats.affinity_seconds += since(affinity_start);
        const auto admission_start = clock_type::now();
        const auto cut_before = diagnostics ? cluster_cut(graph, clusters) : 0;
        std::int64_t valid = 0, accepted = 0, nonempty = 0;
        std::uint64_t predicted_gain = 0;
        GPART_CHECK(gpart_sclp_admit(ctx, n, proposals.data(), gains.data(), ties.data(), vertex_weights64.data(),
                                     stats.capacity, gain_bound, clusters.data(), cluster_weights.data(), &valid,
                                     &accepted, &nonempty, &predicted_gain));
        cluster_count = nonempty;
        stats.lp_accepted += static_cast<std::uint64_t>(accepted);
        stats.capacity_rejected += static_cast<std::uint64_t>(valid - accepted);
        stats.proposal_count += static_cast<std::uint64_t>(valid);
        ++stats.rounds;
        stats.admission_seconds += since(admission_start);
        gpart::detail::log() << "ml_gpu_sclp_round level=" << level << " round=" << round << " proposals=" << valid;
        if (diagnostics) {
            const auto cut_after = cluster_cut(graph, clusters);
            if (cut_after > cut_before) throw std::runtime_error("SCLP synchronous batch increased cluster cut");
            gpart::detail::log() << " accepted=" << accepted << " capacity_rejected=" << (valid - accepted)
                      << " predicted_gain=" << predicted_gain << " capacity_rejection_ratio="
                      << (valid == 0 ? 0.0 : static_cast<double>(valid - accepted) / static_cast<double>(valid))
                      << " cut_before=" << cut_before << " cut_after=" << cut_after
                      << " actual_gain=" << (cut_before - cut_after);
        }
        gpart::detail::log() << " mover_fraction=" << mover_fraction << " clusters=" << cluster_count
                  << " contraction_ratio=" << static_cast<double>(cluster_count) / static_cast<double>(n) << '\n';
        if (static_cast<double>(cluster_count) <= 0.60 * n) break;
    }

    // a level without two isolated vertices has nothing to pack
    if (pack_isolated && classes.isolated_count > 1) {
        std::int64_t isolated = 0, nonempty = cluster_count;
        GPART_CHECK(gpart_sclp_isolated(ctx, &view, vertex_weights64.data(), stats.capacity, clusters.data(),
                                        cluster_weights.data(), &isolated, &nonempty));
        if (isolated > 1) {
            cluster_count = nonempty;
            gpart::detail::log() << "ml_gpu_isolated_pack level=" << level << " isolated_vertices=" << isolated
                      << " coarse_vertices=" << cluster_count << '\n';
        }
    }

    const bool two_hop_triggered = static_cast<double>(cluster_count) > two_hop_threshold * n;
    if (two_hop_triggered) {
        const auto two_hop_start = clock_type::now();
        const auto cut_before = diagnostics ? cluster_cut(graph, clusters) : 0;
        const auto r = make_round(seed, level, 4, false, 0, false, gain_bound);
        GPART_CHECK(gpart_sclp_proposals(ctx, &view, clusters.data(), &classes, &r, proposals.data(), gains.data(),
                                         ties.data()));
        std::int64_t candidates = 0, pairable = 0, merged = 0, nonempty = 0;
        const std::int64_t budget = std::max<std::int64_t>(0, cluster_count - desired_clusters);
        GPART_CHECK(gpart_sclp_two_hop(ctx, n, proposals.data(), gains.data(), ties.data(), vertex_weights64.data(),
                                       stats.capacity, gain_bound, budget, clusters.data(), cluster_weights.data(),
                                       &candidates, &pairable, &merged, &nonempty));
        cluster_count = nonempty;
        stats.singleton_with_favorite = static_cast<std::uint64_t>(candidates);
        stats.pairable_singletons = 2 * static_cast<std::uint64_t>(pairable);
        stats.two_hop_merged = static_cast<std::uint64_t>(merged);
        if (diagnostics) {
            const auto cut_after = cluster_cut(graph, clusters);
            if (cut_after > cut_before) throw std::runtime_error("SCLP two-hop pairing increased cluster cut");
            gpart::detail::log() << "ml_gpu_sclp_two_hop level=" << level << " enabled=1 threshold=" << two_hop_threshold
                      << " triggered=1 singleton_with_favorite=" << candidates
                      << " pairable_singletons=" << 2 * pairable << " merged=" << merged
                      << " cut_before=" << cut_before << " cut_after=" << cut_after
                      << " actual_gain=" << (cut_before - cut_after) << '\n';
        }
        stats.two_hop_seconds = since(two_hop_start);
    }

    const auto compact_start = clock_type::now();
    DeviceAggregateResult<Types> out;
    out.coarse_vertices = static_cast<typename Types::VertexT>(cluster_count);
    out.capacity = stats.capacity;
    out.map.resize_discard(static_cast<std::size_t>(n));
    out.vertex_weights.resize_discard(static_cast<std::size_t>(cluster_count));
    GPART_CHECK(gpart_sclp_compact(ctx, n, clusters.data(), cluster_weights.data(), &view, out.map.data(),
                                   cluster_count, out.vertex_weights.data(), &out.maximum_weight));
    if (out.maximum_weight > stats.capacity) throw std::runtime_error("SCLP coarsening exceeded cluster capacity");
    stats.compact_seconds = since(compact_start);
    stats.aggregate_seconds = since(start);
    gpart::detail::log() << "ml_gpu_sclp level=" << level << " fine_vertices=" << n << " coarse_vertices=" << out.coarse_vertices
              << " contraction_ratio=" << static_cast<double>(out.coarse_vertices) / static_cast<double>(n)
              << " beta=" << cluster_size_factor << " cluster_cap=" << stats.capacity << " lp_rounds=" << stats.rounds
              << " two_hop_enabled=1 two_hop_threshold=" << two_hop_threshold;
    if (diagnostics)
        gpart::detail::log() << " lp_accepted=" << stats.lp_accepted << " capacity_rejected=" << stats.capacity_rejected
                  << " two_hop_merged=" << stats.two_hop_merged;
    gpart::detail::log() << " weight_max=" << out.maximum_weight
              << " weight_max_over_cap=" << static_cast<double>(out.maximum_weight) / static_cast<double>(stats.capacity)
              << " low_vertices=" << classes.low_count << " medium_vertices=" << classes.medium_count
              << " high_vertices=" << classes.high_count << " hub_edge_entries=" << classes.high_edges
              << " hub_fraction=" << (m == 0 ? 0.0 : static_cast<double>(classes.high_edges) / static_cast<double>(m))
              << " seed=" << seed << " diagnostics=" << (diagnostics ? 1 : 0) << '\n';
    return out;
}

namespace {

template <typename Types>
DeviceWeightedGraph<Types> contract_impl(const DeviceWeightedGraph<Types>& fine, DeviceAggregateResult<Types>&& agg,
                                         SclpWorkspace<Types>& workspace, ContractionTimings& timings) {
    const auto start = clock_type::now();
    auto& ws = *workspace.impl;
    auto* ctx = detail::context();
    const auto n = fine.vertices();
    std::uint64_t coarse_row_bound = ~std::uint64_t{0};
    if (ws.record_bound_known) {
        ws.next_row_bounds.resize_discard(static_cast<std::size_t>(agg.coarse_vertices));
        GPART_CHECK(gpart_contract_record_bounds(ctx, n, agg.map.data(), ws.row_bounds.data(), agg.coarse_vertices,
                                                 ws.next_row_bounds.data(), &coarse_row_bound));
    }
    const auto view = fine.view();
    const auto& classes = ws.classes_of(fine);
    void* offsets = nullptr;
    std::int32_t* neighbors = nullptr;
    void* weights = nullptr;
    std::int64_t coarse_m = 0;
    GPART_CHECK(gpart_contract(ctx, &view, &classes, agg.map.data(), agg.coarse_vertices, coarse_row_bound, &offsets,
                               &neighbors, &weights, &coarse_m));
    DeviceWeightedGraph<Types> coarse;
    coarse.offsets = DeviceVector<typename Types::OffsetT>::adopt(static_cast<typename Types::OffsetT*>(offsets),
                                                                  static_cast<std::size_t>(agg.coarse_vertices) + 1);
    coarse.neighbors = DeviceVector<std::int32_t>::adopt(neighbors, static_cast<std::size_t>(coarse_m));
    coarse.edge_weights = DeviceVector<typename Types::WeightT>::adopt(static_cast<typename Types::WeightT*>(weights),
                                                                       static_cast<std::size_t>(coarse_m));
    if (coarse_m == 0) {
        // no cross entries: clear() returns the two (unused) blocks to the context
        coarse.neighbors.clear();
        coarse.edge_weights.clear();
    }
    coarse.vertex_weights = std::move(agg.vertex_weights);
    gpart::detail::log() << "ml_gpu_contract_cross_edge_entries=" << coarse_m << " input_edge_entries=" << fine.edges()
              << " packed_bound=" << coarse_row_bound << '\n';
    if (ws.record_bound_known) {
        ws.maximum_row_weight = coarse_row_bound;
        ws.row_bounds.swap(ws.next_row_bounds);
    }
    ws.forget_classes();
    ws.known_maximum_weight = agg.maximum_weight;
    ws.weight_stats_known = true;
    timings.total_seconds = since(start);
    return coarse;
}

template <typename Types>
void validate_coarsening_step(const WeightedGraph<Types>& fine, const WeightedGraph<Types>& coarse,
                              const std::vector<typename Types::VertexT>& map, std::uint64_t cluster_cap, int level,
                              bool strict) {
    // src/coarsen.cu:1835-1902
    using VertexT = typename Types::VertexT;
    if (map.size() != static_cast<std::size_t>(fine.vertices()))
        throw std::runtime_error("coarsening map has the wrong length");
    if (coarse.vertices() <= 0 || coarse.vertices() > fine.vertices())
        throw std::runtime_error("coarsening produced an invalid vertex count");
    std::vector<std::uint64_t> recomputed(static_cast<std::size_t>(coarse.vertices()), 0);
    std::vector<std::uint64_t> coverage(static_cast<std::size_t>(coarse.vertices()), 0);
    for (std::int64_t v = 0; v < fine.vertices(); ++v) {
        const auto c = map[static_cast<std::size_t>(v)];
        if (c < 0 || c >= coarse.vertices()) throw std::runtime_error("coarsening map contains an out-of-range id");
        recomputed[c] += static_cast<std::uint64_t>(fine.vertex_weights[v]);
        ++coverage[c];
    }
    auto sum = [](const auto& w) {
        return std::accumulate(w.begin(), w.end(), std::uint64_t{0},
                               [](std::uint64_t t, auto x) { return t + static_cast<std::uint64_t>(x); });
    };
    if (sum(fine.vertex_weights) != sum(coarse.vertex_weights))
        throw std::runtime_error("coarsening did not conserve vertex weight");
    for (std::int64_t c = 0; c < coarse.vertices(); ++c) {
        if (coverage[c] == 0 || recomputed[c] != static_cast<std::uint64_t>(coarse.vertex_weights[c]))
            throw std::runtime_error("coarsening cluster weight mismatch");
        if (static_cast<std::uint64_t>(coarse.vertex_weights[c]) > cluster_cap)
            throw std::runtime_error("coarse point exceeds configured weight cap");
    }
    if (strict) validate_weighted_csr(coarse, false);
    else validate_weighted_shape(coarse, false);
    std::vector<VertexT> coarse_labels(static_cast<std::size_t>(coarse.vertices()));
    for (std::int64_t c = 0; c < coarse.vertices(); ++c)
        coarse_labels[c] = static_cast<VertexT>((static_cast<std::uint64_t>(c) * 2654435761ULL + 17) % 23);
    std::vector<VertexT> fine_labels(map.size());
    for (std::size_t v = 0; v < map.size(); ++v) fine_labels[v] = coarse_labels[map[v]];
    if (host_cut(fine, fine_labels) != host_cut(coarse, coarse_labels))
        throw std::runtime_error("cut is not conserved by coarse projection");
    gpart::detail::log() << "ml_layer_verify level=" << level << " map=ok weights=ok csr=ok projection_cut=ok\n";
}

template <typename Types, bool Resident>
auto coarsen_impl(const WeightedGraph<Types>* graph, DeviceWeightedGraph<Types> current,
                  const CoarsenOptions& options) {
    if (options.parts < 2 || options.parts > 32 || options.cluster_size_factor < 2 || options.max_levels < 0 ||
        !std::isfinite(options.stop_contraction_ratio) || options.stop_contraction_ratio <= 0 ||
        options.stop_contraction_ratio > 1)
        throw std::invalid_argument("invalid coarsening options");
    const bool diagnostics = std::getenv("SCLP_DIAGNOSTICS") != nullptr;
    const bool verify = options.strict_verify;
    std::conditional_t<Resident, DeviceHierarchy<Types>, Hierarchy<Types>> hierarchy;
    double input_verify_seconds = 0.0;
    if constexpr (!Resident) {
        hierarchy.levels.push_back(*graph);
        const auto t = clock_type::now();
        if (options.strict_verify) validate_weighted_csr(hierarchy.levels.front(), true);
        else validate_weighted_shape(hierarchy.levels.front(), true);
        input_verify_seconds = since(t);
    }
    gpart::detail::log() << "ml_input_verify_seconds=" << input_verify_seconds
              << " status=ok mode=" << (options.strict_verify ? "strict" : "fast") << '\n';
    const auto device_input_start = clock_type::now();
    if constexpr (!Resident) current = make_device_weighted(hierarchy.levels.front());
    SclpWorkspace<Types> workspace;
    auto* ctx = detail::context();
    if (current.vertices() > 0) {
        auto& ws = *workspace.impl;
        ws.row_bounds.resize_discard(static_cast<std::size_t>(current.vertices()));
        const auto view = current.view();
        int unsorted = 0;
        GPART_CHECK(gpart_record_bounds(ctx, &view, ws.row_bounds.data(), &ws.maximum_row_weight, &unsorted));
        ws.record_bound_known = true;
        ws.sorted_unique_rows = unsorted == 0;
    }
    gpart::detail::log() << "ml_gpu_device_input_seconds=" << since(device_input_start) << " method=sclp\n";

    const std::int64_t cutoff =
        std::max<std::int64_t>(32, options.cluster_size_factor * static_cast<std::int64_t>(options.parts));
    const auto core_start = clock_type::now();
    double snapshot_seconds = 0.0;
    int level = 0;
    for (; level < options.max_levels && current.vertices() > cutoff; ++level) {
        const auto level_seed = options.seed + static_cast<std::uint32_t>(level);
        gpart::detail::log() << "ml_coarsen_begin level=" << level << " vertices=" << current.vertices()
                  << " edges=" << current.edges() << " method=sclp seed=" << level_seed << '\n';
        SclpStats stats;
        auto agg = aggregate(current, workspace, options.parts, level_seed, level, stats, diagnostics,
                             options.cluster_size_factor, options.pack_isolated);
        const double contraction = static_cast<double>(agg.coarse_vertices) / static_cast<double>(current.vertices());
        gpart::detail::log() << "ml_coarsen_map level=" << level << " coarse_vertices=" << agg.coarse_vertices
                  << " ratio=" << contraction << " maximum_weight=" << agg.maximum_weight
                  << " capacity=" << agg.capacity << '\n';
        if (agg.coarse_vertices >= current.vertices() || contraction > options.stop_contraction_ratio) {
            gpart::detail::log() << "ml_coarsen_stop reason=insufficient_contraction\n";
            hierarchy.stop_reason = "insufficient_contraction";
            break;
        }
        ContractionTimings timings;
        auto coarse = contract_impl(current, std::move(agg), workspace, timings);
        gpart::detail::log() << "ml_gpu_contract_seconds level=" << level << " seconds=" << timings.total_seconds
                  << " coarse_edges=" << coarse.edges() << '\n';
        const auto snapshot_start = clock_type::now();
        if constexpr (Resident) {
            if (verify)
                validate_coarsening_step(copy_device_weighted(current), copy_device_weighted(coarse), agg.map.to_host(),
                                         agg.capacity, level, true);
            hierarchy.levels.push_back(std::move(current));
            hierarchy.fine_to_coarse.push_back(std::move(agg.map));
        } else {
            auto host_map = agg.map.to_host();
            auto coarse_host = copy_device_weighted(coarse);
            if (verify)
                validate_coarsening_step(hierarchy.levels.back(), coarse_host, host_map, agg.capacity, level,
                                         options.strict_verify);
            hierarchy.fine_to_coarse.push_back(std::move(host_map));
            hierarchy.levels.push_back(std::move(coarse_host));
            snapshot_seconds += since(snapshot_start);
        }
        current = std::move(coarse);
        // contraction emits each row sorted and duplicate-free
        workspace.impl->sorted_unique_rows = true;
    }
    if (level == options.max_levels && current.vertices() > cutoff) hierarchy.stop_reason = "max_levels";
    if constexpr (Resident) hierarchy.levels.push_back(std::move(current));
    hierarchy.hierarchy_loop_seconds = since(core_start);
    hierarchy.snapshot_seconds = snapshot_seconds;
    gpart::detail::log() << "ml_hierarchy_loop_seconds=" << hierarchy.hierarchy_loop_seconds
              << " includes_snapshot_and_verify=1 method=sclp\n";
    return hierarchy;
}

}  // namespace

template <typename Types>
DeviceWeightedGraph<Types> contract(DeviceWeightedGraph<Types> fine, DeviceAggregateResult<Types>&& aggregate,
                                    SclpWorkspace<Types>& workspace, ContractionTimings& timings) {
    return contract_impl(fine, std::move(aggregate), workspace, timings);
}

template <typename Types>
Hierarchy<Types> coarsen(const WeightedGraph<Types>& graph, const CoarsenOptions& options) {
    return coarsen_impl<Types, false>(&graph, {}, options);
}

template <typename Types>
DeviceHierarchy<Types> coarsen(DeviceWeightedGraph<Types>&& graph, const CoarsenOptions& options) {
    return coarsen_impl<Types, true>(nullptr, std::move(graph), options);
}

#define GPART_INSTANTIATE_COARSEN(T)                                                                             \
    template class SclpWorkspace<T>;                                                                             \
    template DeviceAggregateResult<T> aggregate<T>(const DeviceWeightedGraph<T>&, SclpWorkspace<T>&, int,        \
                                                   std::uint32_t, int, SclpStats&, bool, int, bool);             \
    template DeviceWeightedGraph<T> contract<T>(DeviceWeightedGraph<T>, DeviceAggregateResult<T>&&,              \
                                                SclpWorkspace<T>&, ContractionTimings&);                         \
    template Hierarchy<T> coarsen<T>(const WeightedGraph<T>&, const CoarsenOptions&);                            \
    template DeviceHierarchy<T> coarsen<T>(DeviceWeightedGraph<T>&&, const CoarsenOptions&);
GPART_INSTANTIATE_COARSEN(StandardTypes)
GPART_INSTANTIATE_COARSEN(BigTypes)

}  // namespace gpart
