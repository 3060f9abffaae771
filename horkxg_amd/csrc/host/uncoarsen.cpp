// uncoarsen.cpp -- projection + refinement driver and its report lines
// (reference: src/uncoarsen.cu:39-112).
#include "log.hpp"
#include "uncoarsen.hpp"

#include <algorithm>
#include <chrono>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>

namespace gpart {
namespace {

std::uint64_t sum_weights(const std::vector<std::uint64_t>& weights) {
    std::uint64_t total = 0;
    for (const auto w : weights) {
        if (total > std::numeric_limits<std::uint64_t>::max() - w) throw std::overflow_error("total vertex weight overflow");
        total += w;
    }
    return total;
}

double imbalance_of(const std::vector<std::uint64_t>& weights, std::uint64_t total) {
    if (total == 0) return 0.0;
    const auto maximum = *std::max_element(weights.begin(), weights.end());
    return static_cast<double>(maximum) * weights.size() / static_cast<double>(total);
}

template <typename Types, typename H, typename P>
DeviceUncoarsenResult<Types> uncoarsen_impl(H& hierarchy, P&& coarsest_partition, const RefineOptions& options) {
    const auto start = std::chrono::steady_clock::now();
    if (hierarchy.levels.empty() || hierarchy.fine_to_coarse.size() + 1 != hierarchy.levels.size())
        throw std::invalid_argument("uncoarsen requires a complete hierarchy");
    auto result = refine_hierarchy_device(hierarchy, std::forward<P>(coarsest_partition), options);
    for (const auto& level : result.levels) {
        const auto& s = level.refinement;
        const auto& t = level.timings;
        gpart::detail::log() << std::setprecision(10) << "uncoarsen_level=" << level.level << " vertices=" << level.vertices
                  << " edge_entries=" << level.edge_entries << " projection_cut=" << level.projection_cut
                  << " projection_max_part_weight=" << level.projection_max_part_weight
                  << " projection_imbalance=" << level.projection_imbalance << " refine_rounds=" << s.rounds
                  << " refine_proposals=" << s.proposals << " refine_accepted=" << s.accepted
                  << " admission_iterations=" << s.admission_iterations
                  << " admission_fallbacks=" << s.admission_fallbacks << " refined_cut=" << s.final_cut
                  << " refined_max_part_weight=" << s.final_max_part_weight
                  << " graph_h2d_seconds=" << t.graph_h2d_seconds
                  << " workspace_resize_seconds=" << t.workspace_resize_seconds
                  << " projection_gpu_seconds=" << t.projection_gpu_seconds
                  << " refinement_seconds=" << t.refinement_seconds
                  << " verification_seconds=" << t.verification_seconds << " mover_entries=" << s.mover_entries
                  << " cache_rebuilds=" << s.cache_rebuilds << '\n';
    }
    const auto total = sum_weights(result.final_part_weights);
    const auto maximum = *std::max_element(result.final_part_weights.begin(), result.final_part_weights.end());
    gpart::detail::log() << std::setprecision(10) << "final_edge_cut=" << result.final_cut << '\n'
              << "final_cut_ratio="
              << (result.total_edge_weight == 0
                      ? 0.0
                      : static_cast<double>(result.final_cut) / static_cast<double>(result.total_edge_weight))
              << '\n'
              << "final_max_part_weight=" << maximum << '\n'
              << "final_imbalance=" << imbalance_of(result.final_part_weights, total) << '\n'
              << "final_d2h_seconds=" << result.final_d2h_seconds << '\n'
              << "uncoarsen_total_seconds="
              << std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() << '\n';
    return result;
}

}  // namespace

template <typename Types>
std::vector<typename Types::VertexT> uncoarsen(const Hierarchy<Types>& hierarchy,
                                               const std::vector<typename Types::VertexT>& partition,
                                               const RefineOptions& options) {
    auto result = uncoarsen_impl<Types>(hierarchy, partition, options);
    return std::move(result.partition);
}

template <typename Types>
DeviceUncoarsenResult<Types> uncoarsen(DeviceHierarchy<Types>& hierarchy,
                                       DeviceVector<typename Types::VertexT>&& partition,
                                       const RefineOptions& options) {
    return uncoarsen_impl<Types>(hierarchy, std::move(partition), options);
}

template std::vector<std::int32_t> uncoarsen<StandardTypes>(const Hierarchy<StandardTypes>&,
                                                            const std::vector<std::int32_t>&, const RefineOptions&);
template std::vector<std::int32_t> uncoarsen<BigTypes>(const Hierarchy<BigTypes>&, const std::vector<std::int32_t>&,
                                                       const RefineOptions&);
template DeviceUncoarsenResult<StandardTypes> uncoarsen<StandardTypes>(DeviceHierarchy<StandardTypes>&,
                                                                       DeviceVector<std::int32_t>&&,
                                                                       const RefineOptions&);
template DeviceUncoarsenResult<BigTypes> uncoarsen<BigTypes>(DeviceHierarchy<BigTypes>&, DeviceVector<std::int32_t>&&,
                                                             const RefineOptions&);

}  // namespace gpart
