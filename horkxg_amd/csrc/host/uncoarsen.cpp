// uncoarsen.cpp -- the uncoarsen entry points: projection and refinement over a
// whole hierarchy (done by refine_hierarchy_device), then one report line per
// refined level and the final_* lines.
//
// Reference interface: include/uncoarsen.hpp:10-22, src/uncoarsen.cu:39-112. The
// log keys and their meaning are the drop-in contract that scripts parse.
#include "log.hpp"
#include "uncoarsen.hpp"

#include <chrono>
#include <stdexcept>

namespace gpart {
namespace {

// Total, heaviest part and imbalance (heaviest * parts / total) of the final
// part weights; the total must fit 64 bits.
struct PartLoadSummary {
    std::uint64_t total = 0;
    std::uint64_t heaviest = 0;
    double imbalance = 0.0;
};

PartLoadSummary summarize(const std::vector<std::uint64_t>& loads) {
    PartLoadSummary s;
    for (const std::uint64_t w : loads) {
        if (__builtin_add_overflow(s.total, w, &s.total)) throw std::overflow_error("total vertex weight overflow");
        if (w > s.heaviest) s.heaviest = w;
    }
    if (s.total != 0)
        s.imbalance = static_cast<double>(s.heaviest) * static_cast<double>(loads.size()) /
                      static_cast<double>(s.total);
    return s;
}

void report_level(const RefineLevelResult& r) {
    const RefineStats& s = r.refinement;
    const RefineLevelTimings& t = r.timings;
    detail::ReportLine line;
    line.field("uncoarsen_level", r.level)
        .field("vertices", r.vertices)
        .field("edge_entries", r.edge_entries)
        .field("projection_cut", r.projection_cut)
        .field("projection_max_part_weight", r.projection_max_part_weight)
        .field("projection_imbalance", r.projection_imbalance)
        .field("refine_rounds", s.rounds)
        .field("refine_proposals", s.proposals)
        .field("refine_accepted", s.accepted)
        .field("admission_iterations", s.admission_iterations)
        .field("admission_fallbacks", s.admission_fallbacks)
        .field("refined_cut", s.final_cut)
        .field("refined_max_part_weight", s.final_max_part_weight)
        .field("graph_h2d_seconds", t.graph_h2d_seconds)
        .field("workspace_resize_seconds", t.workspace_resize_seconds)
        .field("projection_gpu_seconds", t.projection_gpu_seconds)
        .field("refinement_seconds", t.refinement_seconds)
        .field("verification_seconds", t.verification_seconds)
        // beyond the reference's keys: the apply path's volume and the rounds
        // that rebuilt the connectivity cache
        .field("mover_entries", s.mover_entries)
        .field("cache_rebuilds", s.cache_rebuilds);
    line.write(detail::log());
}

template <typename Types>
void report_final(const DeviceUncoarsenResult<Types>& result, double seconds) {
    if (result.final_part_weights.empty()) throw std::runtime_error("uncoarsen produced no part weights");
    const PartLoadSummary loads = summarize(result.final_part_weights);
    const double cut_ratio = result.total_edge_weight == 0
                                 ? 0.0
                                 : static_cast<double>(result.final_cut) /
                                       static_cast<double>(result.total_edge_weight);
    auto& out = detail::log();
    detail::ReportLine().field("final_edge_cut", result.final_cut).write(out);
    detail::ReportLine().field("final_cut_ratio", cut_ratio).write(out);
    detail::ReportLine().field("final_max_part_weight", loads.heaviest).write(out);
    detail::ReportLine().field("final_imbalance", loads.imbalance).write(out);
    detail::ReportLine().field("final_d2h_seconds", result.final_d2h_seconds).write(out);
    detail::ReportLine().field("uncoarsen_total_seconds", seconds).write(out);
}

// H is a host Hierarchy (levels uploaded as they are refined) or a DeviceHierarchy
// (levels moved out); Labels is the matching coarsest partition.
template <typename Types, typename H, typename Labels>
DeviceUncoarsenResult<Types> refine_and_report(H& hierarchy, Labels&& coarsest, const RefineOptions& options) {
    using clock = std::chrono::steady_clock;
    const auto begin = clock::now();
    const bool complete = !hierarchy.levels.empty() &&
                          hierarchy.levels.size() == hierarchy.fine_to_coarse.size() + 1;
    if (!complete) throw std::invalid_argument("uncoarsen requires a complete hierarchy");
    DeviceUncoarsenResult<Types> result =
        refine_hierarchy_device(hierarchy, std::forward<Labels>(coarsest), options);
    for (const RefineLevelResult& level : result.levels) report_level(level);
    report_final(result, std::chrono::duration<double>(clock::now() - begin).count());
    return result;
}

}  // namespace

template <typename Types>
std::vector<typename Types::VertexT> uncoarsen(const Hierarchy<Types>& hierarchy,
                                               const std::vector<typename Types::VertexT>& coarsest_partition,
                                               const RefineOptions& options) {
    return refine_and_report<Types>(hierarchy, coarsest_partition, options).partition;
}

template <typename Types>
DeviceUncoarsenResult<Types> uncoarsen(DeviceHierarchy<Types>& hierarchy,
                                       DeviceVector<typename Types::VertexT>&& coarsest_partition,
                                       const RefineOptions& options) {
    return refine_and_report<Types>(hierarchy, std::move(coarsest_partition), options);
}

#define GPART_UNCOARSEN_INSTANCES(T)                                                                        \
    template std::vector<std::int32_t> uncoarsen<T>(const Hierarchy<T>&, const std::vector<std::int32_t>&, \
                                                    const RefineOptions&);                                \
    template DeviceUncoarsenResult<T> uncoarsen<T>(DeviceHierarchy<T>&, DeviceVector<std::int32_t>&&,     \
                                                   const RefineOptions&);
GPART_UNCOARSEN_INSTANCES(StandardTypes)
GPART_UNCOARSEN_INSTANCES(BigTypes)
#undef GPART_UNCOARSEN_INSTANCES

}  // namespace gpart
