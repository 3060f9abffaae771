// quality.hpp -- partition verification and partition-quality metrics over the
// device counters of gpart_part_counters.
//
// verify_partition: tools/verify_partition.py:22-65 (same checks, messages and
// result fields; the capacity is the exact decimal floor of :57).
// analyze_partition: graph_untils/analyze_partition.cpp:176-294 (parts inferred as
// max label + 1 unless given; the same averages, imbalances and deviations). The
// weighted metrics use the graph's integer edge weights; with unit weights they
// equal the tool's output without --values.
#pragma once

#include "gpart_hip.h"

#include <cstdint>
#include <string>
#include <vector>

namespace gpart {

struct VerifyResult {
    std::int64_t vertices = 0;
    std::int64_t undirected_edges = 0;
    std::int64_t cut = 0;
    double cut_ratio = 0.0;
    std::vector<std::int64_t> loads;
    std::int64_t max_load = 0;
    double max_load_ratio = 0.0;
    std::int64_t capacity = 0;
    bool capacity_ok = false;
};

struct QualityResult {
    int parts = 0;
    std::int64_t vertices = 0;
    std::int64_t directed_edges = 0;
    std::uint64_t edge_cut = 0;
    std::uint64_t weighted_edge_cut = 0;
    double edge_cut_ratio = 0.0;
    double avg_local_cut = 0.0;
    std::uint64_t max_local_cut = 0;
    double weighted_avg_local_cut = 0.0;
    std::uint64_t weighted_max_local_cut = 0;
    double vertex_avg = 0.0;
    std::int64_t min_vertices = 0;
    std::int64_t max_vertices = 0;
    double vertex_imbalance = 0.0;
    double vertex_max_deviation = 0.0;
    double edge_set_avg = 0.0;
    std::int64_t min_edge_set = 0;
    std::int64_t max_edge_set = 0;
    double edge_set_imbalance = 0.0;
    double edge_set_max_deviation = 0.0;
    // 7 x parts: vertices, internal, outgoing, incoming, weighted internal,
    // weighted outgoing, weighted incoming
    std::vector<std::uint64_t> per_part;
};

// floor(ratio * n / parts) for a finite decimal ratio >= 1, computed exactly;
// false when ratio is not such a decimal (or has more digits than 128-bit
// arithmetic holds).
bool decimal_capacity(const std::string& ratio, std::int64_t n, int parts, std::int64_t* capacity);

// Labels are device arrays of g.n entries; host_labels (may be null) names the
// first invalid label's value in the error message. Errors are std::invalid_argument
// with the reference's messages.
VerifyResult verify_partition(gpart_ctx* ctx, const gpart_csr_view& g, const std::int32_t* labels,
                              const std::int32_t* host_labels, int parts, const std::string& ratio);
QualityResult analyze_partition(gpart_ctx* ctx, const gpart_csr_view& g, const std::int32_t* labels, int parts);

}  // namespace gpart
