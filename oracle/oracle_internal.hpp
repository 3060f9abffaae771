// oracle_internal.hpp -- C++ internals of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference partitioner,
// used as the parity checker and as the CPU baseline. Never linked by the
// product library.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gpo {

struct Graph {
    std::vector<std::int64_t> off;   // n + 1
    std::vector<std::int32_t> adj;   // m
    std::vector<std::uint64_t> ew;   // m
    std::vector<std::uint64_t> vw;   // n
    std::int64_t n() const { return static_cast<std::int64_t>(vw.size()); }
    std::int64_t m() const { return static_cast<std::int64_t>(adj.size()); }
};

// The public "lowbias32" integer hash (src/coarsen.cu:112-118,
// src/initial_partition_gpu.cu:36-39, src/refine.cu:38-44).
inline std::uint32_t mix32(std::uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    return x ^ (x >> 16);
}

constexpr std::int32_t kSclpInvalid = 0x7fffffff;  // src/coarsen.cu:201

// include/partition_capacity.hpp:12-23
std::uint64_t partition_capacity(std::uint64_t total, int parts, double ratio);

struct AggregateResult {
    std::vector<std::int32_t> map;
    std::vector<std::uint64_t> vw;
    std::int64_t coarse = 0;
    std::uint64_t maximum_weight = 0;
    std::uint64_t capacity = 0;
    int rounds = 0;
    std::uint64_t proposals = 0;
    // Filled only when aggregate() is asked to trace: the clusters after every
    // step (kind 0: LP round, 1: isolated packing, 2: two-hop pairing), with the
    // round number and the mover threshold of LP rounds.
    struct Step {
        int kind = 0;
        int round = 0;
        std::uint32_t threshold = 0;
        std::vector<std::int32_t> clusters;
    };
    std::vector<Step> steps;
};

void sclp_proposals(const Graph& g, const std::vector<std::int32_t>& clusters,
                    std::uint32_t seed, int level, int round, bool filter_roles,
                    std::uint32_t mover_threshold, std::vector<std::int32_t>& proposals,
                    std::vector<std::uint64_t>& gains, std::vector<std::uint32_t>& ties);

AggregateResult aggregate(const Graph& g, int parts, std::uint32_t seed, int level,
                          int cluster_size_factor, bool pack_isolated, bool trace = false);

Graph contract(const Graph& fine, const std::vector<std::int32_t>& map,
               std::int64_t coarse, const std::vector<std::uint64_t>& coarse_weights);

struct Hierarchy {
    std::vector<Graph> levels;
    std::vector<std::vector<std::int32_t>> maps;
    std::string stop_reason = "capacity_floor";
};

struct CoarsenOptions {
    int parts = 4;
    std::uint32_t seed = 0;
    double stop_ratio = 0.90;
    int max_levels = 24;
    int cluster_size_factor = 256;
    bool pack_isolated = false;
};

Hierarchy coarsen(const Graph& g, const CoarsenOptions& options);
void write_hierarchy(const Hierarchy& h, const std::string& path);

std::vector<std::int32_t> initial_partition(const Graph& g, int parts, double ratio,
                                            std::uint32_t seed,
                                            std::vector<double>* coords = nullptr,
                                            int* tasks = nullptr);

struct RefineStats {
    int rounds = 0;
    std::uint64_t proposals = 0;
    std::uint64_t accepted = 0;
    std::uint64_t initial_cut = 0;
    std::uint64_t final_cut = 0;
    std::uint64_t final_max_part_weight = 0;
    int admission_iterations = 0;
    int admission_fallbacks = 0;
};

RefineStats refine(const Graph& g, std::vector<std::int32_t>& labels,
                   std::vector<std::uint64_t>& part_weights, std::uint64_t capacity,
                   int parts, int max_rounds, int stale_rounds);

std::uint64_t cut(const Graph& g, const std::vector<std::int32_t>& labels);
std::vector<std::uint64_t> part_weights(const Graph& g, const std::vector<std::int32_t>& labels,
                                        int parts);
void validate_gain_range(const Graph& g);
void validate_csr(const Graph& g, bool allow_self_loops, bool strict);

struct UncoarsenResult {
    std::vector<std::int32_t> labels;
    std::uint64_t final_cut = 0;
    std::vector<std::uint64_t> final_part_weights;
};
UncoarsenResult uncoarsen(const Hierarchy& h, const std::vector<std::int32_t>& coarsest,
                          int parts, double ratio, int max_rounds, double growth,
                          int stale_rounds = 3);

}  // namespace gpo
