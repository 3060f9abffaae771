// oracle_coarsen.cpp -- CPU restatement of SCLP coarsening and contraction.
//
// TEST INFRASTRUCTURE ONLY (parity checker / CPU baseline).
//
// Follows src/coarsen.cu of the reference. The GPU reference reduces each
// vertex's neighbour-cluster weights with warp shuffles, LDS hash tables or a
// global sort; every route computes the same order-independent quantities:
//   conn[t]      = sum of w(v,u) over entries u != v with cluster(u) == t
//   current      = conn[cluster(v)]
//   best, tie    = lexicographic max of (conn[t], tie(v,t)) over t != cluster(v)
//                  passing the role filter, where affinity 0 never wins a tie
//                  (sclp_better_candidate, src/coarsen.cu:224-229)
// This file computes them directly per vertex.
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <limits>
#include <numeric>
#include <utility>

namespace gpo {
namespace {

inline bool is_mover(std::int32_t cluster, std::uint32_t role_salt, std::uint32_t threshold) {
    // src/coarsen.cu:218-222
    return mix32(static_cast<std::uint32_t>(cluster) ^ role_salt) < threshold;
}

struct Salts {
    std::uint32_t role;
    std::uint32_t tie;
};

Salts make_salts(std::uint32_t seed, int level, int round) {
    // src/coarsen.cu:1152-1157
    Salts s;
    s.role = seed ^ (static_cast<std::uint32_t>(level) * 0x9e3779b9U) ^
             (static_cast<std::uint32_t>(round) * 0x85ebca6bU) ^ 0x27d4eb2dU;
    s.tie = seed ^ (static_cast<std::uint32_t>(level) * 0xc2b2ae35U) ^
            (static_cast<std::uint32_t>(round) * 0x165667b1U) ^ 0xd3a2646cU;
    return s;
}

struct AdmissionEntry {
    std::uint32_t target;
    std::uint64_t gain;
    std::uint32_t tie;
    std::uint32_t vertex;
};

// src/coarsen.cu:600-644 + build_admission_order :1086-1121. The reference sorts
// (target, ~gain, ~tie, vertex) ascending; the same total order written directly.
std::vector<AdmissionEntry> admission_order(const std::vector<std::int32_t>& targets,
                                            const std::vector<std::uint64_t>& gains,
                                            const std::vector<std::uint32_t>& ties) {
    std::vector<AdmissionEntry> out;
    const auto n = static_cast<std::int64_t>(targets.size());
    for (std::int64_t v = 0; v < n; ++v) {
        if (targets[v] == kSclpInvalid) continue;
        out.push_back({static_cast<std::uint32_t>(targets[v]), gains[v], ties[v],
                       static_cast<std::uint32_t>(v)});
    }
    std::sort(out.begin(), out.end(), [](const AdmissionEntry& a, const AdmissionEntry& b) {
        if (a.target != b.target) return a.target < b.target;
        if (a.gain != b.gain) return a.gain > b.gain;
        if (a.tie != b.tie) return a.tie > b.tie;
        return a.vertex < b.vertex;
    });
    return out;
}

std::int64_t count_nonempty(const std::vector<std::uint64_t>& cluster_weights) {
    std::int64_t c = 0;
    for (auto w : cluster_weights) c += w != 0;
    return c;
}

}  // namespace

std::uint64_t partition_capacity(std::uint64_t total, int parts, double ratio) {
    // include/partition_capacity.hpp:12-23
    if (parts < 2 || parts > 32 || !std::isfinite(ratio) || ratio < 1.0)
        throw std::invalid_argument("invalid partition capacity options");
    const long double raw = static_cast<long double>(total) * ratio / parts;
    if (!std::isfinite(raw) ||
        raw >= static_cast<long double>(std::numeric_limits<std::uint64_t>::max()))
        throw std::overflow_error("partition capacity exceeds uint64");
    const auto capacity = static_cast<std::uint64_t>(std::floor(raw));
    if (capacity < total / static_cast<unsigned>(parts) + (total % parts != 0))
        throw std::invalid_argument("integer partition capacity cannot hold total weight");
    return capacity;
}

void sclp_proposals(const Graph& g, const std::vector<std::int32_t>& clusters,
                    std::uint32_t seed, int level, int round, bool filter_roles,
                    std::uint32_t mover_threshold, std::vector<std::int32_t>& proposals,
                    std::vector<std::uint64_t>& gains, std::vector<std::uint32_t>& ties) {
    const auto n = g.n();
    const Salts salt = make_salts(seed, level, round);
    proposals.assign(n, kSclpInvalid);
    gains.assign(n, 0);
    ties.assign(n, 0);
#pragma omp parallel
    {
        std::vector<std::pair<std::int32_t, std::uint64_t>> row;
#pragma omp for schedule(dynamic, 1024)
        for (std::int64_t v = 0; v < n; ++v) {
            row.clear();
            for (auto e = g.off[v]; e < g.off[v + 1]; ++e) {
                const auto u = g.adj[e];
                if (u == v) continue;  // self entries never aggregate (:319,:358,:433,:503)
                row.emplace_back(clusters[u], g.ew[e]);
            }
            std::sort(row.begin(), row.end(),
                      [](const auto& a, const auto& b) { return a.first < b.first; });
            const auto source = clusters[v];
            unsigned long long current = 0, best = 0, best_tie = 0;
            for (std::size_t i = 0; i < row.size();) {
                const auto t = row[i].first;
                unsigned long long conn = 0;
                while (i < row.size() && row[i].first == t) conn += row[i++].second;
                if (t == source) {
                    current = conn;
                    continue;
                }
                if (filter_roles && is_mover(t, salt.role, mover_threshold)) continue;
                const auto hash = mix32(static_cast<std::uint32_t>(v) ^
                                        mix32(static_cast<std::uint32_t>(t)) ^ salt.tie);
                const auto tie = (static_cast<unsigned long long>(hash) << 32) |
                                 (0xffffffffULL - static_cast<std::uint32_t>(t));
                // sclp_better_candidate: affinity 0 never wins (:224-229)
                if (conn > best || (conn == best && conn != 0 && tie > best_tie)) {
                    best = conn;
                    best_tie = tie;
                }
            }
            // sclp_decode_gain_kernel (:562-598)
            const bool source_can_move =
                !filter_roles || is_mover(source, salt.role, mover_threshold);
            const auto role_gain = source_can_move && best > current ? best - current : 0ULL;
            if (role_gain == 0) continue;
            const auto target = static_cast<std::int32_t>(0xffffffffULL - (best_tie & 0xffffffffULL));
            if (target < 0 || target == source) continue;
            proposals[v] = target;
            gains[v] = best - current;
            ties[v] = static_cast<std::uint32_t>(best_tie >> 32);
        }
    }
}

AggregateResult aggregate(const Graph& g, int parts, std::uint32_t seed, int level,
                          int cluster_size_factor, bool pack_isolated, bool trace) {
    // src/coarsen.cu:1661-1671 (option checks) and :937-1659 (aggregate_typed)
    if (parts < 2 || parts > 32 || cluster_size_factor < 2)
        throw std::invalid_argument("invalid SCLP cluster capacity options");
    const auto n = g.n();
    if (n <= 0 || n > std::numeric_limits<std::int32_t>::max())
        throw std::runtime_error("SCLP coarsening received an unsupported graph size");
    for (auto w : g.vw)
        if (w == 0) throw std::invalid_argument("SCLP coarsening requires positive vertex weights");
    constexpr int maximum_rounds = 4;
    constexpr double two_hop_threshold = 0.60;
    std::uint64_t total = 0, maximum_vertex = 0;
    for (auto w : g.vw) {
        total += w;
        maximum_vertex = std::max(maximum_vertex, w);
    }
    const auto denominator = static_cast<std::uint64_t>(cluster_size_factor) *
                             static_cast<std::uint64_t>(parts);
    const auto average_cap = total / denominator + (total % denominator != 0 ? 1 : 0);
    AggregateResult out;
    out.capacity = std::max(maximum_vertex, average_cap);
    const auto capacity = out.capacity;

    std::vector<std::int32_t> clusters(n);
    std::iota(clusters.begin(), clusters.end(), 0);
    std::vector<std::uint64_t> cluster_weights(g.vw.begin(), g.vw.end());

    std::vector<std::int32_t> proposals;
    std::vector<std::uint64_t> gains;
    std::vector<std::uint32_t> ties;

    std::int64_t cluster_count = n;
    const std::int64_t desired_clusters = (n + 1) / 2;
    for (int round = 0; round < maximum_rounds; ++round) {
        // :1307-1408
        if (static_cast<double>(cluster_count) <= 0.60 * n) break;
        const auto removable = std::max<std::int64_t>(0, cluster_count - desired_clusters);
        const double mover_fraction = std::min(
            0.50, static_cast<double>(removable) / static_cast<double>(cluster_count));
        const auto mover_threshold = static_cast<std::uint32_t>(mover_fraction * 4294967296.0);
        sclp_proposals(g, clusters, seed, level, round, true, mover_threshold, proposals,
                       gains, ties);
        const auto order = admission_order(proposals, gains, ties);
        out.proposals += order.size();
        // inclusive_scan_by_key over targets (:1336-1339), base snapshot (:1340),
        // sclp_commit_kernel (:652-680)
        const std::vector<std::uint64_t> base(cluster_weights);
        std::uint64_t prefix = 0;
        for (std::size_t i = 0; i < order.size(); ++i) {
            if (i == 0 || order[i].target != order[i - 1].target) prefix = 0;
            const auto v = static_cast<std::int32_t>(order[i].vertex);
            const auto w = g.vw[v];
            prefix += w;
            const auto t = static_cast<std::int32_t>(order[i].target);
            const auto b = base[t];
            if (b <= capacity && prefix <= capacity - b) {
                const auto source = clusters[v];
                cluster_weights[source] -= w;
                cluster_weights[t] += w;
                clusters[v] = t;
            }
        }
        cluster_count = count_nonempty(cluster_weights);
        ++out.rounds;
        if (trace) out.steps.push_back({0, round, mover_threshold, clusters});
        if (static_cast<double>(cluster_count) <= 0.60 * n) break;
    }

    if (pack_isolated) {
        // :1410-1441 and kernels :741-773
        std::vector<std::int32_t> isolated;
        for (std::int64_t v = 0; v < n; ++v)
            if (g.off[v] == g.off[v + 1]) isolated.push_back(static_cast<std::int32_t>(v));
        const auto count = static_cast<std::int64_t>(isolated.size());
        if (count > 1) {
            std::vector<std::uint64_t> prefix(count);
            std::uint64_t running = 0;
            for (std::int64_t i = 0; i < count; ++i) {
                running += g.vw[isolated[i]];
                prefix[i] = running;
            }
            const auto total_isolated = prefix[count - 1];
            std::vector<std::int32_t> reps(total_isolated / capacity + (total_isolated % capacity != 0),
                                           std::numeric_limits<std::int32_t>::max());
            for (std::int64_t i = 0; i < count; ++i) {
                const auto v = isolated[i];
                const auto first_bin = (prefix[i] - g.vw[v]) / capacity;
                const auto last_bin = (prefix[i] - 1) / capacity;
                if (first_bin == last_bin) reps[first_bin] = std::min(reps[first_bin], v);
            }
            for (std::int64_t i = 0; i < count; ++i) {
                const auto v = isolated[i];
                const auto first_bin = (prefix[i] - g.vw[v]) / capacity;
                const auto last_bin = (prefix[i] - 1) / capacity;
                if (first_bin != last_bin) continue;
                const auto target = reps[first_bin];
                if (target == v) continue;
                clusters[v] = target;
                cluster_weights[v] = 0;
                cluster_weights[target] += g.vw[v];
            }
            cluster_count = count_nonempty(cluster_weights);
        }
        if (trace) out.steps.push_back({1, 0, 0, clusters});
    }

    // Two-hop pairing of singletons sharing a favourite (:1443-1539)
    const bool two_hop_triggered = static_cast<double>(cluster_count) > two_hop_threshold * n;
    if (two_hop_triggered) {
        sclp_proposals(g, clusters, seed, level, 4, false, 0, proposals, gains, ties);
        std::vector<std::int32_t> favorites(n);
        for (std::int64_t v = 0; v < n; ++v)  // sclp_singleton_favorite_kernel :682-692
            favorites[v] = cluster_weights[clusters[v]] == g.vw[v] ? proposals[v] : kSclpInvalid;
        const auto order = admission_order(favorites, gains, ties);
        const auto candidates = static_cast<std::int64_t>(order.size());
        if (candidates > 0) {
            std::vector<std::uint64_t> ranks(candidates);
            std::vector<std::int32_t> flags(candidates), positions(candidates);
            for (std::int64_t i = 0; i < candidates; ++i) {
                ranks[i] = (i > 0 && order[i].target == order[i - 1].target) ? ranks[i - 1] + 1 : 1;
            }
            std::int64_t pair_count = 0;
            for (std::int64_t i = 0; i < candidates; ++i) {  // :694-704
                const bool second = (ranks[i] & 1ULL) == 0ULL;
                const auto wi = g.vw[order[i].vertex];
                flags[i] = second && i > 0 && g.vw[order[i - 1].vertex] <= capacity - wi ? 1 : 0;
                positions[i] = static_cast<std::int32_t>(pair_count);
                pair_count += flags[i];
            }
            const auto merge_budget = static_cast<std::uint64_t>(
                std::max<std::int64_t>(0, cluster_count - desired_clusters));
            const auto selected = std::min<std::uint64_t>(pair_count, merge_budget);
            if (selected > 0) {
                for (std::int64_t i = 0; i < candidates; ++i) {  // :706-724
                    if (!flags[i] || static_cast<std::uint64_t>(positions[i]) >= merge_budget)
                        continue;
                    const auto v = static_cast<std::int32_t>(order[i].vertex);
                    const auto source = clusters[v];
                    const auto target = clusters[order[i - 1].vertex];
                    if (source == target) continue;
                    cluster_weights[source] -= g.vw[v];
                    cluster_weights[target] += g.vw[v];
                    clusters[v] = target;
                }
            }
        }
        cluster_count = count_nonempty(cluster_weights);
        if (trace) out.steps.push_back({2, 4, 0, clusters});
    }

    // Compaction (:1541-1577)
    std::vector<std::int32_t> compact(n);
    std::int32_t running = 0;
    for (std::int64_t c = 0; c < n; ++c) {
        compact[c] = running;
        running += cluster_weights[c] > 0 ? 1 : 0;
    }
    out.coarse = cluster_count;
    out.map.resize(n);
    out.vw.assign(cluster_count, 0);
    for (std::int64_t v = 0; v < n; ++v) {
        out.map[v] = compact[clusters[v]];
        out.vw[out.map[v]] += g.vw[v];
    }
    out.maximum_weight = 0;
    for (auto w : out.vw) out.maximum_weight = std::max(out.maximum_weight, w);
    if (out.maximum_weight > capacity)
        throw std::runtime_error("SCLP coarsening exceeded cluster capacity");
    return out;
}

Graph contract(const Graph& fine, const std::vector<std::int32_t>& map, std::int64_t coarse,
               const std::vector<std::uint64_t>& coarse_weights) {
    // Closed form of src/coarsen.cu:1674-1814: cross entries grouped by
    // (map[v], map[u]) in ascending order, weights summed, internal entries
    // dropped; rows sorted by neighbour.
    const auto n = fine.n();
    std::vector<std::int64_t> member_start(coarse + 1, 0);
    for (std::int64_t v = 0; v < n; ++v) ++member_start[map[v] + 1];
    for (std::int64_t c = 0; c < coarse; ++c) member_start[c + 1] += member_start[c];
    std::vector<std::int32_t> members(n);
    {
        std::vector<std::int64_t> cursor(member_start.begin(), member_start.end() - 1);
        for (std::int64_t v = 0; v < n; ++v) members[cursor[map[v]]++] = static_cast<std::int32_t>(v);
    }
    std::vector<std::vector<std::pair<std::int32_t, std::uint64_t>>> rows(coarse);
#pragma omp parallel
    {
        std::vector<std::pair<std::int32_t, std::uint64_t>> buffer;
#pragma omp for schedule(dynamic, 256)
        for (std::int64_t c = 0; c < coarse; ++c) {
            buffer.clear();
            for (auto i = member_start[c]; i < member_start[c + 1]; ++i) {
                const auto v = members[i];
                for (auto e = fine.off[v]; e < fine.off[v + 1]; ++e) {
                    const auto t = map[fine.adj[e]];
                    if (t != c) buffer.emplace_back(t, fine.ew[e]);
                }
            }
            std::sort(buffer.begin(), buffer.end(),
                      [](const auto& a, const auto& b) { return a.first < b.first; });
            auto& row = rows[c];
            for (std::size_t i = 0; i < buffer.size();) {
                const auto t = buffer[i].first;
                std::uint64_t w = 0;
                while (i < buffer.size() && buffer[i].first == t) w += buffer[i++].second;
                row.emplace_back(t, w);
            }
        }
    }
    Graph out;
    out.vw = coarse_weights;
    out.vw.resize(coarse);
    out.off.assign(coarse + 1, 0);
    for (std::int64_t c = 0; c < coarse; ++c)
        out.off[c + 1] = out.off[c] + static_cast<std::int64_t>(rows[c].size());
    out.adj.resize(out.off[coarse]);
    out.ew.resize(out.off[coarse]);
    for (std::int64_t c = 0; c < coarse; ++c) {
        auto e = out.off[c];
        for (const auto& [t, w] : rows[c]) {
            out.adj[e] = t;
            out.ew[e] = w;
            ++e;
        }
    }
    return out;
}

Hierarchy coarsen(const Graph& g, const CoarsenOptions& options) {
    // src/coarsen.cu:1906-2164 (coarsen_impl)
    if (options.parts < 2 || options.parts > 32 || options.cluster_size_factor < 2 ||
        options.max_levels < 0 || !std::isfinite(options.stop_ratio) ||
        options.stop_ratio <= 0 || options.stop_ratio > 1)
        throw std::invalid_argument("invalid coarsening options");
    Hierarchy h;
    h.levels.push_back(g);
    const std::int64_t cutoff = std::max<std::int64_t>(
        32, static_cast<std::int64_t>(options.cluster_size_factor) * options.parts);
    int level = 0;
    for (; level < options.max_levels && h.levels.back().n() > cutoff; ++level) {
        const auto level_seed = options.seed + static_cast<std::uint32_t>(level);
        const Graph& current = h.levels.back();
        auto agg = aggregate(current, options.parts, level_seed, level,
                             options.cluster_size_factor, options.pack_isolated);
        const double contraction =
            static_cast<double>(agg.coarse) / static_cast<double>(current.n());
        if (agg.coarse >= current.n() || contraction > options.stop_ratio) {
            h.stop_reason = "insufficient_contraction";
            break;
        }
        Graph coarse = contract(current, agg.map, agg.coarse, agg.vw);
        h.maps.push_back(std::move(agg.map));
        h.levels.push_back(std::move(coarse));
    }
    if (level == options.max_levels && h.levels.back().n() > cutoff) h.stop_reason = "max_levels";
    return h;
}

void write_hierarchy(const Hierarchy& h, const std::string& path) {
    // include/hierarchy.hpp:42-126 (int32 format, tmp + rename)
    if (h.levels.empty() || h.maps.size() + 1 != h.levels.size())
        throw std::runtime_error("cannot write an incomplete hierarchy");
    const auto tmp = path + ".tmp";
    std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("cannot open hierarchy temporary file");
    auto put = [&](std::int64_t value) {
        if (value < std::numeric_limits<std::int32_t>::min() ||
            value > std::numeric_limits<std::int32_t>::max())
            throw std::runtime_error("hierarchy value exceeds int32");
        const auto v = static_cast<std::int32_t>(value);
        out.write(reinterpret_cast<const char*>(&v), sizeof(v));
    };
    put(static_cast<std::int64_t>(h.levels.size()));
    for (std::size_t i = 0; i < h.levels.size(); ++i) {
        const auto& g = h.levels[i];
        put(g.n());
        put(g.m());
        for (auto o : g.off) put(o);
        for (auto u : g.adj) put(u);
        for (auto w : g.ew) put(static_cast<std::int64_t>(w));
        for (auto w : g.vw) put(static_cast<std::int64_t>(w));
        if (i > 0)
            for (auto c : h.maps[i - 1]) put(c);
    }
    out.close();
    if (!out) throw std::runtime_error("failed while writing hierarchy");
    if (std::rename(tmp.c_str(), path.c_str()) != 0)
        throw std::runtime_error("cannot commit hierarchy output");
}

}  // namespace gpo
