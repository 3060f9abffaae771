// oracle_refine.cpp -- CPU restatement of incremental bucketed refinement and
// uncoarsening.
//
// TEST INFRASTRUCTURE ONLY (parity checker / CPU baseline).
//
// Follows src/refine_incremental.cuh, src/refine.cu and src/uncoarsen.cu of the
// reference. The reference ranks vertices inside a rebalance bucket by the
// return value of atomicAdd (src/refine_incremental.cuh:185,216), so its order is
// a scheduling accident. This restatement ranks them in increasing vertex id,
// one legal schedule of those atomics; every other quantity is deterministic.
//
// The rebalance gain bucket (balance_gain_bucket, :167-171) is
// clamp(25 + ceil(log(-g)/log(1.5)), 2, 49) for g < 0. It is evaluated here as
// "smallest j with -g <= 1.5^j" against a table of correctly rounded powers,
// which equals the mathematical definition except within one rounding of an
// exact power 1.5^j (where the reference's answer depends on its libm).
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace gpo {
namespace {

constexpr int kBalanceBuckets = 50;
constexpr int kBalanceSections = 16;
constexpr int kBalanceWidth = kBalanceBuckets * kBalanceSections;

// 1.5^j for j = -23..24, each correctly rounded (exact for j >= 0).
struct PowerTable {
    double value[48];
    PowerTable() {
        for (int j = -23; j <= 24; ++j) {
            double p;
            if (j >= 0) {
                double three = 1.0;
                for (int i = 0; i < j; ++i) three *= 3.0;  // exact: 3^24 < 2^53
                p = std::ldexp(three, -j);
            } else {
                double three = 1.0;
                for (int i = 0; i < -j; ++i) three *= 3.0;
                p = std::ldexp(1.0, -j) / three;  // one correctly rounded division
            }
            value[j + 23] = p;
        }
    }
};
const PowerTable kPowers;

int balance_gain_bucket(double gain) {
    if (gain > 0) return 0;
    if (gain == 0) return 1;
    const double x = -gain;
    // j <= -23 -> bucket 2 ; j >= 24 -> bucket 49
    int lo = 0, hi = 48;  // first index with x <= value[index]; 48 = none
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (x <= kPowers.value[mid]) hi = mid;
        else lo = mid + 1;
    }
    const int j = lo == 48 ? 25 : lo - 23;
    return std::max(2, std::min(49, 25 + j));
}

struct Context {
    const Graph& g;
    int k;
    std::uint64_t capacity;
    std::vector<std::int32_t>& labels;
    std::vector<std::uint64_t>& pw;
    std::vector<std::uint64_t> cache;  // n*k
    std::vector<std::int32_t> targets, locks, flags;
    std::vector<std::int64_t> gains;
};

void build_cache(const Graph& g, const std::vector<std::int32_t>& labels, int k,
                 std::vector<std::uint64_t>& cache) {
    // incremental_build_affinity (:7-25)
    const auto n = g.n();
    cache.assign(static_cast<std::size_t>(n) * k, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (std::int64_t v = 0; v < n; ++v) {
        for (auto e = g.off[v]; e < g.off[v + 1]; ++e) {
            const auto u = g.adj[e];
            if (u != v) cache[v * k + labels[u]] += g.ew[e];
        }
    }
}

// incremental_update_cache_and_cut (:69-95) + incremental_commit (:97-109).
// Returns the exact cut decrease of the batch.
std::int64_t apply(Context& c, const std::vector<std::int32_t>& movers, bool lp) {
    const auto& g = c.g;
    const int k = c.k;
    std::int64_t delta = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : delta)
    for (std::size_t i = 0; i < movers.size(); ++i) {
        const auto v = movers[i];
        const auto source = c.labels[v];
        const auto target = c.targets[v];
        long long gain = 0;
        for (auto e = g.off[v]; e < g.off[v + 1]; ++e) {
            const auto u = g.adj[e];
            if (u == v) continue;
            const auto weight = g.ew[e];
            if (!c.flags[u] || u > v) {
                const auto new_neighbor = c.flags[u] ? c.targets[u] : c.labels[u];
                gain += static_cast<long long>(weight) *
                        ((source != c.labels[u]) - (target != new_neighbor));
            }
        }
        delta += gain;
    }
    // cache updates are sums, applied in any order with identical results
    for (const auto v : movers) {
        const auto source = c.labels[v];
        const auto target = c.targets[v];
        for (auto e = g.off[v]; e < g.off[v + 1]; ++e) {
            const auto u = g.adj[e];
            if (u == v) continue;
            c.cache[static_cast<std::size_t>(u) * k + source] -= g.ew[e];
            c.cache[static_cast<std::size_t>(u) * k + target] += g.ew[e];
        }
    }
    for (const auto v : movers) {
        const auto source = c.labels[v];
        const auto target = c.targets[v];
        const auto w = g.vw[v];
        c.pw[source] -= w;
        c.pw[target] += w;
        c.labels[v] = target;
        if (lp) c.locks[v] = 1;
    }
    return delta;
}

std::uint64_t maximum(const std::vector<std::uint64_t>& v) {
    return *std::max_element(v.begin(), v.end());
}

// One rebalance iteration: incremental_rebalance_proposals (:111-157),
// balance_source_histogram (:173-186), balance_scan_buckets (:188-202),
// balance_select_sources (:204-217), balance_select_destinations (:219-232).
// Sets c.flags and c.targets; returns the mover list (ascending ids).
std::vector<std::int32_t> rebalance_select(Context& c) {
    const auto& g = c.g;
    const auto n = g.n();
    const int k = c.k;
    const auto cap = c.capacity;
    std::vector<std::int32_t> bucket(n, -1);
#pragma omp parallel for schedule(static)
    for (std::int64_t v = 0; v < n; ++v) {
        c.targets[v] = -1;
        const auto source = c.labels[v];
        const auto weight = g.vw[v];
        if (c.pw[source] <= cap) continue;
        int best = -1;
        std::uint64_t best_conn = 0, best_slack = 0;
        for (int p = 0; p < k; ++p) {
            if (p == source || c.pw[p] >= cap || weight > cap - c.pw[p]) continue;
            const auto value = c.cache[v * k + p];
            const auto slack = cap - c.pw[p];
            if (best < 0 || value > best_conn || (value == best_conn && slack > best_slack)) {
                best = p;
                best_conn = value;
                best_slack = slack;
            }
        }
        if (best >= 0 && best_conn == 0) {
            std::uint64_t total_slack = 0;
            for (int p = 0; p < k; ++p)
                if (p != source && c.pw[p] < cap && weight <= cap - c.pw[p])
                    total_slack = UINT64_MAX - total_slack < cap - c.pw[p]
                                      ? UINT64_MAX : total_slack + cap - c.pw[p];
            auto ticket = static_cast<std::uint64_t>(mix32(static_cast<std::uint32_t>(v))) % total_slack;
            for (int p = 0; p < k; ++p) {
                if (p == source || c.pw[p] >= cap || weight > cap - c.pw[p]) continue;
                const auto slack = cap - c.pw[p];
                if (ticket < slack) {
                    best = p;
                    break;
                }
                ticket -= slack;
            }
        }
        if (best >= 0) {
            const auto current = c.cache[v * k + source];
            const double gain = best_conn >= current ? static_cast<double>(best_conn - current)
                                                     : -static_cast<double>(current - best_conn);
            const double gain_per_weight = gain / static_cast<double>(weight);
            bucket[v] = source * kBalanceWidth + balance_gain_bucket(gain_per_weight) * kBalanceSections +
                        static_cast<int>(mix32(static_cast<std::uint32_t>(v)) % kBalanceSections);
            c.targets[v] = best;
        }
    }
    // Source quota: ranks in vertex-id order inside each bucket.
    std::vector<std::uint64_t> counts(static_cast<std::size_t>(k) * kBalanceWidth, 0);
    std::vector<std::uint64_t> rank(n, 0);
    for (std::int64_t v = 0; v < n; ++v) {
        if (bucket[v] < 0) continue;
        rank[v] = counts[bucket[v]];
        counts[bucket[v]] += g.vw[v];
    }
    std::vector<std::uint64_t> offsets(counts.size());
    for (int p = 0; p < k; ++p) {
        std::uint64_t run = 0;
        for (int b = 0; b < kBalanceWidth; ++b) {
            offsets[p * kBalanceWidth + b] = run;
            run += counts[p * kBalanceWidth + b];
        }
    }
    std::fill(counts.begin(), counts.end(), 0);
    for (std::int64_t v = 0; v < n; ++v) {
        const int b = bucket[v];
        if (b < 0) continue;
        const int source = b / kBalanceWidth;
        if (offsets[b] + rank[v] >= c.pw[source] - cap) {
            bucket[v] = -1;
            continue;
        }
        const int destination = c.targets[v] * kBalanceWidth + b % kBalanceWidth;
        bucket[v] = destination;
        rank[v] = counts[destination];
        counts[destination] += g.vw[v];
    }
    for (int p = 0; p < k; ++p) {
        std::uint64_t run = 0;
        for (int b = 0; b < kBalanceWidth; ++b) {
            offsets[p * kBalanceWidth + b] = run;
            run += counts[p * kBalanceWidth + b];
        }
    }
    std::vector<std::int32_t> movers;
    for (std::int64_t v = 0; v < n; ++v) {
        const int b = bucket[v];
        c.flags[v] = 0;
        if (b < 0) continue;
        const auto before = offsets[b] + rank[v];
        const auto slack = cap - c.pw[b / kBalanceWidth];
        c.flags[v] = before <= slack && g.vw[v] <= slack - before;
        if (c.flags[v]) movers.push_back(static_cast<std::int32_t>(v));
    }
    return movers;
}

}  // namespace

std::uint64_t cut(const Graph& g, const std::vector<std::int32_t>& labels) {
    // host_cut, src/graph.cpp:192-212
    std::uint64_t directed = 0;
    const auto n = g.n();
#pragma omp parallel for reduction(+ : directed) schedule(static)
    for (std::int64_t v = 0; v < n; ++v)
        for (auto e = g.off[v]; e < g.off[v + 1]; ++e)
            if (labels[v] != labels[g.adj[e]]) directed += g.ew[e];
    if (directed & 1ULL) throw std::runtime_error("weighted CSR cut is not symmetric");
    return directed / 2;
}

std::vector<std::uint64_t> part_weights(const Graph& g, const std::vector<std::int32_t>& labels,
                                        int parts) {
    std::vector<std::uint64_t> w(parts, 0);
    for (std::int64_t v = 0; v < g.n(); ++v) {
        if (labels[v] < 0 || labels[v] >= parts)
            throw std::invalid_argument("refinement partition has an invalid part id");
        w[labels[v]] += g.vw[v];
    }
    return w;
}

void validate_gain_range(const Graph& g) {
    // validate_refinement_gain_range, src/refine.cu:346-365 (Big width: the
    // type bound INT64_MAX/2 / UINT64_MAX is 0, so the sum is always checked)
    constexpr auto limit = static_cast<std::uint64_t>(std::numeric_limits<std::int64_t>::max()) / 2;
    std::uint64_t directed = 0;
    for (auto w : g.ew) {
        if (w > limit || directed > limit - w)
            throw std::overflow_error("refinement signed gain range exceeded");
        directed += w;
    }
}

void validate_csr(const Graph& g, bool allow_self_loops, bool strict) {
    // src/graph.cpp:124-190
    if (g.off.size() != g.vw.size() + 1 || g.off.empty() || g.off.front() != 0 ||
        g.off.back() != g.m() || g.adj.size() != g.ew.size())
        throw std::runtime_error("invalid weighted CSR shape");
    for (std::int64_t v = 0; v < g.n(); ++v)
        if (g.off[v] > g.off[v + 1]) throw std::runtime_error("weighted CSR offsets are not monotone");
    for (auto u : g.adj)
        if (u < 0 || u >= g.n()) throw std::runtime_error("weighted CSR contains an invalid neighbor");
    if (!allow_self_loops)
        for (std::int64_t v = 0; v < g.n(); ++v)
            for (auto e = g.off[v]; e < g.off[v + 1]; ++e)
                if (g.adj[e] == v) throw std::runtime_error("contracted graph contains a self loop");
    if (!strict) return;
    for (std::int64_t v = 0; v < g.n(); ++v) {
        for (auto e = g.off[v]; e < g.off[v + 1]; ++e) {
            const auto u = g.adj[e];
            const auto b = g.adj.begin() + g.off[u], end = g.adj.begin() + g.off[u + 1];
            const auto r = std::lower_bound(b, end, static_cast<std::int32_t>(v));
            if (r == end || *r != v) throw std::runtime_error("weighted CSR is not bidirectional");
            if (g.ew[e] != g.ew[r - g.adj.begin()])
                throw std::runtime_error("weighted CSR reverse edge weights differ");
        }
    }
}

RefineStats refine(const Graph& g, std::vector<std::int32_t>& labels,
                   std::vector<std::uint64_t>& pw, std::uint64_t capacity, int k,
                   int rounds, int stale_rounds) {
    // run_incremental_refinement, src/refine_incremental.cuh:235-361
    validate_gain_range(g);
    const auto n = g.n();
    Context c{g, k, capacity, labels, pw, {}, {}, {}, {}, {}};
    c.targets.assign(n, -1);
    c.locks.assign(n, 0);
    c.flags.assign(n, 0);
    c.gains.assign(n, 0);
    build_cache(g, labels, k, c.cache);
    RefineStats stats;
    stats.initial_cut = cut(g, labels);
    auto current_cut = stats.initial_cut, best_cut = current_cut;
    std::vector<std::int32_t> best_labels = labels;
    std::vector<std::uint64_t> best_pw = pw;

    auto do_apply = [&](const std::vector<std::int32_t>& movers, bool lp) {
        const auto delta = apply(c, movers, lp);
        if (delta >= 0) {
            if (static_cast<std::uint64_t>(delta) > current_cut)
                throw std::runtime_error("incremental cut underflow");
            current_cut -= static_cast<std::uint64_t>(delta);
        } else {
            current_cut += static_cast<std::uint64_t>(-delta);
        }
        stats.accepted += movers.size();
    };
    auto rebalance = [&]() {
        for (int iteration = 0; iteration < 2 * k + 4 && maximum(pw) > capacity; ++iteration) {
            ++stats.admission_iterations;
            const auto movers = rebalance_select(c);
            if (movers.empty()) return false;
            do_apply(movers, false);
        }
        return maximum(pw) <= capacity;
    };
    std::uint64_t maximum_edge = 0;
    for (auto w : g.ew) maximum_edge = std::max(maximum_edge, w);
    const double filter = maximum_edge <= 1 ? 0.25 : 0.75;
    int stale = 0, empty = 0;
    for (int phase = 0; phase < rounds && stale < stale_rounds; ++phase) {
        stats.rounds = phase + 1;
        // incremental_lp_proposals (:27-46)
#pragma omp parallel for schedule(static)
        for (std::int64_t v = 0; v < n; ++v) {
            c.targets[v] = -1;
            c.gains[v] = std::numeric_limits<std::int64_t>::min();
            if (c.locks[v]) {
                c.locks[v] = 0;
                continue;
            }
            const auto source = labels[v];
            std::uint64_t best = 0;
            int target = -1;
            for (int p = 0; p < k; ++p)
                if (p != source && c.cache[v * k + p] > best) {
                    best = c.cache[v * k + p];
                    target = p;
                }
            const auto current = c.cache[v * k + source];
            if (target >= 0 &&
                (best >= current ||
                 current - best < static_cast<std::uint64_t>(filter * static_cast<double>(current)))) {
                c.targets[v] = target;
                c.gains[v] = best >= current ? static_cast<std::int64_t>(best - current)
                                             : -static_cast<std::int64_t>(current - best);
            }
        }
        // incremental_priority_filter (:48-67)
#pragma omp parallel for schedule(dynamic, 1024)
        for (std::int64_t v = 0; v < n; ++v) {
            if (c.targets[v] < 0) {
                c.flags[v] = 0;
                continue;
            }
            const auto source = labels[v], target = c.targets[v];
            const auto gain = c.gains[v];
            long long correction = 0;
            for (auto e = g.off[v]; e < g.off[v + 1]; ++e) {
                const auto u = g.adj[e];
                if (u == v || c.targets[u] < 0 || !(c.gains[u] > gain || (c.gains[u] == gain && u < v)))
                    continue;
                const auto w = static_cast<long long>(g.ew[e]);
                correction += w * ((c.targets[u] == target) - (c.targets[u] == source) -
                                   (labels[u] == target) + (labels[u] == source));
            }
            c.flags[v] = gain + correction >= 0;
        }
        std::vector<std::int32_t> movers;
        for (std::int64_t v = 0; v < n; ++v)
            if (c.flags[v]) movers.push_back(static_cast<std::int32_t>(v));
        stats.proposals += movers.size();
        if (movers.empty()) {
            if (++empty >= 2) break;
            ++stale;
            continue;
        }
        empty = 0;
        do_apply(movers, true);
        if (!rebalance()) {
            ++stats.admission_fallbacks;
            break;
        }
        ++stale;
        if (current_cut < best_cut) {
            if (static_cast<long double>(current_cut) < 0.999L * best_cut) stale = 0;
            best_cut = current_cut;
            best_labels = labels;
            best_pw = pw;
        }
    }
    labels = best_labels;
    pw = best_pw;
    stats.final_cut = cut(g, labels);
    stats.final_max_part_weight = maximum(pw);
    if (stats.final_cut != best_cut || stats.final_cut > stats.initial_cut || maximum(pw) > capacity)
        throw std::runtime_error("incremental refinement postcondition failed");
    return stats;
}

UncoarsenResult uncoarsen(const Hierarchy& h, const std::vector<std::int32_t>& coarsest,
                          int parts, double ratio, int max_rounds, double growth,
                          int stale_rounds) {
    // refine_hierarchy_impl, src/refine.cu:423-646
    if (parts < 2 || parts > 32 || max_rounds < 0 || max_rounds > 64 || stale_rounds < 1 ||
        stale_rounds > 64 || !std::isfinite(growth) || growth < 1.0 || !std::isfinite(ratio) ||
        ratio < 1.0)
        throw std::invalid_argument("invalid refinement options");
    if (h.levels.empty() || h.maps.size() + 1 != h.levels.size())
        throw std::invalid_argument("resident refinement requires a complete hierarchy");
    for (const auto& level : h.levels) validate_gain_range(level);
    const auto& coarse = h.levels.back();
    if (coarsest.size() != static_cast<std::size_t>(coarse.n()))
        throw std::invalid_argument("invalid coarsest device partition");
    for (auto p : coarsest)
        if (p < 0 || p >= parts) throw std::invalid_argument("invalid coarsest device partition");
    std::vector<std::int32_t> labels = coarsest;
    auto pw = part_weights(coarse, labels, parts);
    std::uint64_t total = 0;
    for (auto w : pw) {
        if (total > std::numeric_limits<std::uint64_t>::max() - w)
            throw std::overflow_error("refinement vertex weight overflow");
        total += w;
    }
    const auto capacity = partition_capacity(total, parts, ratio);
    if (maximum(pw) > capacity) throw std::invalid_argument("initial refinement partition is imbalanced");
    auto stats = refine(coarse, labels, pw, capacity, parts, max_rounds, stale_rounds);
    std::uint64_t last_cut = stats.final_cut;
    auto last_refined = coarse.n();
    for (std::size_t coarse_level = h.levels.size() - 1; coarse_level > 0; --coarse_level) {
        const auto fine_level = coarse_level - 1;
        const auto& fine = h.levels[fine_level];
        const auto& map = h.maps[fine_level];
        if (map.size() != static_cast<std::size_t>(fine.n()))
            throw std::invalid_argument("fine-to-coarse map has wrong length");
        std::vector<std::int32_t> projected(fine.n());
        for (std::int64_t v = 0; v < fine.n(); ++v) {
            const auto c = map[v];
            if (c < 0 || c >= static_cast<std::int64_t>(labels.size()))
                throw std::invalid_argument("fine-to-coarse map contains an invalid id");
            projected[v] = labels[c];
        }
        labels.swap(projected);
        const bool refine_this = fine_level == 0 ||
                                 static_cast<double>(fine.n()) >= growth * static_cast<double>(last_refined);
        if (refine_this) {
            stats = refine(fine, labels, pw, capacity, parts, max_rounds, stale_rounds);
            last_refined = fine.n();
            last_cut = stats.final_cut;
        }
    }
    UncoarsenResult out;
    out.labels = std::move(labels);
    out.final_cut = last_cut;
    out.final_part_weights = pw;
    return out;
}

}  // namespace gpo
