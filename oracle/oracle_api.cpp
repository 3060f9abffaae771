// oracle_api.cpp -- extern "C" surface of the CPU oracle (gpart_oracle.h).
//
// TEST INFRASTRUCTURE ONLY (parity checker / CPU baseline).
#include "gpart_oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#ifdef _OPENMP
#include <omp.h>
#endif

struct gpo_graph {
    gpo::Graph g;
};
struct gpo_hierarchy {
    gpo::Hierarchy h;
    std::vector<gpo_graph> views;
};

namespace {
thread_local std::string last_error;

template <typename F>
int guarded(F&& body) {
    try {
        body();
        return 0;
    } catch (const std::invalid_argument& e) {
        last_error = e.what();
        return GPO_INVALID_ARGUMENT;
    } catch (const std::overflow_error& e) {
        last_error = e.what();
        return GPO_OVERFLOW;
    } catch (const std::exception& e) {
        last_error = e.what();
        return GPO_RUNTIME;
    }
}

void refresh_views(gpo_hierarchy* h) {
    h->views.clear();
    h->views.reserve(h->h.levels.size());
    for (auto& level : h->h.levels) h->views.push_back(gpo_graph{level});
}
}  // namespace

extern "C" {

const char* gpo_last_error(void) { return last_error.c_str(); }

void gpo_set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

int gpo_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int gpo_partition_capacity(uint64_t total, int parts, double ratio, uint64_t* out) {
    return guarded([&] { *out = gpo::partition_capacity(total, parts, ratio); });
}

gpo_graph* gpo_graph_create(int64_t n, int64_t m, const int64_t* offsets,
                            const int32_t* neighbors, const uint64_t* edge_weights,
                            const uint64_t* vertex_weights) {
    auto* out = new gpo_graph;
    out->g.off.assign(offsets, offsets + n + 1);
    out->g.adj.assign(neighbors, neighbors + m);
    if (edge_weights) out->g.ew.assign(edge_weights, edge_weights + m);
    else out->g.ew.assign(m, 1);
    if (vertex_weights) out->g.vw.assign(vertex_weights, vertex_weights + n);
    else out->g.vw.assign(n, 1);
    return out;
}

void gpo_graph_free(gpo_graph* g) { delete g; }
int64_t gpo_graph_n(const gpo_graph* g) { return g->g.n(); }
int64_t gpo_graph_m(const gpo_graph* g) { return g->g.m(); }

void gpo_graph_copy(const gpo_graph* g, int64_t* offsets, int32_t* neighbors,
                    uint64_t* edge_weights, uint64_t* vertex_weights) {
    if (offsets) std::copy(g->g.off.begin(), g->g.off.end(), offsets);
    if (neighbors) std::copy(g->g.adj.begin(), g->g.adj.end(), neighbors);
    if (edge_weights) std::copy(g->g.ew.begin(), g->g.ew.end(), edge_weights);
    if (vertex_weights) std::copy(g->g.vw.begin(), g->g.vw.end(), vertex_weights);
}

int gpo_aggregate(const gpo_graph* g, int parts, uint32_t seed, int level,
                  int cluster_size_factor, int pack_isolated, int32_t* map,
                  uint64_t* coarse_weights, int64_t* coarse, uint64_t* capacity, int* rounds) {
    return guarded([&] {
        auto r = gpo::aggregate(g->g, parts, seed, level, cluster_size_factor, pack_isolated != 0);
        std::copy(r.map.begin(), r.map.end(), map);
        std::copy(r.vw.begin(), r.vw.end(), coarse_weights);
        *coarse = r.coarse;
        *capacity = r.capacity;
        if (rounds) *rounds = r.rounds;
    });
}

int gpo_aggregate_trace(const gpo_graph* g, int parts, uint32_t seed, int level, int cluster_size_factor,
                        int pack_isolated, int max_steps, int32_t* clusters, int32_t* kinds, int32_t* rounds,
                        uint32_t* thresholds, int* steps) {
    return guarded([&] {
        auto r = gpo::aggregate(g->g, parts, seed, level, cluster_size_factor, pack_isolated != 0, true);
        if (static_cast<int>(r.steps.size()) > max_steps) throw std::invalid_argument("too many aggregate steps");
        const auto n = g->g.n();
        for (std::size_t i = 0; i < r.steps.size(); ++i) {
            std::copy(r.steps[i].clusters.begin(), r.steps[i].clusters.end(), clusters + i * n);
            kinds[i] = r.steps[i].kind;
            rounds[i] = r.steps[i].round;
            thresholds[i] = r.steps[i].threshold;
        }
        *steps = static_cast<int>(r.steps.size());
    });
}

int gpo_sclp_proposals(const gpo_graph* g, const int32_t* clusters, uint32_t seed, int level,
                       int round, int filter_roles, uint32_t mover_threshold,
                       int32_t* proposals, uint64_t* gains, uint32_t* ties) {
    return guarded([&] {
        std::vector<std::int32_t> c(clusters, clusters + g->g.n()), p;
        std::vector<std::uint64_t> gn;
        std::vector<std::uint32_t> t;
        gpo::sclp_proposals(g->g, c, seed, level, round, filter_roles != 0, mover_threshold, p, gn, t);
        std::copy(p.begin(), p.end(), proposals);
        std::copy(gn.begin(), gn.end(), gains);
        std::copy(t.begin(), t.end(), ties);
    });
}

gpo_graph* gpo_contract(const gpo_graph* fine, const int32_t* map, int64_t coarse,
                        const uint64_t* coarse_weights) {
    gpo_graph* out = nullptr;
    const int rc = guarded([&] {
        std::vector<std::int32_t> m(map, map + fine->g.n());
        std::vector<std::uint64_t> w(coarse_weights, coarse_weights + coarse);
        out = new gpo_graph{gpo::contract(fine->g, m, coarse, w)};
    });
    return rc == 0 ? out : nullptr;
}

gpo_hierarchy* gpo_coarsen(const gpo_graph* g, int parts, uint32_t seed, double stop_ratio,
                           int max_levels, int cluster_size_factor, int pack_isolated) {
    gpo_hierarchy* out = nullptr;
    const int rc = guarded([&] {
        gpo::CoarsenOptions o;
        o.parts = parts;
        o.seed = seed;
        o.stop_ratio = stop_ratio;
        o.max_levels = max_levels;
        o.cluster_size_factor = cluster_size_factor;
        o.pack_isolated = pack_isolated != 0;
        out = new gpo_hierarchy{gpo::coarsen(g->g, o), {}};
        refresh_views(out);
    });
    return rc == 0 ? out : nullptr;
}

int gpo_hierarchy_levels(const gpo_hierarchy* h) { return static_cast<int>(h->h.levels.size()); }
const gpo_graph* gpo_hierarchy_level(const gpo_hierarchy* h, int level) { return &h->views[level]; }
const int32_t* gpo_hierarchy_map(const gpo_hierarchy* h, int level) { return h->h.maps[level].data(); }
const char* gpo_hierarchy_stop_reason(const gpo_hierarchy* h) { return h->h.stop_reason.c_str(); }

int gpo_write_hierarchy(const gpo_hierarchy* h, const char* path) {
    return guarded([&] { gpo::write_hierarchy(h->h, path); });
}

void gpo_hierarchy_free(gpo_hierarchy* h) { delete h; }

int gpo_initial_partition(const gpo_graph* g, int parts, double ratio, uint32_t seed,
                          int32_t* labels) {
    return gpo_initial_partition_trace(g, parts, ratio, seed, labels, nullptr, nullptr);
}

int gpo_initial_partition_trace(const gpo_graph* g, int parts, double ratio, uint32_t seed,
                                int32_t* labels, double* coords, int* tasks) {
    return guarded([&] {
        std::vector<double> x;
        auto l = gpo::initial_partition(g->g, parts, ratio, seed, coords ? &x : nullptr, tasks);
        std::copy(l.begin(), l.end(), labels);
        if (coords) std::copy(x.begin(), x.end(), coords);
    });
}

int gpo_refine(const gpo_graph* g, int32_t* labels, int parts, double ratio, int max_rounds,
               int stale_rounds, gpo_refine_stats* stats) {
    // refine_partition (src/refine.cu:410-421) with RefineDeviceContext's checks
    // (src/refine.cu:259-283)
    return guarded([&] {
        if (parts < 2 || parts > 32 || max_rounds < 0 || max_rounds > 64 || stale_rounds < 1 ||
            stale_rounds > 64 || !std::isfinite(ratio) || ratio < 1.0 || g->g.n() <= 0)
            throw std::invalid_argument("invalid refinement options");
        std::vector<std::int32_t> l(labels, labels + g->g.n());
        auto pw = gpo::part_weights(g->g, l, parts);
        std::uint64_t total = 0;
        for (auto w : pw) total += w;
        const auto cap = gpo::partition_capacity(total, parts, ratio);
        if (*std::max_element(pw.begin(), pw.end()) > cap)
            throw std::invalid_argument("initial refinement partition is imbalanced");
        auto s = gpo::refine(g->g, l, pw, cap, parts, max_rounds, stale_rounds);
        std::copy(l.begin(), l.end(), labels);
        if (stats) {
            stats->rounds = s.rounds;
            stats->proposals = s.proposals;
            stats->accepted = s.accepted;
            stats->initial_cut = s.initial_cut;
            stats->final_cut = s.final_cut;
            stats->final_max_part_weight = s.final_max_part_weight;
            stats->admission_iterations = s.admission_iterations;
            stats->admission_fallbacks = s.admission_fallbacks;
        }
    });
}

int gpo_uncoarsen(gpo_hierarchy* h, const int32_t* coarsest_labels, int parts, double ratio,
                  int max_rounds, double growth, int32_t* labels, uint64_t* final_cut) {
    return guarded([&] {
        const auto& c = h->h.levels.back();
        std::vector<std::int32_t> l(coarsest_labels, coarsest_labels + c.n());
        auto r = gpo::uncoarsen(h->h, l, parts, ratio, max_rounds, growth);
        std::copy(r.labels.begin(), r.labels.end(), labels);
        if (final_cut) *final_cut = r.final_cut;
    });
}

int gpo_partition(const gpo_graph* g, int parts, double ratio, uint32_t seed, int32_t* labels,
                  gpo_partition_result* result) {
    // app/partition_main.cu:42-93 production parameters
    return guarded([&] {
        using clock = std::chrono::steady_clock;
        const auto capacity = gpo::partition_capacity(static_cast<std::uint64_t>(g->g.n()), parts, ratio);
        (void)capacity;
        const auto t0 = clock::now();
        gpo::CoarsenOptions o;
        o.parts = parts;
        o.seed = seed;
        o.cluster_size_factor = std::max(64, (512 + parts - 1) / parts);
        o.pack_isolated = true;
        auto h = gpo::coarsen(g->g, o);
        const auto t1 = clock::now();
        auto coarse_labels = gpo::initial_partition(h.levels.back(), parts, ratio, seed);
        const auto t2 = clock::now();
        auto r = gpo::uncoarsen(h, coarse_labels, parts, ratio, 32, 2.0);
        const auto t3 = clock::now();
        std::copy(r.labels.begin(), r.labels.end(), labels);
        if (result) {
            std::uint64_t total = 0;
            for (auto w : r.final_part_weights) total += w;
            result->final_cut = r.final_cut;
            result->max_part_weight =
                *std::max_element(r.final_part_weights.begin(), r.final_part_weights.end());
            result->capacity = gpo::partition_capacity(total, parts, ratio);
            result->levels = static_cast<int>(h.levels.size());
            result->coarsest_vertices = h.levels.back().n();
            result->initial_cut = gpo::cut(h.levels.back(), coarse_labels);
            result->coarsen_seconds = std::chrono::duration<double>(t1 - t0).count();
            result->initial_seconds = std::chrono::duration<double>(t2 - t1).count();
            result->uncoarsen_seconds = std::chrono::duration<double>(t3 - t2).count();
            result->total_seconds = std::chrono::duration<double>(t3 - t0).count();
        }
    });
}

int gpo_cut(const gpo_graph* g, const int32_t* labels, uint64_t* out) {
    return guarded([&] {
        std::vector<std::int32_t> l(labels, labels + g->g.n());
        *out = gpo::cut(g->g, l);
    });
}

int gpo_part_weights(const gpo_graph* g, const int32_t* labels, int parts, uint64_t* weights) {
    return guarded([&] {
        std::vector<std::int32_t> l(labels, labels + g->g.n());
        auto w = gpo::part_weights(g->g, l, parts);
        std::copy(w.begin(), w.end(), weights);
    });
}

int gpo_validate_csr(const gpo_graph* g, int allow_self_loops, int strict) {
    return guarded([&] { gpo::validate_csr(g->g, allow_self_loops != 0, strict != 0); });
}

}  // extern "C"
