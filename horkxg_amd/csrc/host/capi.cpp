// capi.cpp -- the high-level C ABI (gpart.h) over the C++ API.
//
// gpart_graph_partition reproduces the CLI pipeline of app/partition_main.cu:42-119:
// production coarsening options, initial partition on the coarsest level,
// uncoarsening with refinement, hierarchy release and allocator trim, with the
// same timed scope (coarsen start -> memory released).
#include "gpart.h"

#include "coarsen.hpp"
#include "initial_partition.hpp"
#include "partition_capacity.hpp"
#include "quality.hpp"
#include "uncoarsen.hpp"
#include "../tuned_gfx950.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <variant>
#include <vector>

namespace {

template <typename Types>
struct GraphOf {
    gpart::WeightedGraph<Types> host;
    gpart::DeviceWeightedGraph<Types> device;
    bool host_ready = true;  // false for graphs generated on the device until a host copy is needed
};

// The host copy, made from the device copy on first use.
template <typename Types>
gpart::WeightedGraph<Types>& host_of(GraphOf<Types>& g) {
    if (!g.host_ready) {
        g.host = gpart::copy_device_weighted(g.device);
        g.host_ready = true;
    }
    return g.host;
}

thread_local std::string capi_error;

template <typename F>
int guard(F&& body) {
    try {
        body();
        return GPART_OK;
    } catch (const std::invalid_argument& e) {
        capi_error = e.what();
        gpart_hip_set_error(capi_error.c_str());
        return GPART_ERR_INVALID_ARGUMENT;
    } catch (const std::overflow_error& e) {
        capi_error = e.what();
        gpart_hip_set_error(capi_error.c_str());
        return GPART_ERR_OVERFLOW;
    } catch (const std::exception& e) {
        capi_error = e.what();
        gpart_hip_set_error(capi_error.c_str());
        return GPART_ERR_RUNTIME;
    }
}

template <typename Types>
gpart::WeightedGraph<Types> host_graph(std::int64_t n, const std::int64_t* indptr, std::int64_t m,
                                       const std::int64_t* indices, const std::uint64_t* ew,
                                       const std::uint64_t* vw) {
    auto g = gpart::graph_from_int64<Types>(n, indptr, m, indices);
    using W = typename Types::WeightT;
    if (ew)
        for (std::int64_t e = 0; e < m; ++e) {
            if (static_cast<std::uint64_t>(static_cast<W>(ew[e])) != ew[e])
                throw std::overflow_error("edge weight exceeds the graph width");
            g.edge_weights[e] = static_cast<W>(ew[e]);
        }
    if (vw)
        for (std::int64_t v = 0; v < n; ++v) {
            if (static_cast<std::uint64_t>(static_cast<W>(vw[v])) != vw[v])
                throw std::overflow_error("vertex weight exceeds the graph width");
            g.vertex_weights[v] = static_cast<W>(vw[v]);
        }
    return g;
}

template <typename Types>
void copy_level(const gpart::WeightedGraph<Types>& g, int64_t* offsets, int32_t* neighbors, uint64_t* ew,
                uint64_t* vw) {
    if (offsets) std::copy(g.offsets.begin(), g.offsets.end(), offsets);
    if (neighbors) std::copy(g.neighbors.begin(), g.neighbors.end(), neighbors);
    if (ew) std::copy(g.edge_weights.begin(), g.edge_weights.end(), ew);
    if (vw) std::copy(g.vertex_weights.begin(), g.vertex_weights.end(), vw);
}

// One partition with the CLI's production options, in three phases so that a
// batch can run the initial partitions of several graphs in one launch. The run
// ends, inside the timed scope, by returning every block of the hierarchy to the
// context's cache (the reference's per-run scratch scope); the cache keeps its
// device segments for the next run.
using Clock = std::chrono::steady_clock;

template <typename Types>
struct PipelineRun {
    gpart::DeviceHierarchy<Types> hierarchy;
    std::uint64_t total = 0;
    Clock::time_point t0, t1, t2;
    std::uint64_t initial_cut = 0;
    std::int64_t coarsest_vertices = 0;
};

template <typename Types>
PipelineRun<Types> coarsen_phase(GraphOf<Types>& g, int parts, double ratio, uint32_t seed, int verify) {
    auto* ctx = gpart::detail::context();
    PipelineRun<Types> run;
    std::uint64_t heaviest = 0, lightest = 0;
    {
        const auto view = g.device.view();
        GPART_CHECK(gpart_vertex_weight_stats(ctx, &view, &run.total, &heaviest, &lightest));
    }
    gpart::partition_capacity(run.total, parts, ratio);
    gpart::CoarsenOptions co;
    co.parts = parts;
    co.seed = seed;
    co.strict_verify = verify != 0;
    co.cluster_size_factor = std::max(64, (512 + parts - 1) / parts);
    co.pack_isolated = true;
    // The hierarchy's level 0 is a read-only view of the caller's device graph: no
    // copy, and the input's memory is not held twice.
    gpart::DeviceWeightedGraph<Types> input;
    input.offsets = decltype(input.offsets)::borrow(g.device.offsets);
    input.neighbors = decltype(input.neighbors)::borrow(g.device.neighbors);
    input.edge_weights = decltype(input.edge_weights)::borrow(g.device.edge_weights);
    input.vertex_weights = decltype(input.vertex_weights)::borrow(g.device.vertex_weights);
    GPART_CHECK(gpart_ctx_synchronize(ctx));
    run.t0 = Clock::now();
    run.hierarchy = gpart::coarsen<Types>(std::move(input), co);
    GPART_CHECK(gpart_ctx_synchronize(ctx));
    run.t1 = Clock::now();
    run.coarsest_vertices = run.hierarchy.levels.back().vertices();
    return run;
}

// The initial partition's cut on the coarsest level and the end of its phase.
template <typename Types>
void note_initial(PipelineRun<Types>& run, const gpart::DeviceVector<std::int32_t>& coarse) {
    auto* ctx = gpart::detail::context();
    const auto view = run.hierarchy.levels.back().view();
    GPART_CHECK(gpart_cut(ctx, &view, coarse.data(), &run.initial_cut));
    GPART_CHECK(gpart_ctx_synchronize(ctx));
    run.t2 = Clock::now();
}

template <typename Types>
gpart::DeviceVector<std::int32_t> uncoarsen_phase(PipelineRun<Types>& run, gpart::DeviceVector<std::int32_t> coarse,
                                                  int parts, double ratio, int verify,
                                                  gpart_partition_report* report) {
    auto* ctx = gpart::detail::context();
    const int levels = static_cast<int>(run.hierarchy.levels.size());
    gpart::RefineOptions ro;
    ro.parts = parts;
    ro.imbalance_ratio = ratio;
    ro.max_rounds = 32;
    ro.strict_verify = verify != 0;
    ro.refinement_growth_ratio = 2.0;
    auto result = gpart::uncoarsen<Types>(run.hierarchy, std::move(coarse), ro);
    GPART_CHECK(gpart_ctx_synchronize(ctx));
    const auto t3 = Clock::now();
    run.hierarchy.levels.clear();
    run.hierarchy.fine_to_coarse.clear();
    auto partition = std::move(result.device_partition);
    const auto t4 = Clock::now();
    if (report) {
        report->final_cut = result.final_cut;
        report->max_part_weight =
            *std::max_element(result.final_part_weights.begin(), result.final_part_weights.end());
        report->capacity = gpart::partition_capacity(run.total, parts, ratio);
        report->initial_cut = run.initial_cut;
        report->coarsest_vertices = run.coarsest_vertices;
        report->levels = levels;
        int refined = 0;
        for (const auto& l : result.levels) refined += l.refinement.rounds > 0 ? 1 : 0;
        report->refined_levels = refined;
        auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
        report->coarsen_seconds = secs(run.t0, run.t1);
        report->initial_seconds = secs(run.t1, run.t2);
        report->uncoarsen_seconds = secs(run.t2, t3);
        report->free_seconds = secs(t3, t4);
        report->algorithm_seconds = secs(run.t0, t4);
    }
    return partition;
}

template <typename Types>
gpart::DeviceVector<std::int32_t> run_pipeline(GraphOf<Types>& g, int parts, double ratio, uint32_t seed, int verify,
                                               int mode, gpart_partition_report* report) {
    auto run = coarsen_phase(g, parts, ratio, seed, verify);
    auto coarse = gpart::initial_partition_gpu<Types>(run.hierarchy.levels.back(), parts, ratio, seed,
                                                      static_cast<gpart::SpectralMode>(mode));
    note_initial(run, coarse);
    return uncoarsen_phase(run, std::move(coarse), parts, ratio, verify, report);
}

// A group of graphs of one width: every graph is coarsened, the initial
// partitions of all of them run in one persistent launch, then each graph is
// uncoarsened. Each graph's labels are those run_pipeline gives it; its report
// carries its own coarsening and uncoarsening time and the group's initial time
// from its coarsening end (the hierarchies of the group are alive together).
template <typename Types>
std::vector<gpart::DeviceVector<std::int32_t>> run_group(const std::vector<GraphOf<Types>*>& graphs,
                                                         const std::vector<int>& parts,
                                                         const std::vector<double>& ratios,
                                                         const std::vector<std::uint32_t>& seeds, int mode,
                                                         const std::vector<gpart_partition_report*>& reports) {
    std::vector<PipelineRun<Types>> runs;
    runs.reserve(graphs.size());
    for (std::size_t i = 0; i < graphs.size(); ++i)
        runs.push_back(coarsen_phase(*graphs[i], parts[i], ratios[i], seeds[i], 0));
    std::vector<const gpart::DeviceWeightedGraph<Types>*> coarsest;
    for (auto& r : runs) coarsest.push_back(&r.hierarchy.levels.back());
    auto coarse = gpart::initial_partition_group<Types>(coarsest, parts, ratios, seeds,
                                                        static_cast<gpart::SpectralMode>(mode));
    for (std::size_t i = 0; i < runs.size(); ++i) note_initial(runs[i], coarse[i]);
    std::vector<gpart::DeviceVector<std::int32_t>> out;
    for (std::size_t i = 0; i < runs.size(); ++i)
        out.push_back(uncoarsen_phase(runs[i], std::move(coarse[i]), parts[i], ratios[i], 0, reports[i]));
    return out;
}

std::atomic<int> batch_group{gpart_tuned::kBatchGroup};  // graphs per group of gpart_partition_batch

template <typename Types>
void partition_impl(GraphOf<Types>& g, int parts, double ratio, uint32_t seed, int verify, int mode, int32_t* labels,
                    gpart_partition_report* report) {
    auto partition = run_pipeline(g, parts, ratio, seed, verify, mode, report);
    if (labels) {
        const auto host = partition.to_host();
        std::copy(host.begin(), host.end(), labels);
    }
}

}  // namespace

struct gpart_graph {
    int width;
    std::variant<GraphOf<gpart::StandardTypes>, GraphOf<gpart::BigTypes>> g;
};

struct gpart_hierarchy {
    int width;
    std::variant<gpart::Hierarchy<gpart::StandardTypes>, gpart::Hierarchy<gpart::BigTypes>> h;
};

namespace {
// A device CSR from the kernel library's generators, adopted with unit weights.
template <typename Types>
GraphOf<Types> adopt_generated(void* offsets, std::int32_t* neighbors, std::int64_t n, std::int64_t m) {
    using O = typename Types::OffsetT;
    using W = typename Types::WeightT;
    GraphOf<Types> holder;
    holder.device.offsets = gpart::DeviceVector<O>::adopt(static_cast<O*>(offsets), static_cast<std::size_t>(n + 1));
    holder.device.neighbors = gpart::DeviceVector<std::int32_t>::adopt(neighbors, static_cast<std::size_t>(m));
    holder.device.edge_weights = gpart::DeviceVector<W>(static_cast<std::size_t>(m), W{1});
    holder.device.vertex_weights = gpart::DeviceVector<W>(static_cast<std::size_t>(n), W{1});
    holder.host_ready = false;
    return holder;
}

template <typename F>
int generate_into(int width, gpart_graph** out, F&& generate) {
    return guard([&] {
        void* offsets = nullptr;
        std::int32_t* neighbors = nullptr;
        std::int64_t n = 0, m = 0;
        auto* ctx = gpart::detail::context();
        GPART_CHECK(generate(ctx, &offsets, &neighbors, &n, &m));
        auto* g = new gpart_graph{width, GraphOf<gpart::StandardTypes>{}};
        try {
            if (width == GPART_WIDTH_STANDARD) g->g = adopt_generated<gpart::StandardTypes>(offsets, neighbors, n, m);
            else g->g = adopt_generated<gpart::BigTypes>(offsets, neighbors, n, m);
            GPART_CHECK(gpart_ctx_synchronize(ctx));
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}
}  // namespace

extern "C" {

int gpart_partition_capacity_c(uint64_t total, int parts, double ratio, uint64_t* out) {
    return guard([&] { *out = gpart::partition_capacity(total, parts, ratio); });
}

gpart_ctx* gpart_thread_context(void) {
    gpart_ctx* ctx = nullptr;
    guard([&] { ctx = gpart::detail::context(); });
    return ctx;
}

int gpart_trim(void) {
    return guard([&] { gpart::detail::release_context_cache(); });
}

int gpart_graph_upload(int width, int64_t n, const int64_t* indptr, int64_t m, const int64_t* indices,
                       const uint64_t* edge_weights, const uint64_t* vertex_weights, int validate, gpart_graph** out) {
    return guard([&] {
        auto* g = new gpart_graph{width, GraphOf<gpart::StandardTypes>{}};
        try {
            auto build = [&](auto tag) {
                using T = decltype(tag);
                GraphOf<T> holder;
                if (!edge_weights && !vertex_weights && !validate) {
                    // unit weights, shape checks only: narrowed on the device
                    holder.device = gpart::device_graph_from_int64<T>(n, indptr, m, indices);
                    holder.host_ready = false;
                } else {
                    holder.host = host_graph<T>(n, indptr, m, indices, edge_weights, vertex_weights);
                    if (validate) gpart::validate_weighted_csr(holder.host, true);
                    else gpart::validate_weighted_shape(holder.host, true);
                    holder.device = gpart::make_device_weighted(holder.host);
                }
                g->g = std::move(holder);
            };
            if (width == GPART_WIDTH_STANDARD) build(gpart::StandardTypes{});
            else if (width == GPART_WIDTH_BIG) build(gpart::BigTypes{});
            else throw std::invalid_argument("unknown graph width");
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int gpart_graph_generate_rmat(int width, int scale, int64_t edge_factor, double a, double b, double c, uint64_t seed,
                              gpart_graph** out) {
    return generate_into(width, out, [&](gpart_ctx* ctx, void** o, int32_t** nb, int64_t* n, int64_t* m) {
        return gpart_generate_rmat_device(ctx, width, scale, edge_factor, a, b, c, seed, o, nb, n, m);
    });
}

int gpart_graph_arrays(gpart_graph* g, int64_t* offsets, int32_t* neighbors, uint64_t* edge_weights,
                       uint64_t* vertex_weights) {
    return guard([&] {
        std::visit([&](auto& x) { copy_level(host_of(x), offsets, neighbors, edge_weights, vertex_weights); }, g->g);
    });
}

int gpart_graph_generate_lattice(int width, int64_t nx, int64_t ny, int64_t nz, gpart_graph** out) {
    return generate_into(width, out, [&](gpart_ctx* ctx, void** o, int32_t** nb, int64_t* n, int64_t* m) {
        return gpart_generate_lattice_device(ctx, width, nx, ny, nz, o, nb, n, m);
    });
}

int gpart_graph_load(int width, const char* indptr_path, const char* indices_path, int validate, gpart_graph** out) {
    return guard([&] {
        auto* g = new gpart_graph{width, GraphOf<gpart::StandardTypes>{}};
        try {
            auto build = [&](auto tag) {
                using T = decltype(tag);
                GraphOf<T> holder;
                if (!validate) {
                    holder.device = gpart::load_device_graph<T>(indptr_path, indices_path);
                    holder.host_ready = false;
                } else {
                    holder.host = gpart::load_weighted_graph<T>(indptr_path, indices_path);
                    gpart::validate_weighted_csr(holder.host, true);
                    holder.device = gpart::make_device_weighted(holder.host);
                }
                g->g = std::move(holder);
            };
            if (width == GPART_WIDTH_STANDARD) build(gpart::StandardTypes{});
            else if (width == GPART_WIDTH_BIG) build(gpart::BigTypes{});
            else throw std::invalid_argument("unknown graph width");
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

void gpart_graph_free(gpart_graph* g) { delete g; }

int64_t gpart_graph_vertices(const gpart_graph* g) {
    return std::visit([](const auto& x) { return x.device.vertices(); }, g->g);
}

int64_t gpart_graph_edges(const gpart_graph* g) {
    return std::visit([](const auto& x) { return x.device.edges(); }, g->g);
}

int gpart_graph_cut(gpart_graph* g, const int32_t* labels, uint64_t* cut) {
    return guard([&] {
        std::visit(
            [&](auto& x) {
                gpart::DeviceVector<std::int32_t> d(std::vector<std::int32_t>(labels, labels + x.device.vertices()));
                const auto view = x.device.view();
                GPART_CHECK(gpart_cut(gpart::detail::context(), &view, d.data(), cut));
            },
            g->g);
    });
}

int gpart_decimal_capacity(const char* imbalance, int64_t n, int parts, int64_t* capacity) {
    return guard([&] {
        if (!imbalance || !capacity) throw std::invalid_argument("null imbalance or output");
        if (!gpart::decimal_capacity(imbalance, n, parts, capacity))
            throw std::invalid_argument("invalid parts or imbalance ratio");
    });
}

int gpart_graph_verify(gpart_graph* g, const int32_t* labels, int parts, const char* imbalance,
                       gpart_verify_report* out, int64_t* loads) {
    return guard([&] {
        if (!out || !imbalance) throw std::invalid_argument("null output or imbalance");
        std::visit(
            [&](auto& x) {
                const auto n = x.device.vertices();
                if (n > 0 && !labels) throw std::invalid_argument("null labels");
                gpart::DeviceVector<std::int32_t> d(std::vector<std::int32_t>(labels, labels + n));
                const auto view = x.device.view();
                const auto r = gpart::verify_partition(gpart::detail::context(), view, d.data(), labels, parts,
                                                       imbalance);
                *out = gpart_verify_report{r.vertices, r.undirected_edges, r.cut, r.cut_ratio, r.max_load,
                                           r.max_load_ratio, r.capacity, r.capacity_ok ? 1 : 0, parts};
                if (loads) std::copy(r.loads.begin(), r.loads.end(), loads);
            },
            g->g);
    });
}

int gpart_graph_analyze(gpart_graph* g, const int32_t* labels, int parts, gpart_quality_report* out,
                        uint64_t* per_part, int per_part_capacity) {
    return guard([&] {
        if (!out) throw std::invalid_argument("null output");
        std::visit(
            [&](auto& x) {
                const auto n = x.device.vertices();
                if (n > 0 && !labels) throw std::invalid_argument("null labels");
                gpart::DeviceVector<std::int32_t> d(std::vector<std::int32_t>(labels, labels + n));
                const auto view = x.device.view();
                const auto r = gpart::analyze_partition(gpart::detail::context(), view, d.data(), parts);
                if (per_part && per_part_capacity < r.parts)
                    throw std::invalid_argument("per-part buffer holds fewer parts than the labelling");
                *out = gpart_quality_report{};
                out->parts = r.parts;
                out->vertices = r.vertices;
                out->directed_edges = r.directed_edges;
                out->edge_cut = r.edge_cut;
                out->weighted_edge_cut = r.weighted_edge_cut;
                out->edge_cut_ratio = r.edge_cut_ratio;
                out->avg_local_cut = r.avg_local_cut;
                out->max_local_cut = r.max_local_cut;
                out->weighted_avg_local_cut = r.weighted_avg_local_cut;
                out->weighted_max_local_cut = r.weighted_max_local_cut;
                out->vertex_avg = r.vertex_avg;
                out->min_vertices = r.min_vertices;
                out->max_vertices = r.max_vertices;
                out->vertex_imbalance = r.vertex_imbalance;
                out->vertex_max_deviation = r.vertex_max_deviation;
                out->edge_set_avg = r.edge_set_avg;
                out->min_edge_set = r.min_edge_set;
                out->max_edge_set = r.max_edge_set;
                out->edge_set_imbalance = r.edge_set_imbalance;
                out->edge_set_max_deviation = r.edge_set_max_deviation;
                if (per_part) std::copy(r.per_part.begin(), r.per_part.end(), per_part);
            },
            g->g);
    });
}

int gpart_graph_partition(gpart_graph* g, int parts, double ratio, uint32_t seed, int verify, int spectral_mode,
                          int32_t* labels, gpart_partition_report* report) {
    return guard([&] {
        if (spectral_mode < 0 || spectral_mode > 2) throw std::invalid_argument("unknown spectral mode");
        std::visit([&](auto& x) { partition_impl(x, parts, ratio, seed, verify, spectral_mode, labels, report); }, g->g);
    });
}

int gpart_partition_batch(int count, gpart_graph* const* graphs, const int* parts, const double* ratios,
                          const uint32_t* seeds, int spectral_mode, int workers, int32_t* const* labels,
                          gpart_partition_report* reports, double* wall_seconds) {
    return guard([&] {
        using clock = std::chrono::steady_clock;
        if (count < 0 || workers < 1) throw std::invalid_argument("batch needs count >= 0 and workers >= 1");
        // Concurrent workers are switched off: the full 256-graph mix with four
        // worker threads ended in a host segmentation fault whose cause is not
        // established yet. One worker runs the queue in the calling thread.
        if (workers > 1) throw std::invalid_argument("concurrent batch workers are disabled (workers must be 1)");
        if (spectral_mode < 0 || spectral_mode > 2) throw std::invalid_argument("unknown spectral mode");
        for (int i = 0; i < count; ++i)
            if (!graphs[i]) throw std::invalid_argument("null graph in batch");
        // largest first, so that the last graphs to start are the short ones
        std::vector<int> order(static_cast<std::size_t>(count));
        for (int i = 0; i < count; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return gpart_graph_edges(graphs[a]) + gpart_graph_vertices(graphs[a]) >
                   gpart_graph_edges(graphs[b]) + gpart_graph_vertices(graphs[b]);
        });
        GPART_CHECK(gpart_ctx_synchronize(gpart::detail::context()));  // uploads complete
        // Round graphs pay for a recording per level of every graph and are replayed
        // only within that graph here: the batch launches rounds directly.
        auto* caller_ctx = gpart::detail::context();
        const int graphs_were = gpart_refine_graphs_enabled(caller_ctx);
        GPART_CHECK(gpart_refine_set_graphs(caller_ctx, 0));
        struct RestoreGraphs {
            gpart_ctx* ctx;
            int enabled;
            ~RestoreGraphs() { (void)gpart_refine_set_graphs(ctx, enabled); }
        } restore_graphs{caller_ctx, graphs_were};
        const int threads = std::max(1, std::min(workers, count));
        std::atomic<int> next{0};
        std::atomic<bool> failed{false};
        std::mutex mu;
        std::condition_variable all_done;
        int finished = 0;
        std::exception_ptr error;
        clock::time_point wall_end;
        const auto wall_start = clock::now();
        auto fail = [&]() {
            std::lock_guard<std::mutex> lock(mu);
            if (!error) error = std::current_exception();
            failed = true;
        };
        auto run = [&]() {
            std::vector<std::pair<int, gpart::DeviceVector<std::int32_t>>> done;
            try {
                auto* ctx = gpart::detail::context();
                const int group = spectral_mode == 0 ? 1 : std::max(1, batch_group.load());
                for (;;) {
                    if (failed.load()) break;
                    // up to `group` consecutive graphs of one width
                    const int slot = next.fetch_add(1);
                    if (slot >= count) break;
                    std::vector<int> claim{order[static_cast<std::size_t>(slot)]};
                    const int width = graphs[claim[0]]->width;
                    while (static_cast<int>(claim.size()) < group) {
                        int j = next.load();
                        if (j >= count || graphs[order[static_cast<std::size_t>(j)]]->width != width) break;
                        if (next.compare_exchange_strong(j, j + 1)) claim.push_back(order[static_cast<std::size_t>(j)]);
                    }
                    if (claim.size() == 1) {
                        const int i = claim[0];
                        auto part = std::visit(
                            [&](auto& x) {
                                return run_pipeline(x, parts[i], ratios[i], seeds[i], 0, spectral_mode,
                                                    reports ? reports + i : nullptr);
                            },
                            graphs[i]->g);
                        done.emplace_back(i, std::move(part));
                        continue;
                    }
                    std::visit(
                        [&](auto& first) {
                            using G = std::decay_t<decltype(first)>;
                            std::vector<G*> members;
                            std::vector<int> p;
                            std::vector<double> r;
                            std::vector<std::uint32_t> sd;
                            std::vector<gpart_partition_report*> rep;
                            for (const int i : claim) {
                                members.push_back(&std::get<G>(graphs[i]->g));
                                p.push_back(parts[i]);
                                r.push_back(ratios[i]);
                                sd.push_back(seeds[i]);
                                rep.push_back(reports ? reports + i : nullptr);
                            }
                            auto labels_of = run_group(members, p, r, sd, spectral_mode, rep);
                            for (std::size_t c = 0; c < claim.size(); ++c)
                                done.emplace_back(claim[c], std::move(labels_of[c]));
                        },
                        graphs[claim[0]]->g);
                }
                GPART_CHECK(gpart_ctx_synchronize(ctx));
            } catch (...) {
                fail();
            }
            {
                // the wall clock stops when every worker's last labels are on the device
                std::unique_lock<std::mutex> lock(mu);
                if (++finished == threads) {
                    wall_end = clock::now();
                    all_done.notify_all();
                } else {
                    all_done.wait(lock, [&] { return finished == threads; });
                }
            }
            try {
                if (!failed.load() && labels)
                    for (auto& [i, part] : done)
                        if (labels[i]) {
                            const auto host = part.to_host();
                            std::copy(host.begin(), host.end(), labels[i]);
                        }
            } catch (...) {
                fail();
            }
        };
        // Each worker owns a context for its lifetime and destroys it before the
        // thread ends (never from a thread_local destructor).
        auto work = [&]() {
            gpart_ctx* ctx = nullptr;
            if (gpart_ctx_create(0, &ctx) != GPART_OK) {
                try {
                    throw std::runtime_error(std::string("worker context: ") + gpart_last_error());
                } catch (...) {
                    fail();
                }
                std::unique_lock<std::mutex> lock(mu);
                if (++finished == threads) {
                    wall_end = clock::now();
                    all_done.notify_all();
                }
                return;
            }
            gpart::detail::bind_context(ctx);
            (void)gpart_refine_set_graphs(ctx, 0);
            run();
            gpart::detail::bind_context(nullptr);
            gpart_ctx_destroy(ctx);
        };
        if (threads == 1) {
            run();  // the calling thread's own context
        } else {
            std::vector<std::thread> pool;
            for (int t = 0; t < threads; ++t) pool.emplace_back(work);
            for (auto& t : pool) t.join();
        }
        if (error) std::rethrow_exception(error);
        if (wall_seconds) *wall_seconds = std::chrono::duration<double>(wall_end - wall_start).count();
    });
}

int gpart_batch_set_group(int graphs) {
    return guard([&] {
        if (graphs < 1 || graphs > 4096) throw std::invalid_argument("batch group size out of range");
        batch_group = graphs;
    });
}

int gpart_batch_group(void) { return batch_group.load(); }

int gpart_graph_coarsen(gpart_graph* g, int parts, uint32_t seed, double stop_ratio, int max_levels,
                        int cluster_size_factor, int pack_isolated, int verify, gpart_hierarchy** out) {
    return guard([&] {
        gpart::CoarsenOptions o;
        o.parts = parts;
        o.seed = seed;
        o.stop_contraction_ratio = stop_ratio;
        o.max_levels = max_levels;
        o.cluster_size_factor = cluster_size_factor;
        o.pack_isolated = pack_isolated != 0;
        o.strict_verify = verify != 0;
        std::visit(
            [&](auto& x) {
                auto h = gpart::coarsen(host_of(x), o);
                *out = new gpart_hierarchy{g->width, std::move(h)};
            },
            g->g);
    });
}

int gpart_graph_aggregate(gpart_graph* g, int parts, uint32_t seed, int level, int cluster_size_factor,
                          int pack_isolated, int32_t* map, uint64_t* cluster_weights, int64_t* coarse_vertices,
                          uint64_t* capacity, uint64_t* maximum_weight) {
    return guard([&] {
        std::visit(
            [&](auto& x) {
                using T = std::conditional_t<std::is_same_v<std::decay_t<decltype(x)>, GraphOf<gpart::StandardTypes>>,
                                             gpart::StandardTypes, gpart::BigTypes>;
                gpart::SclpWorkspace<T> workspace;
                gpart::SclpStats stats;
                auto result = gpart::aggregate(x.device, workspace, parts, seed, level, stats, false,
                                               cluster_size_factor, pack_isolated != 0);
                const auto mapping = result.map.to_host();
                const auto weights = result.vertex_weights.to_host();
                std::copy(mapping.begin(), mapping.end(), map);
                for (std::size_t i = 0; i < weights.size(); ++i) cluster_weights[i] = static_cast<uint64_t>(weights[i]);
                *coarse_vertices = static_cast<int64_t>(result.coarse_vertices);
                *capacity = result.capacity;
                *maximum_weight = result.maximum_weight;
            },
            g->g);
    });
}

int gpart_hierarchy_read(int width, const char* path, gpart_hierarchy** out) {
    return guard([&] {
        if (width == GPART_WIDTH_STANDARD)
            *out = new gpart_hierarchy{width, gpart::read_hierarchy<gpart::StandardTypes>(path)};
        else
            *out = new gpart_hierarchy{width, gpart::read_hierarchy<gpart::BigTypes>(path)};
    });
}

int gpart_hierarchy_levels(const gpart_hierarchy* h) {
    return std::visit([](const auto& x) { return static_cast<int>(x.levels.size()); }, h->h);
}

int gpart_hierarchy_level_size(const gpart_hierarchy* h, int level, int64_t* n, int64_t* m) {
    return guard([&] {
        std::visit(
            [&](const auto& x) {
                if (level < 0 || level >= static_cast<int>(x.levels.size()))
                    throw std::invalid_argument("hierarchy level out of range");
                *n = x.levels[level].vertices();
                *m = x.levels[level].edges();
            },
            h->h);
    });
}

int gpart_hierarchy_level_copy(const gpart_hierarchy* h, int level, int64_t* offsets, int32_t* neighbors,
                               uint64_t* edge_weights, uint64_t* vertex_weights) {
    return guard([&] {
        std::visit(
            [&](const auto& x) {
                if (level < 0 || level >= static_cast<int>(x.levels.size()))
                    throw std::invalid_argument("hierarchy level out of range");
                copy_level(x.levels[level], offsets, neighbors, edge_weights, vertex_weights);
            },
            h->h);
    });
}

int gpart_hierarchy_map_copy(const gpart_hierarchy* h, int level, int32_t* map) {
    return guard([&] {
        std::visit(
            [&](const auto& x) {
                if (level < 0 || level >= static_cast<int>(x.fine_to_coarse.size()))
                    throw std::invalid_argument("hierarchy map out of range");
                std::copy(x.fine_to_coarse[level].begin(), x.fine_to_coarse[level].end(), map);
            },
            h->h);
    });
}

const char* gpart_hierarchy_stop_reason(const gpart_hierarchy* h) {
    return std::visit([](const auto& x) { return x.stop_reason.c_str(); }, h->h);
}

int gpart_hierarchy_write(const gpart_hierarchy* h, const char* path) {
    return guard([&] { std::visit([&](const auto& x) { gpart::write_hierarchy(x, path); }, h->h); });
}

void gpart_hierarchy_free(gpart_hierarchy* h) { delete h; }

int gpart_hierarchy_uncoarsen(const gpart_hierarchy* h, const int32_t* coarsest_labels, int parts, double ratio,
                              int max_rounds, double growth, int verify, int32_t* labels, uint64_t* final_cut) {
    return guard([&] {
        std::visit(
            [&](const auto& x) {
                gpart::RefineOptions o;
                o.parts = parts;
                o.imbalance_ratio = ratio;
                o.max_rounds = max_rounds;
                o.refinement_growth_ratio = growth;
                o.strict_verify = verify != 0;
                const auto& coarsest = x.levels.back();
                std::vector<std::int32_t> c(coarsest_labels, coarsest_labels + coarsest.vertices());
                auto r = gpart::refine_hierarchy_device(x, c, o);
                std::copy(r.partition.begin(), r.partition.end(), labels);
                if (final_cut) *final_cut = r.final_cut;
            },
            h->h);
    });
}

int gpart_hierarchy_uncoarsen_host(const gpart_hierarchy* h, const int32_t* coarsest_labels, int parts,
                                   double ratio, int max_rounds, double growth, int verify, int32_t* labels) {
    return guard([&] {
        std::visit(
            [&](const auto& x) {
                gpart::RefineOptions o;
                o.parts = parts;
                o.imbalance_ratio = ratio;
                o.max_rounds = max_rounds;
                o.refinement_growth_ratio = growth;
                o.strict_verify = verify != 0;
                const auto& coarsest = x.levels.back();
                const std::vector<std::int32_t> c(coarsest_labels, coarsest_labels + coarsest.vertices());
                const auto out = gpart::uncoarsen(x, c, o);
                std::copy(out.begin(), out.end(), labels);
            },
            h->h);
    });
}

int gpart_graph_initial_partition(gpart_graph* g, int parts, double ratio, uint32_t seed, int spectral_mode,
                                  int32_t* labels, double* coords) {
    return guard([&] {
        if (spectral_mode < 0 || spectral_mode > 2) throw std::invalid_argument("unknown spectral mode");
        std::visit(
            [&](auto& x) {
                using Types = std::conditional_t<std::is_same_v<std::decay_t<decltype(x)>, GraphOf<gpart::StandardTypes>>,
                                                 gpart::StandardTypes, gpart::BigTypes>;
                gpart::DeviceVector<double> c;
                auto l = gpart::initial_partition_gpu<Types>(x.device, parts, ratio, seed,
                                                             static_cast<gpart::SpectralMode>(spectral_mode),
                                                             coords ? &c : nullptr);
                const auto host = l.to_host();
                std::copy(host.begin(), host.end(), labels);
                if (coords) {
                    const auto hc = c.to_host();
                    std::copy(hc.begin(), hc.end(), coords);
                }
            },
            g->g);
    });
}

int gpart_graph_spectral_compare(gpart_graph* g, int parts, double ratio, uint32_t seed,
                                 gpart_spectral_comparison* out) {
    return guard([&] {
        std::visit(
            [&](auto& x) {
                using Types = std::conditional_t<std::is_same_v<std::decay_t<decltype(x)>, GraphOf<gpart::StandardTypes>>,
                                                 gpart::StandardTypes, gpart::BigTypes>;
                const auto c = gpart::compare_spectral_precision<Types>(x.device, parts, ratio, seed);
                if (c.depths.size() > 16) throw std::runtime_error("more than 16 recursion depths");
                *out = gpart_spectral_comparison{};
                out->depths = static_cast<int32_t>(c.depths.size());
                for (std::size_t i = 0; i < c.depths.size(); ++i) {
                    const auto& e = c.depths[i];
                    out->level[i] = gpart_spectral_level_error{e.depth, e.tasks, e.vertices, e.max_abs_error,
                                                               e.rms_error, e.splits_differ, 0, e.labels_differ};
                }
                out->cut_fp64 = c.cut_fp64;
                out->cut_fp32 = c.cut_fp32;
                out->max_load_fp64 = c.max_load_fp64;
                out->max_load_fp32 = c.max_load_fp32;
                out->labels_differ = c.labels_differ;
            },
            g->g);
    });
}

int gpart_graph_refine(gpart_graph* g, int32_t* labels, int parts, double ratio, int max_rounds, int stale_rounds,
                       int verify, gpart_refine_report* report) {
    return guard([&] {
        std::visit(
            [&](auto& x) {
                gpart::RefineOptions o;
                o.parts = parts;
                o.imbalance_ratio = ratio;
                o.max_rounds = max_rounds;
                o.stale_rounds = stale_rounds;
                o.strict_verify = verify != 0;
                std::vector<std::int32_t> l(labels, labels + x.device.vertices());
                gpart::RefineStats s;
                gpart::refine_partition(host_of(x), l, o, &s);
                std::copy(l.begin(), l.end(), labels);
                if (report) {
                    report->rounds = s.rounds;
                    report->admission_iterations = s.admission_iterations;
                    report->admission_fallbacks = s.admission_fallbacks;
                    report->pad = 0;
                    report->proposals = s.proposals;
                    report->accepted = s.accepted;
                    report->initial_cut = s.initial_cut;
                    report->final_cut = s.final_cut;
                    report->final_max_part_weight = s.final_max_part_weight;
                }
            },
            g->g);
    });
}

}  // extern "C"
