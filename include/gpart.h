/*
 * gpart.h -- high-level C ABI of the partitioner (libgpart.so).
 *
 * The drop-in surface for foreign-language callers (ctypes, cgo, JNI): the same
 * operations the reference exposes as C++ templates and CLIs, on plain host
 * arrays. Status codes as in gpart_hip.h (0 ok, 1 invalid_argument,
 * 2 overflow_error, 3 runtime_error); message via gpart_last_error().
 *
 * Reference counterparts:
 *   gpart_graph_upload            load_weighted_graph + validate + make_device_weighted
 *                                 (src/graph.cpp:47-223; app/partition_main.cu:42-66)
 *   gpart_graph_partition         app/partition_main.cu:67-119 (timed region, production options)
 *   gpart_graph_coarsen           coarsen() (include/coarsen.hpp:93-101; src/coarsen.cu:1906-2174)
 *   gpart_graph_aggregate         aggregate() (include/coarsen.hpp:80-85; src/coarsen.cu:937-1659)
 *   gpart_graph_initial_partition initial_partition_gpu (include/initial_partition.hpp:9-12)
 *   gpart_graph_refine            refine_partition (include/refine.hpp:62-66)
 *   gpart_hierarchy_uncoarsen     refine_hierarchy_device, host hierarchy + host labels
 *                                 (include/refine.hpp:69-73; src/refine.cu:648-653)
 *   gpart_hierarchy_uncoarsen_host uncoarsen, host hierarchy + host labels -> host labels
 *                                 (include/uncoarsen.hpp:10-14; src/uncoarsen.cu:100-106)
 *   gpart_hierarchy_write         write_hierarchy (include/hierarchy.hpp:42-126)
 *   gpart_partition_capacity_c    partition_capacity (include/partition_capacity.hpp:12-23)
 */
#ifndef GPART_H
#define GPART_H

#include <stdint.h>

#include "gpart_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpart_graph gpart_graph;          /* device-resident input graph */
typedef struct gpart_hierarchy gpart_hierarchy;  /* host copy of a coarsening hierarchy */
typedef struct gpart_host_csr gpart_host_csr;    /* generated int64 CSR on the host */

int gpart_partition_capacity_c(uint64_t total, int parts, double ratio, uint64_t* out);

/* The calling thread's kernel-library context (for profiling hooks). */
gpart_ctx* gpart_thread_context(void);
/* Synchronise and return cached device memory (ScratchPoolSession::finish). */
int gpart_trim(void);
/* Progress/report lines on stdout (the reference's drivers always print them):
 * 1 on (default), 0 off. Process-wide. */
void gpart_set_log(int enabled);

/* Upload an int64 CSR. edge_weights / vertex_weights may be NULL (unit weights, as
 * the CLI loader sets them). validate: 0 shape checks, 1 strict (symmetry). */
int gpart_graph_upload(int width, int64_t n, const int64_t* indptr, int64_t m, const int64_t* indices,
                       const uint64_t* edge_weights, const uint64_t* vertex_weights, int validate,
                       gpart_graph** out);
/* Load the reference's binary int64 files (native byte order). */
int gpart_graph_load(int width, const char* indptr_path, const char* indices_path, int validate,
                     gpart_graph** out);
/* Synthetic graphs built on the GPU, identical to the host generators below
 * (gpart_generate_rmat, gpart_generate_grid2d with rows = ny and cols = nx, and
 * gpart_generate_mesh3d; nz == 1 selects the grid), with unit weights. The host
 * copy is made on first use by an entry point that needs it. No reference
 * counterpart (graph_untils/ prepares graphs on the host). */
int gpart_graph_generate_rmat(int width, int scale, int64_t edge_factor, double a, double b, double c, uint64_t seed,
                              gpart_graph** out);
int gpart_graph_generate_lattice(int width, int64_t nx, int64_t ny, int64_t nz, gpart_graph** out);
/* The graph's CSR on the host (offsets n + 1, neighbors m, weights m and n); any
 * pointer may be NULL. */
int gpart_graph_arrays(gpart_graph* g, int64_t* offsets, int32_t* neighbors, uint64_t* edge_weights,
                       uint64_t* vertex_weights);
void gpart_graph_free(gpart_graph* g);
int64_t gpart_graph_vertices(const gpart_graph* g);
int64_t gpart_graph_edges(const gpart_graph* g);
int gpart_graph_cut(gpart_graph* g, const int32_t* labels, uint64_t* cut);

/* tools/verify_partition.py:22-65 on the device: labels (host, n entries) must lie
 * in [0, parts), neighbours in [0, n), the directed cut and non-loop entry counts
 * must be even; capacity = floor(imbalance * n / parts) with imbalance a decimal
 * string, exactly as Decimal computes it. Errors carry the reference's messages
 * ("invalid parts or imbalance ratio", "invalid label: <value>", "invalid neighbor
 * id", "odd undirected edge/cut count; input must be symmetric"). loads (host, parts
 * entries) may be NULL. */
/* floor(imbalance * n / parts) for a decimal string, exactly (the verifier's
 * capacity, tools/verify_partition.py:57). Host only. Returns GPART_OK, or an
 * invalid-argument status when imbalance is not a finite decimal >= 1. */
int gpart_decimal_capacity(const char* imbalance, int64_t n, int parts, int64_t* capacity);

typedef struct gpart_verify_report {
    int64_t vertices;
    int64_t undirected_edges;
    int64_t cut;
    double cut_ratio;
    int64_t max_load;
    double max_load_ratio;
    int64_t capacity;
    int32_t capacity_ok;
    int32_t parts;
} gpart_verify_report;
int gpart_graph_verify(gpart_graph* g, const int32_t* labels, int parts, const char* imbalance,
                       gpart_verify_report* out, int64_t* loads);

/* graph_untils/analyze_partition.cpp:176-294 on the device. parts == 0 infers max
 * label + 1. The weighted metrics use the graph's edge weights (unit weights give
 * the tool's default output). per_part (host, may be NULL) receives 7 x out->parts
 * words, as gpart_part_counters writes them, when per_part_capacity >= out->parts
 * (otherwise it is an error). */
typedef struct gpart_quality_report {
    int32_t parts;
    int32_t pad;
    int64_t vertices;
    int64_t directed_edges;
    uint64_t edge_cut;
    uint64_t weighted_edge_cut;
    double edge_cut_ratio;
    double avg_local_cut;
    uint64_t max_local_cut;
    double weighted_avg_local_cut;
    uint64_t weighted_max_local_cut;
    double vertex_avg;
    int64_t min_vertices;
    int64_t max_vertices;
    double vertex_imbalance;
    double vertex_max_deviation;
    double edge_set_avg;
    int64_t min_edge_set;
    int64_t max_edge_set;
    double edge_set_imbalance;
    double edge_set_max_deviation;
} gpart_quality_report;
int gpart_graph_analyze(gpart_graph* g, const int32_t* labels, int parts, gpart_quality_report* out,
                        uint64_t* per_part, int per_part_capacity);

typedef struct gpart_partition_report {
    uint64_t final_cut;
    uint64_t max_part_weight;
    uint64_t capacity;
    uint64_t initial_cut;
    int64_t coarsest_vertices;
    int32_t levels;
    int32_t refined_levels;
    double coarsen_seconds;
    double initial_seconds;
    double uncoarsen_seconds;
    double free_seconds;
    double algorithm_seconds; /* coarsen start -> hierarchy freed (the reference's scope) */
} gpart_partition_report;

/* Full pipeline with the CLI's production options (cluster_size_factor =
 * max(64, ceil(512/k)), pack_isolated, 32 refinement rounds, growth 2.0). The
 * device input is copied first (outside the timed region) so a graph can be
 * partitioned repeatedly. spectral_mode: 0 plain FP64, 1 batched FP64, 2 batched
 * FP32. labels (host, n entries) may be NULL. */
int gpart_graph_partition(gpart_graph* g, int parts, double ratio, uint32_t seed, int verify, int spectral_mode,
                          int32_t* labels, gpart_partition_report* report);

/* Many independent partitions (the batched graph mix; no reference counterpart:
 * the reference runs one graph per process). Graphs are taken largest first from
 * one shared queue. workers must be 1 for now (the queue runs in the calling
 * thread); several worker threads, each with its own context (HIP stream and
 * allocator), are implemented but switched off until a host crash seen with four
 * of them on the full mix is explained. Each graph gets gpart_graph_partition's pipeline and options with
 * parts[i], ratios[i], seeds[i] and spectral_mode; a worker keeps its allocator
 * cache from one graph to the next. With a batched spectral_mode, consecutive
 * graphs of one width are taken in groups (gpart_batch_set_group, default 16):
 * the group is coarsened, the initial partitions of all its graphs run in one
 * persistent launch (gpart_spectral_queue), then each graph is uncoarsened; each
 * graph's labels are those it gets alone. labels[i] (host, n_i entries) and
 * reports may be NULL. *wall_seconds runs from the first coarsening start to the
 * moment the last graph's labels are on the device (label copies come after). */
int gpart_partition_batch(int count, gpart_graph* const* graphs, const int* parts, const double* ratios,
                          const uint32_t* seeds, int spectral_mode, int workers, int32_t* const* labels,
                          gpart_partition_report* reports, double* wall_seconds);
/* Graphs per group of gpart_partition_batch (1..4096; 1: one graph at a time). */
int gpart_batch_set_group(int graphs);
int gpart_batch_group(void);

/* Coarsening hierarchy (host copy). */
int gpart_graph_coarsen(gpart_graph* g, int parts, uint32_t seed, double stop_ratio, int max_levels,
                        int cluster_size_factor, int pack_isolated, int verify, gpart_hierarchy** out);
/* One SCLP aggregation of a graph (aggregate, include/coarsen.hpp:80-85 of the
 * reference; src/coarsen.cu:937-1659). map (n entries) and cluster_weights (at
 * least coarse entries; n suffices) are host arrays. */
int gpart_graph_aggregate(gpart_graph* g, int parts, uint32_t seed, int level, int cluster_size_factor,
                          int pack_isolated, int32_t* map, uint64_t* cluster_weights, int64_t* coarse_vertices,
                          uint64_t* capacity, uint64_t* maximum_weight);
int gpart_hierarchy_read(int width, const char* path, gpart_hierarchy** out);
int gpart_hierarchy_levels(const gpart_hierarchy* h);
int gpart_hierarchy_level_size(const gpart_hierarchy* h, int level, int64_t* n, int64_t* m);
int gpart_hierarchy_level_copy(const gpart_hierarchy* h, int level, int64_t* offsets, int32_t* neighbors,
                               uint64_t* edge_weights, uint64_t* vertex_weights);
int gpart_hierarchy_map_copy(const gpart_hierarchy* h, int level, int32_t* map);
const char* gpart_hierarchy_stop_reason(const gpart_hierarchy* h);
int gpart_hierarchy_write(const gpart_hierarchy* h, const char* path);
void gpart_hierarchy_free(gpart_hierarchy* h);
/* Projection + refinement of an (imported) hierarchy from labels of its coarsest
 * level: levels are uploaded one at a time as they are refined. labels receives the
 * level-0 partition (n_0 entries); final_cut may be NULL. */
int gpart_hierarchy_uncoarsen(const gpart_hierarchy* h, const int32_t* coarsest_labels, int parts, double ratio,
                              int max_rounds, double growth, int verify, int32_t* labels, uint64_t* final_cut);
/* The same through the host uncoarsen overload, which also writes the per-level
 * and final_* report lines while logging is on. */
int gpart_hierarchy_uncoarsen_host(const gpart_hierarchy* h, const int32_t* coarsest_labels, int parts,
                                   double ratio, int max_rounds, double growth, int verify, int32_t* labels);

/* Initial partition of a graph; coords (n doubles) may be NULL. */
int gpart_graph_initial_partition(gpart_graph* g, int parts, double ratio, uint32_t seed, int spectral_mode,
                                  int32_t* labels, double* coords);

/* FP32 against FP64 diffusion of the batched bisection on one graph (normally a
 * coarsest level). No reference counterpart: the reference diffuses in FP64 only.
 * Every recursion depth runs both precisions on the same tasks from the FP64
 * state of the depth before. Per depth: the largest and RMS difference of the
 * normalised vectors, the tasks whose split boundary differs and the vertices that
 * land on a different side. Then the cut and the largest part weight of complete
 * FP64 and FP32 recursions, and the vertices whose final label differs. */
typedef struct gpart_spectral_level_error {
    int32_t depth;
    int32_t tasks;
    int64_t vertices;
    double max_abs_error;
    double rms_error;
    int32_t splits_differ;
    int32_t pad;
    int64_t labels_differ;
} gpart_spectral_level_error;

typedef struct gpart_spectral_comparison {
    int32_t depths; /* entries of level[] used (at most 16) */
    int32_t pad;
    gpart_spectral_level_error level[16];
    uint64_t cut_fp64;
    uint64_t cut_fp32;
    uint64_t max_load_fp64;
    uint64_t max_load_fp32;
    int64_t labels_differ;
} gpart_spectral_comparison;

int gpart_graph_spectral_compare(gpart_graph* g, int parts, double ratio, uint32_t seed,
                                 gpart_spectral_comparison* out);

typedef struct gpart_refine_report {
    int32_t rounds;
    int32_t admission_iterations;
    int32_t admission_fallbacks;
    int32_t pad;
    uint64_t proposals;
    uint64_t accepted;
    uint64_t initial_cut;
    uint64_t final_cut;
    uint64_t final_max_part_weight;
} gpart_refine_report;

/* Refine labels (in/out, host) of a graph. */
int gpart_graph_refine(gpart_graph* g, int32_t* labels, int parts, double ratio, int max_rounds, int stale_rounds,
                       int verify, gpart_refine_report* report);

/* Graph generators (host, int64 CSR, rows sorted, no self loops, no duplicates). */
gpart_host_csr* gpart_generate_grid2d(int64_t rows, int64_t cols);
gpart_host_csr* gpart_generate_mesh3d(int64_t nx, int64_t ny, int64_t nz);
/* R-MAT: 2^scale vertices, edge_factor * 2^scale samples, quadrant probabilities
 * (a, b, c, 1-a-b-c), counter-based RNG seeded by seed; symmetrized, deduplicated,
 * self loops dropped. */
gpart_host_csr* gpart_generate_rmat(int scale, int64_t edge_factor, double a, double b, double c, uint64_t seed);
int64_t gpart_host_csr_vertices(const gpart_host_csr* h);
int64_t gpart_host_csr_entries(const gpart_host_csr* h);
const int64_t* gpart_host_csr_indptr(const gpart_host_csr* h);
const int64_t* gpart_host_csr_indices(const gpart_host_csr* h);
int gpart_host_csr_write(const gpart_host_csr* h, const char* indptr_path, const char* indices_path);
void gpart_host_csr_free(gpart_host_csr* h);

#ifdef __cplusplus
}
#endif
#endif /* GPART_H */
