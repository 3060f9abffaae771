/*
 * gpart_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a serial/OpenMP C++ restatement
 * of the reference partitioner (SCLP coarsening, weighted recursive spectral
 * bisection, incremental bucketed refinement). It is the parity checker and
 * the CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository root).
 *
 * Conventions: offsets are int64, neighbour ids int32, weights uint64 (the
 * reference's Big width; Standard-width inputs are the same values). Return
 * value 0 = success; nonzero = error, message via gpo_last_error(). Error
 * codes mirror the reference's exception classes:
 *   1 = std::invalid_argument, 2 = std::overflow_error, 3 = std::runtime_error.
 */
#ifndef GPART_ORACLE_H
#define GPART_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPO_INVALID_ARGUMENT 1
#define GPO_OVERFLOW 2
#define GPO_RUNTIME 3

const char* gpo_last_error(void);
void gpo_set_threads(int threads);
int gpo_get_threads(void);

/* include/partition_capacity.hpp:12-23 */
int gpo_partition_capacity(uint64_t total, int parts, double ratio, uint64_t* out);

/* Opaque graph (CSR with weights). */
typedef struct gpo_graph gpo_graph;
gpo_graph* gpo_graph_create(int64_t n, int64_t m, const int64_t* offsets,
                            const int32_t* neighbors, const uint64_t* edge_weights,
                            const uint64_t* vertex_weights);
void gpo_graph_free(gpo_graph* g);
int64_t gpo_graph_n(const gpo_graph* g);
int64_t gpo_graph_m(const gpo_graph* g);
/* Copy out arrays (caller-sized buffers; any pointer may be NULL). */
void gpo_graph_copy(const gpo_graph* g, int64_t* offsets, int32_t* neighbors,
                    uint64_t* edge_weights, uint64_t* vertex_weights);

/* One SCLP level: src/coarsen.cu:937-1659 (aggregate_typed).
 * map (n entries) and coarse vertex weights (up to n entries) are written;
 * *coarse receives the coarse vertex count, *capacity the cluster cap.
 * sorted_unique_rows selects nothing observable (the singleton fast path is
 * result-identical); it is accepted to mirror the reference signature. */
int gpo_aggregate(const gpo_graph* g, int parts, uint32_t seed, int level,
                  int cluster_size_factor, int pack_isolated,
                  int32_t* map, uint64_t* coarse_weights, int64_t* coarse,
                  uint64_t* capacity, int* rounds);

/* The same level, traced: the clusters (n entries per step, at most max_steps
 * steps) after every LP round (kind 0), isolated packing (1) and two-hop pairing
 * (2), with each step's round number and each LP round's mover threshold. */
int gpo_aggregate_trace(const gpo_graph* g, int parts, uint32_t seed, int level,
                        int cluster_size_factor, int pack_isolated, int max_steps,
                        int32_t* clusters, int32_t* kinds, int32_t* rounds,
                        uint32_t* thresholds, int* steps);

/* Per-vertex SCLP proposal of one LP round (the affinity + decode semantics of
 * src/coarsen.cu:306-598). clusters: n entries. Outputs n entries each. */
int gpo_sclp_proposals(const gpo_graph* g, const int32_t* clusters, uint32_t seed,
                       int level, int round, int filter_roles,
                       uint32_t mover_threshold, int32_t* proposals,
                       uint64_t* gains, uint32_t* ties);

/* Contraction closed form: src/coarsen.cu:1674-1814. */
gpo_graph* gpo_contract(const gpo_graph* fine, const int32_t* map, int64_t coarse,
                        const uint64_t* coarse_weights);

/* Hierarchy: src/coarsen.cu:1906-2164 (coarsen_impl). */
typedef struct gpo_hierarchy gpo_hierarchy;
gpo_hierarchy* gpo_coarsen(const gpo_graph* g, int parts, uint32_t seed,
                           double stop_ratio, int max_levels,
                           int cluster_size_factor, int pack_isolated);
int gpo_hierarchy_levels(const gpo_hierarchy* h);
const gpo_graph* gpo_hierarchy_level(const gpo_hierarchy* h, int level);
/* map from level l to level l+1 (n_l entries) */
const int32_t* gpo_hierarchy_map(const gpo_hierarchy* h, int level);
const char* gpo_hierarchy_stop_reason(const gpo_hierarchy* h);
/* include/hierarchy.hpp:42-126 int32 file format. */
int gpo_write_hierarchy(const gpo_hierarchy* h, const char* path);
void gpo_hierarchy_free(gpo_hierarchy* h);

/* Initial partition: src/initial_partition_gpu.cu:170-276. labels: n. */
int gpo_initial_partition(const gpo_graph* g, int parts, double ratio,
                          uint32_t seed, int32_t* labels);
/* Same, also returning the final FP64 diffusion coordinate of every vertex
 * (the x vector after the last normalisation of the task that split it) and
 * the number of split tasks. coords may be NULL. */
int gpo_initial_partition_trace(const gpo_graph* g, int parts, double ratio,
                                uint32_t seed, int32_t* labels, double* coords,
                                int* tasks);

/* Refinement of one graph from given labels: src/refine.cu:410-421 with
 * run_incremental_refinement src/refine_incremental.cuh:235-361. */
typedef struct gpo_refine_stats {
    int rounds;
    uint64_t proposals;
    uint64_t accepted;
    uint64_t initial_cut;
    uint64_t final_cut;
    uint64_t final_max_part_weight;
    int admission_iterations;
    int admission_fallbacks;
} gpo_refine_stats;
int gpo_refine(const gpo_graph* g, int32_t* labels, int parts, double ratio,
               int max_rounds, int stale_rounds, gpo_refine_stats* stats);

/* Uncoarsen a hierarchy from coarsest labels: src/refine.cu:423-646. */
int gpo_uncoarsen(gpo_hierarchy* h, const int32_t* coarsest_labels, int parts,
                  double ratio, int max_rounds, double growth, int32_t* labels,
                  uint64_t* final_cut);

/* Full pipeline with the CLI's production parameters:
 * app/partition_main.cu:42-93. Returns final cut. */
typedef struct gpo_partition_result {
    uint64_t final_cut;
    uint64_t max_part_weight;
    uint64_t capacity;
    int levels;
    int64_t coarsest_vertices;
    uint64_t initial_cut;
    double coarsen_seconds;
    double initial_seconds;
    double uncoarsen_seconds;
    double total_seconds;
} gpo_partition_result;
int gpo_partition(const gpo_graph* g, int parts, double ratio, uint32_t seed,
                  int32_t* labels, gpo_partition_result* result);

/* Helpers. */
int gpo_cut(const gpo_graph* g, const int32_t* labels, uint64_t* cut);
int gpo_part_weights(const gpo_graph* g, const int32_t* labels, int parts,
                     uint64_t* weights);
/* src/graph.cpp:124-190 */
int gpo_validate_csr(const gpo_graph* g, int allow_self_loops, int strict);

#ifdef __cplusplus
}
#endif
#endif
