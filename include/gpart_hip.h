/*
 * gpart_hip.h -- C ABI of the gfx950 kernel library (libgpart_hip.so).
 *
 * This is the seam between the C++ host scheduling (libgpart.so, built with g++
 * against this header only) and the hand-written HIP kernels. Every entry point
 * takes plain pointers and sizes, enqueues work on the context's stream and
 * returns an int status (0 = success). On failure, gpart_last_error() returns a
 * thread-local message; no C++ exception crosses this boundary. Entry points
 * whose outputs are host values ("host out") synchronise the stream; all others
 * are stream-ordered and asynchronous.
 *
 * Each entry point names the reference code it replaces (paths relative to the
 * reference repository root). The reference has no C ABI: its phases are C++
 * templates calling CUDA kernels directly; these are the kernel families of
 * src/coarsen.cu, src/initial_partition_gpu.cu, src/refine.cu and
 * src/refine_incremental.cuh.
 *
 * Widths: GPART_WIDTH_STANDARD = int32 offsets / uint32 weights,
 *         GPART_WIDTH_BIG      = int64 offsets / uint64 weights
 * (include/graph_types.hpp:7-17). Vertex ids are always int32.
 */
#ifndef GPART_HIP_H
#define GPART_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPART_OK 0
#define GPART_ERR_INVALID_ARGUMENT 1 /* std::invalid_argument in the reference */
#define GPART_ERR_OVERFLOW 2         /* std::overflow_error */
#define GPART_ERR_RUNTIME 3          /* std::runtime_error, including device failures */

#define GPART_WIDTH_STANDARD 0
#define GPART_WIDTH_BIG 1

/* ---------------- runtime (include/check.hpp, include/detail/scratch_pool.hpp) ---------- */

typedef struct gpart_ctx gpart_ctx;

const char* gpart_last_error(void);
/* Record a message for gpart_last_error (used by layers above this ABI). */
void gpart_hip_set_error(const char* message);
int gpart_device_count(int* count);
/* One stream (default priority) and a caching device allocator per context. */
int gpart_ctx_create(int device, gpart_ctx** out);
int gpart_ctx_destroy(gpart_ctx* ctx);
int gpart_ctx_synchronize(gpart_ctx* ctx);
void* gpart_ctx_stream(gpart_ctx* ctx);
/* Stream-ordered caching allocator; replaces ScratchAllocator/ScratchPoolSession
 * (include/detail/scratch_pool.hpp:15-106). Blocks are cut best-fit from device
 * segments the context keeps; a freed block serves later allocations on the same
 * stream. Segments go back to the device only through gpart_release_cache, context
 * destruction, or an allocation the device cannot otherwise serve: ROCm scrubs
 * recycled device memory inside the allocation call (about 30 ms per GiB on
 * MI355X), so the reference's per-run pool trim would charge every run for it. */
int gpart_device_alloc(gpart_ctx* ctx, size_t bytes, void** out);
int gpart_device_free(gpart_ctx* ctx, void* ptr);
/* Synchronise, then return every segment without live blocks to the device
 * (ScratchPoolSession::finish, scratch_pool.hpp:42-49). */
int gpart_release_cache(gpart_ctx* ctx);
/* Allocator state: cached (free) blocks and bytes, live bytes, and for the last
 * release the number of segments freed and host seconds (0, waiting for the stream,
 * freeing segments) in release_seconds[3]. */
int gpart_cache_stats(gpart_ctx* ctx, int64_t* cached_blocks, uint64_t* cached_bytes, uint64_t* live_bytes,
                      int64_t* released_blocks, double* release_seconds);
/* Full allocator counters (no counterpart in the reference). */
typedef struct gpart_cache_info_t {
    uint64_t reserved_bytes;        /* held in segments */
    uint64_t live_bytes;            /* in live blocks */
    uint64_t free_bytes;            /* in cached free blocks */
    uint64_t peak_reserved_bytes;
    uint64_t peak_live_bytes;
    int64_t segments;
    int64_t live_blocks;
    int64_t free_blocks;
    int64_t segment_allocations;    /* device allocations made, all time */
    double segment_allocation_seconds; /* host seconds inside them */
    int64_t segment_releases;       /* segments returned, all time */
    int64_t out_of_memory_releases; /* releases forced by a failed device allocation */
} gpart_cache_info_t;
int gpart_cache_info(gpart_ctx* ctx, gpart_cache_info_t* out);
/* Tuning of the refinement rounds (no counterpart in the reference, whose rounds
 * always update the cache with atomics): a round rebuilds the n x k cache from the
 * CSR instead when the level has >= min_entries entries and its movers hold more
 * than m / share of them; share <= 0 never rebuilds. Defaults 2^23 and 16. Results
 * are identical either way. */
int gpart_refine_set_rebuild(gpart_ctx* ctx, int64_t min_entries, int64_t share);
int gpart_memcpy_htod(gpart_ctx* ctx, void* dst, const void* src, size_t bytes); /* host out: sync */
int gpart_memcpy_dtoh(gpart_ctx* ctx, void* dst, const void* src, size_t bytes); /* host out: sync */
int gpart_memcpy_dtod(gpart_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Fill with a byte value using a kernel (safe under stream capture). */
int gpart_memset(gpart_ctx* ctx, void* dst, int value, size_t bytes);
/* count copies of a 4- or 8-byte pattern (kernel; safe under stream capture). */
int gpart_fill_pattern(gpart_ctx* ctx, void* dst, uint64_t pattern, int pattern_bytes, int64_t count);
/* iota: out[i] = i */
int gpart_sequence_i32(gpart_ctx* ctx, int32_t* out, int64_t n);
/* Widen vertex weights of either width to uint64. */
int gpart_widen_weights(gpart_ctx* ctx, const void* weights, int width, int64_t n, uint64_t* out);

/* Live kernel timing for roofline measurement: when enabled, every launch of the
 * named family's main kernel is bracketed by HIP events on the context stream,
 * and its algorithmic HBM bytes are accumulated. */
#define GPART_PROFILE_SCLP_AFFINITY 0
#define GPART_PROFILE_CONTRACT_SORT 1
#define GPART_PROFILE_REFINE_CACHE 2
#define GPART_PROFILE_REFINE_PROPOSALS 3
#define GPART_PROFILE_REFINE_FILTER 4
#define GPART_PROFILE_SPECTRAL 5
#define GPART_PROFILE_REBALANCE 6      /* balance_proposals_kernel; units: bucket histogram adds (R17) */
#define GPART_PROFILE_CACHE_UPDATE 7   /* movers' cache updates (R13); units: global atomics */
#define GPART_PROFILE_SCLP_COMMIT 8    /* admission commit (K18); units: global atomics */
#define GPART_PROFILE_FAMILIES 9
int gpart_profile_enable(gpart_ctx* ctx, int family, int enable);
/* host out: total milliseconds, launch count, algorithmic bytes */
int gpart_profile_read(gpart_ctx* ctx, int family, double* total_ms, int64_t* launches,
                       double* bytes);
/* host out: the family's atomic operations over the same launches (0 for families
 * that are not atomic-bound). */
int gpart_profile_read_units(gpart_ctx* ctx, int family, double* units);
int gpart_profile_reset(gpart_ctx* ctx);

/* ---------------- primitives (P1-P13 of the reference's CUB/Thrust calls) ---------------- */

/* Stable LSD radix sort of (key, value) pairs over key bits [begin_bit, end_bit), in
 * place (keys/values are overwritten with the sorted result). values may be NULL.
 * value_bytes is 4 or 8. Replaces cub::DeviceRadixSort::SortKeys/SortPairs
 * (src/coarsen.cu:898-923,1238-1246,1746-1755). */
int gpart_sort_pairs_u64(gpart_ctx* ctx, uint64_t* keys, void* values, int value_bytes,
                         int64_t n, int begin_bit, int end_bit);
/* Exclusive prefix sum (thrust::exclusive_scan, src/coarsen.cu:1017,1486,1550). */
int gpart_exclusive_scan_u64(gpart_ctx* ctx, const uint64_t* in, uint64_t* out, int64_t n);
/* Stable compaction of indices with nonzero flags (thrust::copy_if,
 * src/coarsen.cu:997; src/refine_incremental.cuh:277). host out: count. */
int gpart_compact_flags(gpart_ctx* ctx, const int32_t* flags, int64_t n, int32_t* out,
                        int64_t* count);

/* ---------------- graph view ---------------- */

typedef struct gpart_csr_view {
    int64_t n;                    /* vertices */
    int64_t m;                    /* directed CSR entries */
    int32_t width;                /* GPART_WIDTH_* */
    const void* offsets;          /* n + 1 entries of the width's offset type */
    const int32_t* neighbors;     /* m */
    const void* edge_weights;     /* m entries of the width's weight type */
    const void* vertex_weights;   /* n entries of the width's weight type */
} gpart_csr_view;

/* Degree classes of one level (sclp_degree_class_kernel, src/coarsen.cu:204-216,
 * and copy_if :994-1004). Lists are device arrays owned by the caller (n each). */
typedef struct gpart_degree_classes {
    int32_t* low;       /* degree <= low_max, 8 lanes per vertex */
    int32_t* medium;    /* low_max < degree <= medium_max, one wave per vertex */
    int32_t* high;      /* degree > medium_max, one workgroup per vertex */
    int64_t low_count;
    int64_t medium_count;
    int64_t high_count;
    int64_t high_edges;  /* sum of degrees of high vertices */
    int64_t max_degree;
    int64_t isolated_count;  /* vertices of degree 0 */
} gpart_degree_classes;

/* host out: counts. */
int gpart_degree_classes_build(gpart_ctx* ctx, const gpart_csr_view* g,
                               gpart_degree_classes* classes);

/* ---------------- int64 CSR input on the device ---------------- */

/* Device arrays of one CSR, context allocations owned by the caller. */
typedef struct gpart_device_csr {
    void* offsets;          /* n + 1 entries of the width's offset type */
    int32_t* neighbors;     /* m */
    void* edge_weights;     /* m entries of the width's weight type, all 1 */
    void* vertex_weights;   /* n entries of the width's weight type, all 1 */
    int64_t n;
    int64_t m;
} gpart_device_csr;

/* The narrowing of src/graph.cpp:95-116 (int64 indptr / indices to the width's
 * offsets and int32 neighbours, unit weights), done on the device: the int64
 * arrays stream through device staging in 64 MiB chunks and a kernel narrows each.
 * indptr has n + 1 entries, indices m; host memory (pageable). host out: *status, a
 * bit set: 1 offsets decrease or are negative, 2 an offset exceeds the width's
 * offset type, 4 a neighbour lies outside [0, n), 8 indptr[0] != 0 or indptr[n] !=
 * m. On a non-zero status nothing is allocated; dimension limits are the caller's
 * check. */
int gpart_csr_from_int64(gpart_ctx* ctx, int width, int64_t n, const int64_t* indptr, int64_t m,
                         const int64_t* indices, gpart_device_csr* out, int32_t* status);
/* The same from two int64 files of n + 1 and m entries (src/graph.cpp:47-122),
 * read through two pinned staging buffers that alternate, so that reading the next
 * chunk overlaps the transfer and narrowing of the last. */
int gpart_csr_from_int64_files(gpart_ctx* ctx, int width, const char* indptr_path, const char* indices_path,
                               int64_t n, int64_t m, gpart_device_csr* out, int32_t* status);

/* ---------------- partition verification and quality ---------------- */

/* Counters of a labelling for the verifier (tools/verify_partition.py:33-54) and
 * the quality metrics (graph_untils/analyze_partition.cpp:219-254). */
typedef struct gpart_part_counters_t {
    int64_t first_invalid_label;    /* first vertex whose label is outside [0, parts), -1 if none */
    int64_t first_invalid_neighbor; /* first entry whose neighbour is outside [0, n), -1 if none */
    int32_t min_label;              /* over all vertices; 0 and -1 when n == 0 */
    int32_t max_label;
    uint64_t nonloop_entries;       /* entries with neighbour != row */
    uint64_t directed_cut;          /* entries whose endpoints carry different valid labels */
    uint64_t weighted_directed_cut; /* their edge weights */
} gpart_part_counters_t;

/* Two device passes over the labels (device, n) and the CSR. per_part (host, may
 * be NULL) receives 7 x parts words: vertices, internal entries, outgoing cut
 * entries, incoming cut entries, then the edge weights of the last three, each
 * indexed by part. Entries with an out-of-range neighbour or an endpoint whose
 * label is outside [0, parts) are left out of the per-part counters. parts == 0
 * computes the label range only. host out: *out, per_part. */
int gpart_part_counters(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, int parts,
                        gpart_part_counters_t* out, uint64_t* per_part);

/* ---------------- synthetic graphs on the device ---------------- */

/* The host generators' graphs (gpart_generate_rmat/grid2d/mesh3d of gpart.h),
 * entry for entry, built on the device: unit weights, rows sorted ascending, no
 * self loops or duplicates. offsets (n + 1 entries of the width's offset type) and
 * neighbors (m int32) are context allocations for the caller (gpart_device_free).
 * No reference counterpart (graph_untils/ prepares graphs on the host).
 * host out: *n, *m. */
int gpart_generate_rmat_device(gpart_ctx* ctx, int width, int scale, int64_t edge_factor, double a, double b,
                               double c, uint64_t seed, void** offsets, int32_t** neighbors, int64_t* n,
                               int64_t* m);
/* nz == 1: the 2D grid of rows = ny, cols = nx; otherwise the nx x ny x nz mesh. */
int gpart_generate_lattice_device(gpart_ctx* ctx, int width, int64_t nx, int64_t ny, int64_t nz, void** offsets,
                                  int32_t** neighbors, int64_t* n, int64_t* m);

/* ---------------- SCLP coarsening (src/coarsen.cu) ---------------- */

/* record_bound_kernel (:44-70): saturating row weight sums, their maximum and the
 * "rows not strictly ascending" flag. host out: *maximum, *unsorted. */
int gpart_record_bounds(gpart_ctx* ctx, const gpart_csr_view* g, uint64_t* row_bounds,
                        uint64_t* maximum, int* unsorted);
/* contract_record_bounds (:72-85) with its maximum (:1690-1700). coarse_bounds has
 * n_coarse entries and is overwritten. host out: *maximum. */
int gpart_contract_record_bounds(gpart_ctx* ctx, int64_t n, const int32_t* map,
                                 const uint64_t* fine_bounds, int64_t n_coarse,
                                 uint64_t* coarse_bounds, uint64_t* maximum);

typedef struct gpart_sclp_round {
    uint32_t role_salt;        /* src/coarsen.cu:1152-1154 */
    uint32_t tie_salt;         /* :1155-1157 */
    uint32_t mover_threshold;  /* :1311-1315 */
    int32_t filter_roles;
    int32_t singleton;         /* round 0 on sorted duplicate-free rows (:1161) */
    uint64_t value_bound;      /* bound on every row weight (0: unknown); a bound below
                                  2^32 lets the affinity tables keep 32-bit sums */
} gpart_sclp_round;

/* Affinity + decode of one LP round, fused: sclp_singleton_affinity_kernel,
 * sclp_affinity_low_match_kernel, sclp_affinity_medium_hash_kernel, the hub
 * sort/reduce path and sclp_decode_gain_kernel (:306-598, :1142-1303).
 * Writes proposals (INVALID = 0x7fffffff), gains and ties for all n vertices. */
int gpart_sclp_proposals(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* clusters,
                         const gpart_degree_classes* classes, const gpart_sclp_round* round,
                         int32_t* proposals, uint64_t* gains, uint32_t* ties);

/* Admission + commit of one LP round: sclp_admission_keys_kernel, copy_if, the
 * admission radix sort, decode, inclusive_scan_by_key and sclp_commit_kernel
 * (:600-680, :1086-1121, :1329-1373). gain_bound bounds every gain (row weight
 * bound). Updates clusters and cluster_weights.
 * host out: *valid (proposals), *accepted, *nonempty (cluster count after),
 * *predicted_gain (sum of accepted gains; may be NULL). */
int gpart_sclp_admit(gpart_ctx* ctx, int64_t n, const int32_t* proposals,
                     const uint64_t* gains, const uint32_t* ties,
                     const uint64_t* vertex_weights, uint64_t capacity, uint64_t gain_bound,
                     int32_t* clusters, uint64_t* cluster_weights, int64_t* valid,
                     int64_t* accepted, int64_t* nonempty, uint64_t* predicted_gain);

/* Isolated-vertex packing (:1410-1441, kernels :741-773).
 * host out: *isolated (zero-degree count), *nonempty. */
int gpart_sclp_isolated(gpart_ctx* ctx, const gpart_csr_view* g, const uint64_t* vertex_weights,
                        uint64_t capacity, int32_t* clusters, uint64_t* cluster_weights,
                        int64_t* isolated, int64_t* nonempty);

/* Two-hop pairing from round-4 unfiltered proposals (:1443-1539).
 * host out: *candidates, *pairable (pairs found), *merged, *nonempty. */
int gpart_sclp_two_hop(gpart_ctx* ctx, int64_t n, const int32_t* proposals,
                       const uint64_t* gains, const uint32_t* ties,
                       const uint64_t* vertex_weights, uint64_t capacity, uint64_t gain_bound,
                       int64_t merge_budget, int32_t* clusters, uint64_t* cluster_weights,
                       int64_t* candidates, int64_t* pairable, int64_t* merged,
                       int64_t* nonempty);

/* Compaction to coarse ids (:1541-1577): map[v] = rank of clusters[v] among
 * nonempty clusters; coarse vertex weights (width type, n_coarse entries).
 * host out: *maximum_weight. */
int gpart_sclp_compact(gpart_ctx* ctx, int64_t n, const int32_t* clusters,
                       const uint64_t* cluster_weights, const gpart_csr_view* g, int32_t* map,
                       int64_t n_coarse, void* coarse_vertex_weights, uint64_t* maximum_weight);

/* CSR contraction (contract_typed :1674-1814): cross entries keyed by
 * (map[v], map[u]), sorted, duplicates summed, self entries dropped.
 * coarse_row_bound bounds every coarse row weight (UINT64_MAX when unknown) and
 * selects the 8-byte packed record layout when it fits
 * (include/detail/sclp_record.hpp:16-49). Output arrays are allocated from the
 * context allocator and owned by the caller (free with gpart_device_free).
 * host out: *coarse_m. */
int gpart_contract(gpart_ctx* ctx, const gpart_csr_view* fine,
                   const gpart_degree_classes* classes, const int32_t* map,
                   int64_t n_coarse, uint64_t coarse_row_bound, void** coarse_offsets,
                   int32_t** coarse_neighbors, void** coarse_edge_weights, int64_t* coarse_m);
/* Contraction tuning of a context (no counterpart in the reference): levels of at
 * least min_vertices fine vertices store every entry's coarse neighbour in the
 * count pass and read it back in the encode pass instead of gathering map twice
 * (default 2^24, where map outgrows the caches; 0 always, INT64_MAX never).
 * Results are identical either way. */
int gpart_contract_set_stash(gpart_ctx* ctx, int64_t min_vertices);

/* The contraction record layout for n vertices and a row-weight bound: vertex bits,
 * weight bits (0 = unpacked 8-byte key + separate value) and weight mask
 * (SclpRecordLayout::checked(n, bound, 1), include/detail/sclp_record.hpp:20-35).
 * Host only; needs no device. */
void gpart_record_layout(int64_t vertices, uint64_t bound, int* vertex_bits, int* weight_bits,
                         uint64_t* weight_mask);

/* ---------------- initial partition (src/initial_partition_gpu.cu) ---------------- */

typedef struct gpart_bisect_task {
    int32_t begin;       /* segment of order[] */
    int32_t count;
    int32_t first_part;  /* label of every vertex of the task */
    int32_t parts;
    uint64_t weight;
} gpart_bisect_task;

#define GPART_SPECTRAL_FP64 0   /* FP64 vectors, reference summation order */
#define GPART_SPECTRAL_FP32 1   /* FP32 vectors/SpMV, FP64 moments/normalisation/prefix */

/* Plain per-task form (the reference's launch structure :206-229): spectral_start,
 * finish_moments, 65 normalize and 64 spectral_step + finish_moments launches, then
 * the (x, id) sort of order[begin, begin+count) and the inclusive vertex-weight
 * prefix of the sorted segment (prefix: count entries, device). */
int gpart_spectral_task(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels,
                        int32_t* order, const uint64_t* vertex_weights,
                        const gpart_bisect_task* task, uint32_t seed, double* x, double* y,
                        double* degree, uint64_t* prefix);
/* assign_split (:149-153). */
int gpart_spectral_assign(gpart_ctx* ctx, const int32_t* order, const gpart_bisect_task* task,
                          int32_t boundary, int32_t left_parts, int32_t* labels);

/* Batched form: every task of one recursion depth shares each launch (the grid is
 * the tasks' 8-vertex blocks): start, 64 diffusion steps that also finish the
 * previous moments and normalise on read, then one workgroup per task for the
 * (x, id) sort, weight prefix, feasible split and label assignment, all on the
 * device. tasks: host array (at most 32 tasks, each at most
 * gpart_spectral_depth_max_task() vertices); next: device array receiving the
 * child tasks. precision: GPART_SPECTRAL_FP64 (bit-identical to the plain form)
 * or GPART_SPECTRAL_FP32. host out: *next_count, *status (0 ok, 1 infeasible
 * split, 2 leaf over capacity). */
int gpart_spectral_depth(gpart_ctx* ctx, const gpart_csr_view* g, int32_t* labels,
                         int32_t* order, const uint64_t* vertex_weights,
                         const gpart_bisect_task* tasks, int32_t task_count, uint32_t seed,
                         uint64_t capacity, int precision, double* coords,
                         gpart_bisect_task* next, int32_t* next_count, int32_t* status);
/* Largest task (vertices) the batched form accepts. */
int64_t gpart_spectral_depth_max_task(void);

/* Persistent form of the batched bisection, replacing the reference's host task
 * loop (src/initial_partition_gpu.cu:198-255) for every task of at most
 * gpart_spectral_depth_max_task() vertices: one launch serves the root tasks and
 * every task they produce, at every depth, of every graph of the call, from one
 * device work queue. A task of c vertices is served by ceil(ceil(c/8) / B)
 * workgroups (B: own_blocks of gpart_spectral_queue_tuning). Labels, order and coordinates are
 * bit-identical to gpart_spectral_depth run depth by depth on the same roots.
 * roots: host array; root_graph[r] names the graph of roots[r]; each root has at
 * least two parts and its vertices carry the label first_part. Per graph status
 * out: 0 ok, 1 infeasible split, 2 leaf above capacity (the graph's labels are then
 * incomplete). A wait beyond its spin limit or a queue overflow is an error. */
typedef struct gpart_queue_graph {
    gpart_csr_view graph;            /* every graph of one call has the same width */
    int32_t* labels;                 /* n, device */
    int32_t* order;                  /* n, device: the roots' segments of vertex ids */
    const uint64_t* vertex_weights;  /* n, device */
    double* coords;                  /* n, device, or NULL: final coordinates */
    uint64_t capacity;
    uint32_t seed;
    int32_t status;                  /* out */
} gpart_queue_graph;
typedef struct gpart_queue_report {
    int64_t tasks;       /* tasks served */
    int64_t slots;       /* team slots written (roots included) */
    int64_t entries;     /* CSR entries of the served tasks' rows */
    int64_t vertices;    /* vertices of the served tasks */
    int32_t workgroups;  /* grid of the launch */
    int32_t pad;
} gpart_queue_report;
int gpart_spectral_queue(gpart_ctx* ctx, gpart_queue_graph* graphs, int32_t graph_count,
                         const gpart_bisect_task* roots, const int32_t* root_graph, int32_t root_count,
                         int precision, gpart_queue_report* report);
/* Whether the single-graph host driver uses the persistent form for batched
 * bisection (1) or gpart_spectral_depth depth by depth (0, the default: faster for
 * one graph). Groups of graphs (gpart::initial_partition_group) always use it. */
int gpart_spectral_set_queue(int32_t enabled);
int32_t gpart_spectral_queue_enabled(void);
/* own_blocks: blocks of 8 vertices per queue workgroup in the team-size rule
 * (8..128, default 8); max_task: the single-graph driver's largest queue task
 * (1..4096, default 4096), larger batched tasks run depth by depth. */
int gpart_spectral_queue_set_tuning(int32_t own_blocks, int32_t max_task);
void gpart_spectral_queue_tuning(int32_t* own_blocks, int32_t* max_task);

/* ---------------- refinement (src/refine.cu, src/refine_incremental.cuh) ---------------- */

/* refine_cut_kernel + device_cut_value (src/refine.cu:52-87,117-142).
 * host out: *cut (undirected; odd directed total is an error). */
int gpart_cut(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, uint64_t* cut);
/* gpart_cut without any allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid allocation). The same kernels as gpart_cut,
 * with their scratch (about 4 MiB: total, workgroup partials, tile rows) in one
 * static device array per GPU, held until the process ends and resolved when a
 * context is created on that GPU. The CSR may hold at most
 * gpart_cut_noalloc_max_entries() entries (2^32); larger graphs are refused with
 * GPART_ERR_INVALID_ARGUMENT. Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (two contexts or threads at once) race. host out: *cut. */
int gpart_cut_noalloc(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, uint64_t* cut);
int64_t gpart_cut_noalloc_max_entries(void);
/* accumulate_part_weights_kernel (src/refine.cu:102-115); part_weights (device,
 * k entries) is overwritten. */
int gpart_part_weights(gpart_ctx* ctx, const gpart_csr_view* g, const int32_t* labels, int k,
                       uint64_t* part_weights);
/* project_partition_kernel (src/refine.cu:89-100). */
int gpart_project(gpart_ctx* ctx, int64_t n, const int32_t* map, const int32_t* coarse_labels,
                  int32_t* fine_labels);
/* incremental_build_affinity (:7-25): cache[v*k + p], entries of cache_bytes
 * (8, or 4 when every row's edge-weight sum is below 2^32: the entries and every
 * value they pass through during a round then fit 32 bits). */
int gpart_refine_build_cache(gpart_ctx* ctx, const gpart_csr_view* g,
                             const gpart_degree_classes* classes, const int32_t* labels, int k,
                             void* cache, int cache_bytes);
/* Outcome of one refinement round, read back at its end. */
typedef struct gpart_refine_status {
    int64_t movers;              /* vertices moved (0: nothing was applied) */
    int64_t cut_delta;           /* exact cut decrease; negative when the cut grew */
    int64_t candidates;          /* rebalance: vertices of overloaded parts */
    uint64_t max_part_weight;    /* after the round */
    uint64_t part_weights[32];   /* after the round; first k valid */
    int64_t mover_entries;       /* CSR entries of the movers */
    int32_t rebuilt;             /* 1: cache and cut were recomputed from the CSR after
                                    the commit (gpart_refine_set_rebuild); cut_delta is
                                    then 0 and cut_after holds the cut */
    uint64_t cut_after;          /* rebuilt rounds: the cut after the round */
} gpart_refine_status;

/* One LP round of run_incremental_refinement (src/refine_incremental.cuh:338-352):
 * incremental_lp_proposals (:27-46), incremental_priority_filter (:48-67),
 * mover compaction (:276-279), incremental_update_cache_and_cut (:69-95) and
 * incremental_commit (:97-109) with locks set. cache, labels, locks, targets,
 * gains, flags and part_weights are device arrays the round updates in place;
 * one host synchronisation. host out: *status. */
int gpart_refine_lp_round(gpart_ctx* ctx, const gpart_csr_view* g,
                          const gpart_degree_classes* classes, void* cache, int cache_bytes,
                          int32_t* labels, int k, double filter, int32_t* locks, int32_t* targets,
                          int64_t* gains, int32_t* flags, uint64_t* part_weights,
                          gpart_refine_status* status);
/* One iteration of the rebalance loop (src/refine_incremental.cuh:303-331):
 * incremental_rebalance_proposals (:111-157), the source and destination bucket
 * quotas (:167-232) and the same apply without locks. Members of a bucket are
 * ranked by the weight of its lower-id members, one legal outcome of the
 * reference's atomicAdd ranking (:185,216). One host synchronisation.
 * host out: *status (movers == 0: nothing could move). */
int gpart_refine_rebalance_round(gpart_ctx* ctx, const gpart_csr_view* g,
                                 const gpart_degree_classes* classes, void* cache, int cache_bytes,
                                 int32_t* labels, int k, uint64_t* part_weights, uint64_t capacity,
                                 int32_t* targets, int32_t* flags, gpart_refine_status* status);
/* Round graphs (no reference counterpart: the reference launches every kernel of a
 * round from the host). gpart_refine_lp_round and gpart_refine_rebalance_round
 * record the launches of a round the first time they see its level, arrays and
 * settings, as hipGraphs, and replay them on later calls; the results are those
 * of the launches themselves. A profiled family's kernels run outside the graphs.
 * A round is recorded on the second call of its key. Levels above 2^23 vertices are
 * not recorded. gpart_partition_batch launches directly (its graphs do not repeat
 * a key often enough to repay recording). enabled = 0 drops the recordings
 * and launches every kernel directly. The stats count the plans held, the calls
 * that recorded and the calls that replayed since the plans were last dropped. */
int gpart_refine_set_graphs(gpart_ctx* ctx, int enabled);
/* Workgroups of the rebalance proposals launch at most (default from the tuning
 * header); results do not depend on it. */
int gpart_refine_set_balance_blocks(gpart_ctx* ctx, int64_t blocks);
int gpart_refine_graphs_enabled(gpart_ctx* ctx);
int gpart_refine_graph_stats(gpart_ctx* ctx, int64_t* plans, int64_t* captures, int64_t* replays);
/* Outcome of gpart_refine_level. */
typedef struct gpart_refine_level_stats {
    int32_t rounds;               /* LP rounds run */
    int32_t admission_iterations; /* rebalance iterations run */
    int32_t admission_fallbacks;  /* 1: a rebalance could not restore the capacity */
    int32_t error;                /* 0 (errors are returned as status codes) */
    uint64_t proposals;           /* movers of LP rounds */
    uint64_t accepted;            /* movers of applied rounds and iterations */
    uint64_t mover_entries;       /* CSR entries of those movers */
    uint64_t best_cut;            /* cut of the labels left behind */
    uint64_t steps;               /* LP rounds and rebalance iterations with a stage pass */
} gpart_refine_level_stats;

/* The whole of run_incremental_refinement (src/refine_incremental.cuh:235-361) for
 * one level in a single launch of `grid` resident workgroups: LP rounds
 * (:338-352), their rebalance loops (:303-331), the stale-round rule, checkpoints
 * and the restore of the best labels (:353-354). Inputs as for the per-round entry
 * points: the cache built from labels, part_weights (device, k) matching labels,
 * initial_cut the cut of labels; filter 0.25 or 0.75 (:332-333). On return labels
 * and part_weights hold the best partition and stats->best_cut its cut. grid comes
 * from gpart_refine_engine_grid. The labels equal those of the per-round entry
 * points driven by the reference's loop. host out: *stats. */
int gpart_refine_level(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* classes, void* cache,
                       int cache_bytes, int32_t* labels, int k, uint64_t* part_weights, uint64_t capacity,
                       double filter, int rounds, int stale_rounds, uint64_t initial_cut, int32_t* targets,
                       int64_t* gains, int32_t* flags, int32_t* locks, int grid, gpart_refine_level_stats* stats);
/* The grid gpart_refine_level would use for this level, 0 when the level takes the
 * per-round path (more vertices or a longer row than the context's engine tuning
 * allows). host out: *grid. */
int gpart_refine_engine_grid(gpart_ctx* ctx, const gpart_csr_view* g, const gpart_degree_classes* classes,
                             int cache_bytes, int* grid);
/* Engine tuning of a context (no counterpart in the reference): levels of at most
 * max_vertices vertices and rows of at most max_degree entries take the
 * single-launch engine with ceil(n / vertices_per_group) workgroups, at most
 * max_grid (<= 128). max_vertices = 0 switches the engine off. Defaults 0 (off),
 * 512, 128, 4096: measured slower than the per-round path on every level of the
 * 128^3 mesh. Results do not depend on it. */
int gpart_refine_set_engine(gpart_ctx* ctx, int64_t max_vertices, int64_t vertices_per_group, int max_grid,
                            int64_t max_degree);
/* The stale-round test of the reference's loop, (long double)current < 0.999L *
 * best (src/refine_incremental.cuh:347-350), as the engine evaluates it on the
 * device; on the host. best < 2^63. */
int gpart_refine_keep_ratio(uint64_t current, uint64_t best);
/* Largest edge weight (src/refine_incremental.cuh:332). host out. */
int gpart_max_edge_weight(gpart_ctx* ctx, const gpart_csr_view* g, uint64_t* maximum);
/* Edge weight sums. host out: *checked saturates at UINT64_MAX (CheckedSum,
 * src/initial_partition_gpu.cu:31-35); *limited saturates at INT64_MAX/2 + 1
 * (RefinementWeightSum, src/refine.cu:367-383). */
int gpart_edge_weight_sum(gpart_ctx* ctx, const gpart_csr_view* g, uint64_t* checked,
                          uint64_t* limited);
/* Saturating sum, maximum and minimum of vertex weights (src/coarsen.cu:953-958).
 * host out. */
int gpart_vertex_weight_stats(gpart_ctx* ctx, const gpart_csr_view* g, uint64_t* total,
                              uint64_t* maximum, uint64_t* minimum);

#ifdef __cplusplus
}
#endif
#endif /* GPART_HIP_H */
