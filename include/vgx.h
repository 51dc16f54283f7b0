/*
 * vgx.h -- C ABI of the MI355X (gfx950) traversal engine.
 *
 * The engine keeps a VeloGraphX DynamicGraph resident in HBM and runs BFS,
 * incremental BFS repair, PageRank, multi-source BFS and weighted
 * personalized PageRank on it with hand-written HIP kernels.  This header is
 * the drop-in boundary: plain pointers and sizes, no HIP or torch types,
 * status codes instead of exceptions.  Each entry point names the reference
 * interface it replaces (paths relative to the VeloGraphX source tree).
 *
 * Conventions
 *  - Every call returns vgx_status; on failure vgx_last_error() holds a
 *    message for the calling thread.
 *  - Host pointers ("const uint32_t* src") are read during the call only.
 *    Arguments named *_dev are device pointers (HBM), used on (or, for
 *    update batches, waited for on) the stream passed alongside, a
 *    hipStream_t; NULL is the default (null) stream, torch's default.
 *  - Results are answer-ready when a call returns (the call synchronises
 *    its stream), unless the name ends in _async.
 *  - Vertex ids are uint32 below 2^31 (bit 31 marks tombstones on device).
 *  - Algorithm handles borrow their graph (non-owning, like the reference's
 *    Graph& members, incremental/bfs.hpp:356); destroy them before the
 *    graph.  Calls on one graph and its algorithms must not overlap.
 */
#ifndef VGX_H
#define VGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  VGX_OK = 0,
  VGX_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
  VGX_OUT_OF_RANGE = 2,     /* std::out_of_range */
  VGX_DEVICE_ERROR = 3,     /* HIP runtime failure */
  VGX_OUT_OF_MEMORY = 4,    /* std::bad_alloc */
  VGX_UNSUPPORTED = 5
} vgx_status;

#define VGX_UNREACHABLE 0xFFFFFFFFu /* types.hpp:7 kUnreachable */

typedef struct vgx_graph vgx_graph;
typedef struct vgx_bfs vgx_bfs;
typedef struct vgx_pagerank vgx_pagerank;
typedef struct vgx_msbfs vgx_msbfs;
typedef struct vgx_kcore vgx_kcore;
typedef struct vgx_components vgx_components;

/* Message of the last failed call on this thread ("" if none). */
const char* vgx_last_error(void);
/* Version string of the library. */
const char* vgx_version(void);
/* Number of visible gfx950 devices; 0 when none (no error). */
int vgx_device_count(void);

/* ------------------------------------------------------------------ graph */

/* DynamicGraph(vertices, directed) -- storage/dynamic_graph.hpp:481-488. */
vgx_status vgx_graph_create(uint64_t vertices, int directed, vgx_graph** out);
vgx_status vgx_graph_destroy(vgx_graph* g);

/* DynamicGraph::bulk_load_edges -- dynamic_graph.hpp:502-529: replaces the
 * graph with the canonical (sorted, de-duplicated, loop-free) CSR of the m
 * edges, mirrored when undirected; grows to max id + 1.  The edges are
 * uploaded and checked on the device; a list of 2^31 arcs or more (m, twice m
 * when undirected) is refused with VGX_INVALID_ARGUMENT. */
vgx_status vgx_graph_bulk_load(vgx_graph* g, const uint32_t* src, const uint32_t* dst, uint64_t m);
/* No counterpart in the reference: vgx_graph_bulk_load for callers that hold
 * the m edges as one host array of (src, dst) pairs, pairs[2i] and
 * pairs[2i + 1] (an (m, 2) NumPy edge array, a vector of uint32 pairs).  One
 * upload; the pairs are split and checked on the device.  Same result and
 * id rule as vgx_graph_bulk_load (grows to max id + 1, below 2^31). */
vgx_status vgx_graph_bulk_load_pairs(vgx_graph* g, const uint32_t* pairs, uint64_t m);

/* vgx_graph_bulk_load over an edge list already in device memory -- the
 * device form of from_numpy_edges / from_arrow_columns (python/bindings.cpp
 * :21-25, :34), whose host forms copy the list through an UpdateBatch.
 * Edge i is (src_dev[i * stride], dst_dev[i * stride]); ids are id_bytes
 * wide (4: uint32 or int32, a negative int32 being out of range; 8: int64)
 * and must lie in [0, 2^31 - 1) (VGX_INVALID_ARGUMENT otherwise, checked on
 * the device before the graph changes).  `stream` is the stream the list was
 * written on (null: the default stream); the call waits for it.  Replaces
 * the adjacency, grows to max id + 1 over non-loop edges, version += 1. */
vgx_status vgx_graph_bulk_load_device(vgx_graph* g, const void* src_dev, const void* dst_dev, uint64_t m,
                                      uint64_t stride, uint32_t id_bytes, void* stream);

/* from_scipy_csr (python/bindings.cpp:26-32) for a CSR already in device
 * memory: indptr_dev[vertices + 1] and indices_dev[nnz], both index_bytes
 * wide (4: int32, 8: int64).  Same checks and messages as the host form
 * ("invalid CSR offsets", "CSR contains out-of-range vertex id"), made on the
 * device before the graph changes; an undirected graph keeps the entries with
 * v >= u and mirrors them; loops are dropped.  The graph grows to `vertices`;
 * when at least one entry is kept its adjacency is replaced and version
 * += 1 (the host form applies one UpdateBatch of those entries). */
vgx_status vgx_graph_bulk_load_csr_device(vgx_graph* g, const void* indptr_dev, const void* indices_dev,
                                          uint64_t vertices, uint64_t nnz, uint32_t index_bytes, void* stream);

/* DynamicGraph::apply(UpdateBatch) -- dynamic_graph.hpp:543-549, 717-733.
 * add[i] != 0 inserts (src[i], dst[i]), 0 removes it; sequential semantics
 * (last op on an edge wins), self-loops ignored, vertex growth to max id+1,
 * version += 1 when k > 0. */
vgx_status vgx_graph_apply(vgx_graph* g, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                           uint64_t k);
/* vgx_graph_apply for a batch in device memory (see vgx_bfs_apply_device
 * for the pointer and stream rules). */
vgx_status vgx_graph_apply_device(vgx_graph* g, const uint32_t* src_dev, const uint32_t* dst_dev,
                                  const uint8_t* add_dev, uint64_t k, void* stream);

/* HIP device ordinal the graph lives on (no reference counterpart: the
 * device of the arrays vgx_*_device and the DLPack views return). */
vgx_status vgx_graph_device(const vgx_graph* g, int* device);

/* vertex_count / version / directed / edge_count_directed / delta_edge_count
 * -- dynamic_graph.hpp:486-491, 589-597. */
vgx_status vgx_graph_info(const vgx_graph* g, uint64_t* vertex_count, uint64_t* version, int* directed,
                          uint64_t* edge_count_directed);

/* DynamicGraph::neighbors / in_neighbors -- dynamic_graph.hpp:551-557.
 * Writes at most cap ids (ascending) to out and the full count to *count. */
vgx_status vgx_graph_neighbors(vgx_graph* g, uint32_t u, int in_direction, uint32_t* out, uint64_t cap,
                               uint64_t* count);

/* DynamicGraph::has_edge -- dynamic_graph.hpp:583-587. */
vgx_status vgx_graph_has_edge(vgx_graph* g, uint32_t u, uint32_t v, int* present);

/* Live out-degree of every vertex (neighbor_count, graph_access.hpp:142-153)
 * into out[vertex_count]. */
vgx_status vgx_graph_degrees(vgx_graph* g, int in_direction, uint32_t* out);

/* DynamicGraph::compact / consolidate_to_csr_snapshot --
 * dynamic_graph.hpp:631-639, storage/consolidation.hpp:132-163: rewrites
 * every row contiguously without tombstones (same logical graph). */
vgx_status vgx_graph_compact(vgx_graph* g);

/* Storage introspection (storage_bytes, delta counts; dynamic_graph.hpp:597-618). */
typedef struct {
  uint64_t arena_entries;    /* allocated id slots, both directions */
  uint64_t arena_used;       /* slots handed out to rows */
  uint64_t tombstones;       /* deleted ids still stored */
  uint64_t storage_bytes;    /* device bytes held by the graph */
  uint64_t relocated_rows;   /* rows moved since the last compaction */
} vgx_graph_storage;
vgx_status vgx_graph_storage_info(vgx_graph* g, vgx_graph_storage* out);

/* DynamicGraph::ensure_vertex -- dynamic_graph.hpp:493-500: grows the
 * vertex count to v + 1 when it is smaller (no arcs, version unchanged). */
vgx_status vgx_graph_ensure_vertex(vgx_graph* g, uint32_t v);

/* Delta introspection -- base_edge_count_directed, delta_edge_count,
 * delta_ratio, dirty_out/in_segment_count, is_compact
 * (dynamic_graph.hpp:569-618).  Rows are updated in place here, so a
 * row's delta is read off its state against its last bulk load or
 * compaction: entries inserted since plus tombstones (an arc inserted and
 * deleted again since counts twice, where the reference's overlay drops
 * it; a base arc deleted and re-added counts nothing, as there). */
typedef struct {
  uint64_t base_arcs;      /* out-row entries at each row's last bulk load or compaction */
  uint64_t delta_arcs;     /* out-row delta: inserted entries + tombstones */
  uint64_t delta_in_arcs;  /* the same over the in-rows (= delta_arcs when undirected) */
  uint64_t dirty_out_rows; /* out-rows with a nonzero delta */
  uint64_t dirty_in_rows;  /* in-rows with a nonzero delta (= dirty_out_rows when undirected) */
} vgx_graph_delta;
vgx_status vgx_graph_delta_info(vgx_graph* g, vgx_graph_delta* out);

/* DynamicGraph::maybe_compact -- dynamic_graph.hpp:620-631, 764-789: every
 * row (each direction) whose delta is at least threshold * max(8, base
 * length) drops its tombstones in place and becomes its own base; other
 * rows keep their delta.  *compacted (may be null) = 1 when any row was
 * compacted.  The logical graph and version are unchanged. */
vgx_status vgx_graph_maybe_compact(vgx_graph* g, double threshold, int* compacted);

/* consolidate_to_csr_snapshot -- storage/consolidation.hpp:132-163: a new
 * graph (*out, owned by the caller) holding the current logical graph in
 * compacted rows (no tombstones, no relocated rows), version 1, on the
 * source's device; the source graph is not changed. */
vgx_status vgx_graph_snapshot(vgx_graph* src, vgx_graph** out);

/* The neighbour-latency probe of the steady-state storage benchmark --
 * benchmarks/steady_state_storage.cpp:61-86 (probe_once / probe_median):
 * `count` probes of out-rows u_i = (i * 11400714819323198485 mod 2^64) mod
 * vertex_count, one wave per probe on the device, timed over `repeats`
 * launches; *ns_per_probe is the median launch time per probe.  *digest is
 * the sum mod 2^64 of mix(live length) + sum of mix(v) over each probed
 * row (mix = splitmix64's output function), equal for every graph with the
 * same logical rows whatever their storage; a digest that differs between
 * repeats is VGX_DEVICE_ERROR.  An empty graph gives 0 and 0. */
vgx_status vgx_graph_probe_neighbors(vgx_graph* g, uint64_t count, uint32_t repeats, double* ns_per_probe,
                                     uint64_t* digest);

/* Export the logical graph as CSR (ascending rows) into host buffers:
 * off[vertex_count+1], adj[edge_count_directed]. */
vgx_status vgx_graph_export_csr(vgx_graph* g, int in_direction, uint64_t* off, uint32_t* adj);

/* -------------------------------------------------------------------- BFS */

/* BasicIncrementalBFS(graph, source, deletion_fallback_fraction) --
 * incremental/bfs.hpp:18-21; runs recompute() (direction-optimizing). */
vgx_status vgx_bfs_create(vgx_graph* g, uint32_t source, double deletion_fallback_fraction, vgx_bfs** out);
vgx_status vgx_bfs_destroy(vgx_bfs* b);

/* BasicIncrementalBFS::recompute -- bfs.hpp:100-120. */
vgx_status vgx_bfs_recompute(vgx_bfs* b);
/* A new source for the same object, then recompute (no reference
 * counterpart: there the source is fixed at construction, bfs.hpp:18-21,
 * and a new root means a new object).  Reuses the object's device buffers,
 * so traversals from many roots allocate nothing; like construction, a
 * source outside the graph gives all-unreachable distances. */
vgx_status vgx_bfs_set_source(vgx_bfs* b, uint32_t source);

/* Allocation-free BFS.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under stream capture, serving loops
 * that forbid allocation).  The distances of a full direction-optimizing
 * BFS from `source` (UINT32_MAX = unreachable; a source outside the graph
 * leaves every entry unreachable) are written to dist_dev[0, vertex_count),
 * device memory on the graph's device, with the same kernels and results as
 * vgx_bfs_create / vgx_bfs_recompute; *reachable (if not null) gets the
 * reached count.  The traversal's scratch is one array per GPU of about
 * 53 MB, held until the process ends, so the graph may have at most
 * 1,048,576 vertices (more is VGX_INVALID_ARGUMENT, as is cap below the
 * vertex count).  Calls that overlap on one GPU (from several host threads)
 * race on that array and must be serialised by the caller.  Runs on the
 * graph's stream and returns answer-ready. */
vgx_status vgx_bfs_distances_noalloc(vgx_graph* g, uint32_t source, uint32_t* dist_dev, uint64_t cap,
                                     uint64_t* reachable);

/* BasicIncrementalBFS::apply -- bfs.hpp:29-98: applies the batch to the
 * graph and repairs the distances exactly (localized invalidation and
 * re-relaxation, or a full recompute past the fallback fraction). */
vgx_status vgx_bfs_apply(vgx_bfs* b, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                         uint64_t k);
/* The same batch already in device memory (zero-copy ingestion,
 * python/bindings.cpp:21-36's from_numpy path without the host copy):
 * src_dev, dst_dev, add_dev are device pointers of this graph's device,
 * written on `stream` (a hipStream_t, synchronised first unless it is the
 * graph's own; null = the default stream).  Results equal vgx_bfs_apply's
 * on the same ops. */
vgx_status vgx_bfs_apply_device(vgx_bfs* b, const uint32_t* src_dev, const uint32_t* dst_dev, const uint8_t* add_dev,
                                uint64_t k, void* stream);

/* distances() -- bfs.hpp:23: copies *count (= vertex entries) uint32 values
 * into out (cap entries max). */
vgx_status vgx_bfs_distances(vgx_bfs* b, uint32_t* out, uint64_t cap, uint64_t* count);
/* Device pointer to the distance array (valid until the next call). */
vgx_status vgx_bfs_distances_device(vgx_bfs* b, const uint32_t** dist_dev, uint64_t* count);
/* Zero-copy views of that array (no reference counterpart; the DLPack
 * views of the Python module use it): delta = +1 when a view is made, with
 * the stream its reader queues work on (null: the default stream), -1 when
 * it is released.  While views are held, every call that writes the
 * distances first makes the graph's stream wait (on the device) for the
 * work queued so far on those streams, and an array that has to grow keeps
 * its old allocation until the last view is released (then, once those
 * streams have finished, it is freed).  The streams must outlive their
 * views. */
vgx_status vgx_bfs_hold_view(vgx_bfs* b, int delta, void* stream);
/* The graph a BFS object borrows (no reference counterpart; the device of
 * its arrays is vgx_graph_device of it). */
vgx_status vgx_bfs_graph(const vgx_bfs* b, vgx_graph** g);

typedef struct {
  uint64_t reachable_count;        /* bfs.hpp:24 */
  uint64_t last_deletion_candidates;
  uint64_t last_affected_vertices;
  int last_used_full_recompute;
  uint32_t levels;                 /* BFS levels of the last full recompute */
  uint32_t bottom_up_levels;
  uint64_t last_relaxed_vertices;  /* vertices whose distance changed in repair */
  /* With profiling on: level-kernel launches and their summed duration
   * (HIP events on the graph's stream) since the last reset. */
  uint64_t level_launches;
  double level_kernel_ms;
  uint64_t steps;          /* launches the last traversal needed: its levels, tails and packs */
  double traversal_ms;     /* profiling mode 2: GPU time of the traversals since enabled */
  uint64_t traversals;
} vgx_bfs_stats;
vgx_status vgx_bfs_stats_get(vgx_bfs* b, vgx_bfs_stats* out);
/* Profiling, and a reset of the accumulated launch counts and durations:
 * 0: off; 1: HIP events around every step launch (level_kernel_ms, and
 * each level's steps summed in vgx_bfs_level_log); 2: events around every
 * whole traversal (traversal_ms). */
vgx_status vgx_bfs_set_profiling(vgx_bfs* b, int on);
/* Per-level record of the last full traversal (diagnostics; no reference
 * counterpart): direction, frontier vertices and out-edges, bottom-up rows
 * left for the second phase, and the level's kernel time when profiling is
 * on.  At most 64 levels are recorded. */
typedef struct {
  uint32_t bottom_up;  /* 0 top-down (queue), 1 bottom-up, 2 top-down from the frontier bitmap */
  uint64_t frontier_vertices;
  uint64_t frontier_edges;
  uint64_t pending_vertices;
  uint64_t pending_edges;
  double kernel_ms;
} vgx_bfs_level;
vgx_status vgx_bfs_level_log(vgx_bfs* b, vgx_bfs_level* out, uint32_t cap, uint32_t* count);

/* Direction-optimizing switch and launch shape (autotuned).  Every choice
 * is made on the device from the traversal's own counters, so the result
 * and the launch program never depend on an earlier traversal. */
typedef struct {
  double alpha;          /* top-down -> bottom-up when m_f > m_u / alpha */
  double beta;           /* bottom-up -> top-down when n_f < n_u / beta */
  uint32_t blocks;       /* step-kernel workgroups (0 = auto) */
  uint32_t levels_per_sync; /* steps enqueued before the first host check (0 = as many as
                               this graph version's traversals have needed) */
  double tail_ratio;     /* a bottom-up level leaves rows beyond the head to a tail step
                            when n_f * tail_ratio < n_u */
  double claim_ratio;    /* a top-down level claims by distance (pack step after it)
                            when m_f * claim_ratio >= n_u */
} vgx_bfs_tuning;
vgx_status vgx_bfs_set_tuning(vgx_bfs* b, const vgx_bfs_tuning* t);
vgx_status vgx_bfs_get_tuning(vgx_bfs* b, vgx_bfs_tuning* t);
/* Autotuner (no reference counterpart): times answer-ready full BFS runs
 * from this object's source, `reps` per candidate, by coordinate descent
 * over alpha {5, 10, 15, 25, 40}, beta {12, 18, 24, 48}, workgroups
 * {auto, 512, 1024, 2048}, tail ratio {8, 16, 32, 64, 128} and claim ratio
 * {2, 3, 5, 8, 12};
 * installs the fastest setting and returns it with its mean time.  The
 * distances afterwards are those of the last run (unchanged: BFS distances
 * do not depend on the tuning). */
vgx_status vgx_bfs_autotune(vgx_bfs* b, uint32_t reps, vgx_bfs_tuning* best, double* best_ms);

/* --------------------------------------------------------------- PageRank */

/* BasicIncrementalPageRank(graph, damping) -- incremental/pagerank.hpp:32-38.
 * precision: 64 = FP64 ranks and sums (reference semantics), 32 = FP32 ranks
 * with FP32 compensated accumulation. */
vgx_status vgx_pagerank_create(vgx_graph* g, double damping, int precision, vgx_pagerank** out);
vgx_status vgx_pagerank_destroy(vgx_pagerank* p);
/* recompute(max_iterations, tol) -- pagerank.hpp:193-202 (L-inf stop). */
vgx_status vgx_pagerank_recompute(vgx_pagerank* p, uint64_t max_iterations, double tol);
/* Full solve with an L1 stop (the static push pagerank's criterion,
 * src/algorithms.cpp:56-81), optionally warm-started from the current ranks. */
vgx_status vgx_pagerank_solve_l1(vgx_pagerank* p, uint64_t max_iterations, double l1_tol, int warm_start);
/* pagerank(const CsrGraph&, damping, max_iterations, tolerance) --
 * src/algorithms.cpp:56-81, include/velographx/algorithms.hpp:10: FP64 ranks
 * from 1/n, at most max_iterations sweeps, stop once the L1 change is below
 * tolerance; no solver object is kept.  Copies min(cap, n) ranks to out,
 * sets *count = n and, if not null, *iterations to the sweeps run.  Damping
 * outside [0, 1] is VGX_INVALID_ARGUMENT. */
vgx_status vgx_pagerank_static(vgx_graph* g, double damping, uint64_t max_iterations, double tolerance, double* out,
                               uint64_t cap, uint64_t* count, uint64_t* iterations);
/* apply(batch, local_iterations, tol, full_fallback_fraction) -- pagerank.hpp:47-191. */
vgx_status vgx_pagerank_apply(vgx_pagerank* p, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                              uint64_t k, uint64_t local_iterations, double tol, double full_fallback_fraction);
/* Localized residual-push repair (no reference counterpart: the
 * reference's apply, pagerank.hpp:47-191, runs Jacobi sweeps over an active
 * set and falls back whenever a dangling vertex exists).  The object keeps
 * the exact residual r = b + d M x - x of its ranks x.  The batch is applied
 * to the graph; the changed rows' old contributions leave r and their new
 * ones enter it (a dangling row's share goes to a uniform scalar); then
 * rounds of pushes move every entry above `threshold` (0 = l1_tol / 1024)
 * into its rank and d r_u / deg(u) to each out-neighbour, lowering the
 * threshold 16-fold whenever no entry is above it, until sum |r| <= l1_tol;
 * the uniform part is folded into the ranks exactly (A^-1 1 = n/(1-d) x*).
 * Past max_push_fraction x stored arcs pushed in total, it falls back: the
 * ranks from before the pushes are restored and warm full sweeps (at most
 * max_sweeps) run until the L1 change is below l1_tol.  After a batch
 * whose first push round alone was over budget, batches of at least half
 * its size go to the sweeps directly (every 8th of them tries the pushes
 * again; a push phase that finishes clears this).  Either way the L1 norm of the residual of the reported ranks is
 * below l1_tol (the L1 error is at most that over 1 - d).  A change of the
 * vertex count recomputes from scratch. */
typedef struct {
  uint64_t touched_rows;    /* rows whose out-arcs the batch's ops address */
  uint64_t rounds;          /* push rounds */
  uint64_t pushed_vertices; /* pushes (a vertex pushed in two rounds counts twice) */
  uint64_t pushed_arcs;     /* out-arcs those pushes scanned */
  uint64_t fallback_sweeps; /* full sweeps run (fallback, or the residual's first sync) */
  int used_fallback;
  int skipped_push;         /* no push phase: an earlier batch of at least half this size had
                               its first push round over budget (every 8th such batch tries) */
  double residual_l1;       /* sum |r| of the reported ranks */
  double threshold;         /* the last push threshold */
} vgx_pagerank_push_stats;
vgx_status vgx_pagerank_apply_push(vgx_pagerank* p, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                                   uint64_t k, double l1_tol, double threshold, double max_push_fraction,
                                   uint64_t max_sweeps, vgx_pagerank_push_stats* out);
/* values() as float64 -- pagerank.hpp:40. */
vgx_status vgx_pagerank_values(vgx_pagerank* p, double* out, uint64_t cap, uint64_t* count);

/* validate_against_full -- pagerank.hpp:204-232: FP64 reference solve on the
 * device; out_errors = {l1_error, linf_error, reference_residual_l1,
 * reference_residual_linf}. */
vgx_status vgx_pagerank_validate(vgx_pagerank* p, uint64_t reference_max_iterations, double reference_tol,
                                 double* out_errors, uint64_t* reference_iterations, int* reference_converged);
/* apply_validated's fallback -- pagerank.hpp:246-256: install the FP64
 * reference solve as the current values. */
vgx_status vgx_pagerank_install_reference(vgx_pagerank* p, uint64_t reference_max_iterations, double reference_tol);
/* Device pointer to the current ranks (float32 or float64 per precision). */
vgx_status vgx_pagerank_values_device(vgx_pagerank* p, const void** dev, uint64_t* count, int* precision);
/* vgx_bfs_hold_view for views of the ranks. */
vgx_status vgx_pagerank_hold_view(vgx_pagerank* p, int delta, void* stream);
/* The graph a PageRank object borrows (no reference counterpart). */
vgx_status vgx_pagerank_graph(const vgx_pagerank* p, vgx_graph** g);

typedef struct {
  uint64_t last_repaired_vertices;  /* pagerank.hpp:41-45 */
  uint64_t last_repair_iterations;
  double last_residual_l1;
  double last_residual_linf;
  int last_full_recompute_converged;
  uint64_t dangling_vertices;
  int last_used_full_recompute;
} vgx_pagerank_stats;
vgx_status vgx_pagerank_stats_get(vgx_pagerank* p, vgx_pagerank_stats* out);

/* ------------------------------------------------------ multi-source BFS */

/* 1..64 roots traversed together (bit-parallel frontiers, uint8 levels). */
vgx_status vgx_msbfs_create(vgx_graph* g, const uint32_t* roots, uint32_t k, vgx_msbfs** out);
vgx_status vgx_msbfs_destroy(vgx_msbfs* m);
vgx_status vgx_msbfs_run(vgx_msbfs* m);
/* Distances of root r as uint32 (VGX_UNREACHABLE) into out[vertex_count]. */
vgx_status vgx_msbfs_distances(vgx_msbfs* m, uint32_t r, uint32_t* out, uint64_t cap);
/* Root-major uint8 levels (255 = unreachable) of all roots: out[k * vertex_count]. */
vgx_status vgx_msbfs_levels_u8(vgx_msbfs* m, uint8_t* out, uint64_t cap);
vgx_status vgx_msbfs_reached(vgx_msbfs* m, uint64_t* reached_per_root, uint32_t cap);
/* Levels stored by the last run (deepest distance + 1). */
vgx_status vgx_msbfs_level_count(vgx_msbfs* m, uint32_t* levels);
/* Tuning (no reference counterpart): workgroups of the short-row kernels
 * (0 = auto) and the direction switch (push while the frontier's out-arcs
 * are fewer than arcs / alpha); the other parameters keep their values (see
 * vgx_msbfs_set_params).  The autotuner times `reps` runs per candidate of
 * every parameter of vgx_msbfs_tuning (see there) and reports workgroups and
 * alpha; vgx_msbfs_get_params reads all of them. */
vgx_status vgx_msbfs_set_tuning(vgx_msbfs* m, uint32_t blocks, double alpha);
/* Run the pass as a captured hipGraph (no reference counterpart): the
 * launch program of the first 16 levels is captured once and replayed by
 * later runs, also after the graph changed (its arguments live in device
 * memory and are rewritten before each replay); a change of vertex count or
 * size class captures again.  Results are identical to the eager path. */
vgx_status vgx_msbfs_set_graph(vgx_msbfs* m, int on);
/* Levels that ran in the push direction in the last run, and hipGraph
 * replays so far. */
vgx_status vgx_msbfs_stats(vgx_msbfs* m, uint32_t* push_levels, uint64_t* graph_replays);
vgx_status vgx_msbfs_autotune(vgx_msbfs* m, uint32_t reps, uint32_t* best_blocks, double* best_alpha,
                              double* best_ms);
/* All launch parameters of a pass (no reference counterpart): the
 * short-row kernels' workgroups (0 = auto); alpha (push while the frontier's
 * out-arcs are fewer than arcs / alpha); beta (a pulling pass turns back to
 * push only once its new (vertex, root) pairs * beta < vertices * roots; 0 =
 * off); the degree buckets -- rows of up to short_row arcs (1..1024) take one
 * lane, longer rows chunks of `chunk` arcs (64..2^20), one wave each.  None
 * changes a result.  vgx_msbfs_autotune sweeps, one at a time, alpha over
 * {2, 4, 8, 16, 32}, beta over {0, 4, 16, 64}, short_row over {8, 16, 32,
 * 64, 128}, chunk over {256, ..., 4096} and the workgroups, and keeps the
 * fastest of each. */
typedef struct {
  uint32_t blocks;
  double alpha, beta;
  uint32_t short_row, chunk;
} vgx_msbfs_tuning;
vgx_status vgx_msbfs_set_params(vgx_msbfs* m, const vgx_msbfs_tuning* t);
vgx_status vgx_msbfs_get_params(vgx_msbfs* m, vgx_msbfs_tuning* t);

/* IncrementalBFS::apply (incremental/bfs.hpp:29-98) for every root at once:
 * applies the batch to the graph (DynamicGraph::apply semantics) and makes
 * every root's distances those of a BFS of the new graph.  The repair keeps
 * the reference's steps per root -- pairs whose shortest parents were all
 * deleted or affected are affected, reset, seeded from unaffected
 * in-neighbours, the added arcs relaxed and decreases propagated -- run for
 * all roots together (no reference counterpart for several roots); a full
 * pass runs instead when the repair's affected pairs pass the fallback
 * fraction of vertices x roots, when the cost model prefers it (mode 0),
 * when the graph changed since the distances were computed, or when the
 * batch brings a root's id into the graph.  Before the first run() the
 * batch is applied and a full pass runs. */
vgx_status vgx_msbfs_apply(vgx_msbfs* m, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k);
/* vgx_msbfs_apply for a batch in device memory (see vgx_bfs_apply_device). */
vgx_status vgx_msbfs_apply_device(vgx_msbfs* m, const uint32_t* src_dev, const uint32_t* dst_dev,
                                  const uint8_t* add_dev, uint64_t k, void* stream);
/* Repair policy (no reference counterpart): mode 0 chooses per batch from
 * running means of the answer-ready time of each arm -- the full pass, and
 * the repair per batch-size class floor(log2 ops), a class without its own
 * mean borrowing the nearest smaller class's scaled by the size ratio; a
 * repair that allocated its arrays or captured its programs is not counted,
 * and a class predicted within twice the pass time that chose the pass 16
 * times tries the repair again
 * --, 1 always repairs, 2 always runs a full pass.  A repair whose affected
 * (vertex, root) pairs exceed fallback_fraction * vertices * roots stops and
 * runs a full pass (default 0.1; the reference's deletion_fallback_fraction,
 * 0.35 at bfs.hpp:18, plays this part for one root). */
vgx_status vgx_msbfs_set_repair(vgx_msbfs* m, int mode, double fallback_fraction);
typedef struct {
  uint64_t batches, repaired, full_passes, fallbacks;
  int last_repaired, last_fallback;
  uint64_t last_candidates, last_affected_pairs, last_lowered;
  double pass_ms, repair_us_per_op;
} vgx_msbfs_repair_info;
vgx_status vgx_msbfs_repair_stats(vgx_msbfs* m, vgx_msbfs_repair_info* out);

/* -------------------------------------- repair-vs-recompute selection */

/* Adaptive selector of the reference's publication policy
 * (benchmarks/publication_policy_bfs.cpp:162-233 decision, :297-330 model
 * update, constants benchmarks/adaptive_policy_bfs.cpp:21-31).  One record
 * per batch: the features, the prediction, the decision and its outcome. */
typedef struct {
  double update_fraction;                  /* ops / directed arcs before the batch */
  double shallow_parent_deletion_fraction; /* first batch only (:126-148) */
  double predicted_incremental_us;
  double predicted_full_us;
  double execution_us;                     /* answer-ready cost of the arm that ran */
  double decision_us;
  uint64_t affected_vertices;              /* incremental arm: invalidated vertices */
  int first_batch;
  int chose_full;
  int internal_fallback;                   /* incremental arm fell back to a full BFS */
  int reason;                              /* see vgx_policy_reason_name */
} vgx_selector_trace;

typedef struct vgx_policy vgx_policy;
typedef struct vgx_selector vgx_selector;

/* Name of a decision reason ("large_preflight_full", ...). */
const char* vgx_policy_reason_name(int reason);

/* The decision model alone (host; replays recorded features):
 * create with the graph's vertex count, the measured cost of the initial
 * full BFS and its reachable fraction; decide() fills predicted_*,
 * chose_full and reason from update_fraction, shallow_parent_deletion_fraction
 * and first_batch; observe() takes the same record with execution_us,
 * internal_fallback and affected_vertices filled in. */
vgx_status vgx_policy_create(uint64_t vertices, double initial_full_us, double initial_reachable_fraction,
                             vgx_policy** out);
vgx_status vgx_policy_destroy(vgx_policy* p);
vgx_status vgx_policy_decide(vgx_policy* p, vgx_selector_trace* t);
vgx_status vgx_policy_observe(vgx_policy* p, const vgx_selector_trace* t);

/* The selector bound to an IncrementalBFS, as the reference harness runs it
 * (publication_policy_bfs.cpp:243-256): create() times one full BFS as the
 * initial full-arm sample; apply() computes the features on the GPU (update
 * fraction from the device arc count, shallow-parent deletions from the
 * device distances), decides, runs either graph apply + full BFS or the
 * localized repair, times it answer-ready and updates the model.  The BFS
 * object's own fallback fraction still applies inside the repair arm (the
 * reference harness constructs it with 2.0 to disable it). */
vgx_status vgx_selector_create(vgx_bfs* b, vgx_selector** out);
vgx_status vgx_selector_destroy(vgx_selector* s);
vgx_status vgx_selector_apply(vgx_selector* s, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                              uint64_t k, vgx_selector_trace* trace);

/* ------------------------------------------ weighted personalized PageRank */

/* Forward: X = (1-d) S + d P^T X with P_uv = w_uv / sum_v' w_uv' over the
 * out-CSR snapshot (csr_off/csr_adj device arrays of the graph, obtained
 * with vgx_graph_csr_device).  S, X are n x K row-major float32 device
 * arrays.  Fixed iteration count `iterations`. */
vgx_status vgx_graph_csr_device(vgx_graph* g, int in_direction, const uint32_t** off_dev,
                                const uint32_t** adj_dev, uint64_t* vertices, uint64_t* arcs);
vgx_status vgx_ppr_forward(vgx_graph* g, const float* w_dev, const float* s_dev, float* x_dev, uint32_t K,
                           double damping, uint32_t iterations, void* stream);
/* Adjoint: L = C + d P L; grad_s = (1-d) L; grad_w (raw weights, through the
 * row normalisation) per arc of the out-CSR snapshot. */
vgx_status vgx_ppr_backward(vgx_graph* g, const float* w_dev, const float* x_dev, const float* c_dev,
                            float* lambda_dev, float* grad_s_dev, float* grad_w_dev, uint32_t K,
                            double damping, uint32_t iterations, void* stream);

/* ----------------------------------------------------------------- k-core */

/* BasicIncrementalKCore(graph) -- incremental/kcore.hpp:20-26, 58-63: core
 * numbers of an undirected graph (a directed one is VGX_INVALID_ARGUMENT,
 * as the reference throws).  Computed on the device by level-synchronous
 * peeling; the object borrows the graph. */
vgx_status vgx_kcore_create(vgx_graph* g, vgx_kcore** out);
vgx_status vgx_kcore_destroy(vgx_kcore* kc);
/* recompute() -- kcore.hpp:58-63. */
vgx_status vgx_kcore_recompute(vgx_kcore* kc);
/* apply(batch) -- kcore.hpp:31-56: applies the batch to the graph and
 * re-peels the reference's repair region, the components of the batch's
 * endpoints before and after the update (the other vertices keep their core
 * numbers).  k = 0 changes nothing and sets last_repaired_vertices to 0, as
 * the reference's empty batch does. */
vgx_status vgx_kcore_apply(vgx_kcore* kc, const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint64_t k);
/* core() -- kcore.hpp:28: copies min(cap, n) core numbers, *count = n. */
vgx_status vgx_kcore_core(vgx_kcore* kc, uint32_t* out, uint64_t cap, uint64_t* count);
/* last_repaired_vertices() -- kcore.hpp:29: the vertices re-peeled by the
 * last apply (its region) or recompute (all); and the level loop's levels
 * (levels whose scan claimed vertices) and peeling rounds. */
vgx_status vgx_kcore_stats(vgx_kcore* kc, uint64_t* last_repaired_vertices, uint32_t* levels, uint32_t* rounds);

/* ------------------------------------------------------------ components */

/* BasicIncrementalComponents(graph) -- incremental/connected_components.hpp:
 * 19-25: connected components of an undirected graph (a directed one is
 * VGX_INVALID_ARGUMENT, as the reference throws).  A component's label is
 * its smallest vertex id (the reference's is a union-find root that depends
 * on its union order: the partition is the same, the representative may
 * differ).  Min-label hooking with pointer jumping on the device. */
vgx_status vgx_components_create(vgx_graph* g, vgx_components** out);
vgx_status vgx_components_destroy(vgx_components* cc);
vgx_status vgx_components_recompute(vgx_components* cc);
/* apply(batch) -- connected_components.hpp:35-97: applies the batch to the
 * graph and relabels (k = 0 changes nothing and sets last_repaired_vertices
 * to 0, as the reference's empty batch does). */
vgx_status vgx_components_apply(vgx_components* cc, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                                uint64_t k);
/* component(v) -- :27-30: VGX_OUT_OF_RANGE for v outside the graph. */
vgx_status vgx_components_component(vgx_components* cc, uint32_t v, uint32_t* label);
/* Every vertex's label: min(cap, n) values, *count = n. */
vgx_status vgx_components_labels(vgx_components* cc, uint32_t* out, uint64_t cap, uint64_t* count);
/* last_repaired_vertices() -- :31-33 (here: the vertex count after a
 * non-empty batch or recompute, 0 after an empty batch) and the hooking
 * rounds of the last computation. */
vgx_status vgx_components_stats(vgx_components* cc, uint64_t* last_repaired_vertices, uint32_t* rounds);

/* ------------------------------------------------------- workload helpers */

/* Graph500-style R-MAT tuple stream of scripts/prepare_publication_datasets.py
 * :80-129 (A,B,C,D = .57,.19,.19,.05, no permutation), driven by the MT19937
 * stream of Python's random.Random(seed).random().  Writes m = edgefactor <<
 * scale tuples into src/dst (host). */
vgx_status vgx_rmat_generate(uint32_t scale, uint32_t edgefactor, uint64_t seed, uint32_t* src, uint32_t* dst);

#ifdef __cplusplus
}
#endif
#endif /* VGX_H */
