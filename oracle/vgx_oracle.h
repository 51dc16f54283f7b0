/*
 * CPU oracle for the traversal hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference engine's algorithms, used solely as
 * the checker by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py.  Nothing in the shipped library links, loads or calls this code.
 *
 * Graphs are passed as canonical CSR (sorted, de-duplicated, loop-free rows,
 * uint64 offsets), the layout of SegmentedCsr::build
 * (include/velographx/storage/dynamic_graph.hpp:79-111) flattened over
 * segments.  Each function cites the reference lines it restates.
 *
 * Parity pins (each function is checked by tests/test_oracle.py):
 *   orc_bfs                 reference known answers (tests/test_static.cpp:17-18,
 *                           ci.yml smoke), scipy.sparse.csgraph, BFS invariants,
 *                           oracle/_ref (compiled reference) on random graphs.
 *   orc_incbfs_apply        the 8 hand cases of tests/test_incremental_bfs_deletion.cpp
 *                           (distances, affected counts, fallback flags) and
 *                           oracle/_ref diagnostics on randomized streams.
 *   orc_pagerank_full       dense linear solve of (I - d P^T) x = b with the
 *                           dangling term; sum-to-one; oracle/_ref.
 *   orc_pagerank_apply      tests/test_pagerank_localized.cpp cases and oracle/_ref.
 *   orc_pagerank_static     algorithms.cpp:56-81 (push form), dense solve.
 *   orc_csr_build           the ModelGraph of oracle.py (bulk load) on random edge
 *                           lists with loops, duplicates and both directions.
 *   orc_csr_apply           ModelGraph.apply on random batch streams (loops, repeated
 *                           arcs with mixed ops, vertex growth by removals).
 *   orc_ppr_sweeps          oracle/ppr.py's Jacobi sweeps (forward and adjoint).
 * The last two, with orc_bfs and orc_pagerank_full, carry the at-size checks
 * (graphs of 2^20-2^22 vertices), where the numpy model is too slow.
 */
#ifndef VGX_ORACLE_H
#define VGX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_UNREACHABLE 0xFFFFFFFFu

/* Queue BFS: src/algorithms.cpp:11-27, incremental/bfs.hpp:100-120.
 * Returns the number of reached vertices; source >= n gives all unreachable. */
uint64_t orc_bfs(uint32_t n, const uint64_t* off, const uint32_t* adj,
                 uint32_t source, uint32_t* dist);

typedef struct {
  uint64_t reachable_count;
  uint64_t deletion_candidates;
  uint64_t affected_vertices;
  int used_full_recompute;
} orc_incbfs_diag;

/* BasicIncrementalBFS::apply (incremental/bfs.hpp:29-98, 206-354).
 * pre_* describe the graph before the batch, post_* after it (the caller
 * applies the batch with DynamicGraph semantics).  dist_pre has n_pre entries
 * and reachable_pre is the object's cached reachable count; dist_post
 * receives n_post entries.  ops: k updates (src, dst, add). */
int orc_incbfs_apply(uint32_t n_pre, const uint64_t* pre_out_off, const uint32_t* pre_out_adj,
                     const uint64_t* pre_in_off, const uint32_t* pre_in_adj,
                     uint32_t n_post, const uint64_t* post_out_off, const uint32_t* post_out_adj,
                     const uint64_t* post_in_off, const uint32_t* post_in_adj,
                     int directed, uint32_t source, double fallback_fraction,
                     uint64_t k, const uint32_t* src, const uint32_t* dst, const uint8_t* add,
                     const uint32_t* dist_pre, uint64_t reachable_pre,
                     uint32_t* dist_post, orc_incbfs_diag* diag);

typedef struct {
  uint64_t iterations;
  double residual_l1;
  double residual_linf;
  int converged;
} orc_pr_info;

/* BasicIncrementalPageRank::full_solve (incremental/pagerank.hpp:277-321):
 * FP64 pull iteration from the uniform vector, dangling mass spread
 * uniformly, stop when the L-inf residual <= tol. */
int orc_pagerank_full(uint32_t n, const uint64_t* in_off, const uint32_t* in_adj,
                      const uint32_t* outdeg, double damping, uint64_t max_iterations,
                      double tol, double* values, orc_pr_info* info);

typedef struct {
  uint64_t last_repaired_vertices;
  uint64_t last_repair_iterations;
  double last_residual_l1;
  double last_residual_linf;
  int last_full_recompute_converged;
  int used_full_recompute;
  uint64_t dangling_vertices;
} orc_pr_apply_diag;

/* BasicIncrementalPageRank::apply (incremental/pagerank.hpp:47-191) on the
 * post-batch graph.  rank holds n_prev values on entry and n_post on exit.
 * dangling_before is the object's dangling count before the batch. */
int orc_pagerank_apply(uint32_t n_prev, uint64_t dangling_before,
                       uint32_t n_post, const uint64_t* out_off, const uint32_t* out_adj,
                       const uint64_t* in_off, const uint32_t* in_adj, int directed,
                       double damping, uint64_t k, const uint32_t* src, const uint32_t* dst,
                       uint64_t local_iterations, double tol, double full_fallback_fraction,
                       double* rank, orc_pr_apply_diag* diag);

/* Static push PageRank (src/algorithms.cpp:56-81): L1 stop. */
int orc_pagerank_static(uint32_t n, const uint64_t* out_off, const uint32_t* out_adj,
                        double damping, uint64_t max_iterations, double tolerance,
                        double* rank);

/* Canonical CSR of an edge list, the state after DynamicGraph::bulk_load_edges
 * (dynamic_graph.hpp:502-529) with SegmentedCsr::build's row order (:79-111):
 * self-loops dropped, both arcs of an edge when undirected, rows sorted and
 * de-duplicated.  n is the vertex count (the caller makes it at least
 * max id + 1 over the non-loop edges); off gets n + 1 entries, adj room for
 * m (directed) or 2m (undirected) arcs.  Returns the arc count. */
uint64_t orc_csr_build(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, int directed,
                       uint64_t* off, uint32_t* adj);

/* DynamicGraph::apply of one batch to a canonical CSR (dynamic_graph.hpp:543-549,
 * 717-733): ops applied in order (so the last op on an arc wins), self-loops
 * ignored, undirected ops on both arcs.  n_new is the post-batch vertex count
 * (at least max id + 1 over the non-loop ops, removals included); off_new gets
 * n_new + 1 entries, adj_new room for the old arcs plus 2k.  Returns the arc
 * count. */
uint64_t orc_csr_apply(uint32_t n_old, const uint64_t* off, const uint32_t* adj, int directed, uint64_t k,
                       const uint32_t* src, const uint32_t* dst, const uint8_t* add, uint32_t n_new,
                       uint64_t* off_new, uint32_t* adj_new);

/* Weighted personalized PageRank sweeps (no reference counterpart; the
 * mathematics of oracle/ppr.py): Y_0 = B, then `iterations` times
 * Y'[v] = B[v] + d * sum_{e in row v} p[e] * Y[idx[e]], K <= 64 columns,
 * row-major n x K.  Over the in-rows with B = (1 - d) S this is the forward
 * X = (1 - d) S + d P^T X; over the out-rows with B = C the adjoint
 * L = C + d P L.  Pinned against oracle/ppr.py's numpy sweeps. */
int orc_ppr_sweeps(uint32_t n, const uint64_t* off, const uint32_t* idx, const double* p, uint32_t K,
                   const double* B, double d, uint64_t iterations, double* Y);

#ifdef __cplusplus
}
#endif
#endif
