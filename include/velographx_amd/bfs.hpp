// BasicIncrementalBFS<Graph> -- the interface of
// include/velographx/incremental/bfs.hpp:13-383 (IncrementalBFS is
// BasicIncrementalBFS<DynamicGraph>, bfs.hpp:383) over the gfx950 C ABI.
//
// The traversal and the batch repair run on the device graph that
// GraphBinding<Graph> supplies: the caller's DynamicGraph, a CsrGraph's
// rows, or a device copy of any other ReadableGraph kept in step through
// graph_access.hpp.  As in the reference, apply() updates the caller's
// graph; the algorithm objects borrow their graph (Graph& members).
#pragma once

#include <concepts>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "velographx_amd/dynamic_graph.hpp"
#include "velographx_amd/graph_binding.hpp"

namespace velographx {

template <class Graph>
class BasicIncrementalBFS {
 public:
  static constexpr std::uint32_t unreachable = kUnreachable;

  // bfs.hpp:18-21: the distances of a full traversal from `source`.
  BasicIncrementalBFS(Graph& g, VertexId source, double deletion_fallback_fraction = 0.35) : bind_(g) {
    vgx_check(vgx_bfs_create(bind_.device().handle(), source, deletion_fallback_fraction, &b_));
  }
  ~BasicIncrementalBFS() {
    if (b_) vgx_bfs_destroy(b_);
  }
  BasicIncrementalBFS(const BasicIncrementalBFS&) = delete;
  BasicIncrementalBFS& operator=(const BasicIncrementalBFS&) = delete;

  // bfs.hpp:23 (a copy: the distances live in HBM).  One entry per vertex
  // of the caller's graph.
  [[nodiscard]] std::vector<std::uint32_t> distances() const {
    std::uint64_t count = 0;
    vgx_check(vgx_bfs_distances(b_, nullptr, 0, &count));
    std::vector<std::uint32_t> d(count);
    if (count) vgx_check(vgx_bfs_distances(b_, d.data(), count, &count));
    if constexpr (!detail::GraphBinding<Graph>::native) d.resize(bind_.vertex_count(), unreachable);
    return d;
  }
  // bfs.hpp:24-27
  [[nodiscard]] std::size_t reachable_count() const { return stats().reachable_count; }
  [[nodiscard]] std::size_t last_deletion_candidates() const { return stats().last_deletion_candidates; }
  [[nodiscard]] std::size_t last_affected_vertices() const { return stats().last_affected_vertices; }
  [[nodiscard]] bool last_used_full_recompute() const { return stats().last_used_full_recompute != 0; }

  // bfs.hpp:29-98: updates the graph and repairs the distances.
  void apply(const UpdateBatch& batch) {
    const UpdateBatch& ops = bind_.forward(batch);
    auto c = ops.columns();
    vgx_check(vgx_bfs_apply(b_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
  }
  // bfs.hpp:100-120
  void recompute() {
    bind_.refresh();
    vgx_check(vgx_bfs_recompute(b_));
  }

  // addition: a batch as columns (no UpdateBatch), on a DynamicGraph.
  void apply_arrays(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k)
    requires std::same_as<Graph, DynamicGraph>
  {
    vgx_check(vgx_bfs_apply(b_, src, dst, add, k));
  }

  [[nodiscard]] vgx_bfs_stats stats() const {
    vgx_bfs_stats s{};
    vgx_check(vgx_bfs_stats_get(b_, &s));
    return s;
  }
  [[nodiscard]] vgx_bfs* handle() const noexcept { return b_; }

 private:
  detail::GraphBinding<Graph> bind_;
  vgx_bfs* b_ = nullptr;
};

// bfs.hpp:383
using IncrementalBFS = BasicIncrementalBFS<DynamicGraph>;

// Addition without a reference counterpart: up to 64 roots traversed in one
// pass (the reference runs independent roots on threads,
// benchmarks/thread_scaling_bfs.cpp:78-108).
class MultiSourceBFS {
 public:
  MultiSourceBFS(DynamicGraph& g, const std::vector<VertexId>& roots)
      : g_(&g), n_(g.vertex_count()), k_(roots.size()) {
    vgx_check(vgx_msbfs_create(g.handle(), roots.data(), static_cast<std::uint32_t>(roots.size()), &m_));
  }
  ~MultiSourceBFS() {
    if (m_) vgx_msbfs_destroy(m_);
  }
  MultiSourceBFS(const MultiSourceBFS&) = delete;
  MultiSourceBFS& operator=(const MultiSourceBFS&) = delete;

  void run(std::size_t vertex_count) {
    n_ = vertex_count;
    vgx_check(vgx_msbfs_run(m_));
  }
  // IncrementalBFS::apply (bfs.hpp:29-98) for every root: the batch goes
  // into the graph and every root's distances are repaired on the device
  // (or recomputed by a full pass, see vgx_msbfs_apply).
  void apply(const UpdateBatch& batch) {
    const auto c = batch.columns();
    apply_arrays(c.src.data(), c.dst.data(), c.add.data(), c.src.size());
  }
  void apply_arrays(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k) {
    vgx_check(vgx_msbfs_apply(m_, src, dst, add, k));
    n_ = g_->vertex_count();
  }
  [[nodiscard]] std::vector<std::uint32_t> distances(std::uint32_t r) const {
    std::vector<std::uint32_t> d(n_);
    vgx_check(vgx_msbfs_distances(m_, r, d.data(), d.size()));
    return d;
  }
  [[nodiscard]] std::vector<std::uint8_t> levels_u8() const {
    std::vector<std::uint8_t> d(n_ * k_);
    vgx_check(vgx_msbfs_levels_u8(m_, d.data(), d.size()));
    return d;
  }
  [[nodiscard]] std::vector<std::uint64_t> reached() const {
    std::vector<std::uint64_t> r(k_);
    vgx_check(vgx_msbfs_reached(m_, r.data(), static_cast<std::uint32_t>(k_)));
    return r;
  }
  [[nodiscard]] std::uint32_t level_count() const {
    std::uint32_t l = 0;
    vgx_check(vgx_msbfs_level_count(m_, &l));
    return l;
  }
  [[nodiscard]] std::size_t root_count() const noexcept { return k_; }
  [[nodiscard]] vgx_msbfs* handle() const noexcept { return m_; }

 private:
  DynamicGraph* g_ = nullptr;
  vgx_msbfs* m_ = nullptr;
  std::size_t n_ = 0, k_ = 0;
};

}  // namespace velographx
