// Unweighted incremental SSSP -- the interface of
// include/velographx/incremental/sssp.hpp:14-69 (BasicIncrementalSSSP<Graph>;
// IncrementalSSSP is the DynamicGraph instance, :69).
//
// Every edge has weight 1, so the reference's distances are hop distances
// (uint64, kDijkstraInf = 2^64 / 4 - 1 where unreached).  They are kept here
// by the device BFS engine on the device graph GraphBinding<Graph>
// supplies: apply() runs the BFS repair on the batch and recompute() a full
// BFS.  One corner is kept from the reference: a source outside the graph
// gives all-infinite distances, and only a recompute -- a batch with a
// deletion triggers one (sssp.hpp:21-28) -- picks the source up once the
// graph has grown to include it; insert-only batches keep the distances
// infinite (they only relax from already reached endpoints).
#pragma once

#include <concepts>
#include <cstdint>
#include <limits>
#include <vector>

#include "velographx_amd/dynamic_graph.hpp"
#include "velographx_amd/graph_binding.hpp"

namespace velographx {

namespace incremental_detail {
// incremental/dijkstra.hpp:14
inline constexpr std::uint64_t kDijkstraInf = std::numeric_limits<std::uint64_t>::max() / 4;
}  // namespace incremental_detail

template <class Graph>
class BasicIncrementalSSSP {
 public:
  static constexpr std::uint64_t infinity = incremental_detail::kDijkstraInf;

  // sssp.hpp:17
  BasicIncrementalSSSP(Graph& g, VertexId source) : bind_(g), source_(source) {
    vgx_check(vgx_bfs_create(bind_.device().handle(), source, 0.35, &b_));
    stale_ = source_ >= bind_.vertex_count();
  }
  ~BasicIncrementalSSSP() {
    if (b_) vgx_bfs_destroy(b_);
  }
  BasicIncrementalSSSP(const BasicIncrementalSSSP&) = delete;
  BasicIncrementalSSSP& operator=(const BasicIncrementalSSSP&) = delete;

  // sssp.hpp:19: hop distances, `infinity` where unreached.
  [[nodiscard]] std::vector<std::uint64_t> distances() const {
    const auto n = bind_.vertex_count();
    std::vector<std::uint64_t> out(n, infinity);
    if (stale_) return out;
    std::uint64_t count = 0;
    vgx_check(vgx_bfs_distances(b_, nullptr, 0, &count));
    std::vector<std::uint32_t> d(count);
    if (count) vgx_check(vgx_bfs_distances(b_, d.data(), count, &count));
    for (std::size_t i = 0; i < d.size() && i < n; ++i)
      if (d[i] != kUnreachable) out[i] = d[i];
    return out;
  }

  // sssp.hpp:21-28: any removal op in the batch means a full recompute.
  void apply(const UpdateBatch& batch) {
    bool deletion = false;
    for (const auto& e : batch.updates) deletion |= !e.add;
    auto c = bind_.forward(batch).columns();
    vgx_check(vgx_bfs_apply(b_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
    if (deletion) recompute_after_update();
  }

  void apply_arrays(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k)
    requires std::same_as<Graph, DynamicGraph>
  {
    bool deletion = false;
    for (std::size_t i = 0; i < k; ++i) deletion |= add[i] == 0;
    vgx_check(vgx_bfs_apply(b_, src, dst, add, k));
    if (deletion) recompute_after_update();
  }

  // sssp.hpp:29-36
  void recompute() {
    bind_.refresh();
    vgx_check(vgx_bfs_recompute(b_));
    stale_ = source_ >= bind_.vertex_count();
  }

  [[nodiscard]] vgx_bfs* handle() const noexcept { return b_; }

 private:
  void recompute_after_update() {
    // the repair already left exact distances for a source inside the graph
    // since the last recompute; a source that has just come into range needs
    // a traversal of its own
    if (stale_ && source_ < bind_.vertex_count()) vgx_check(vgx_bfs_recompute(b_));
    stale_ = source_ >= bind_.vertex_count();
  }

  detail::GraphBinding<Graph> bind_;
  VertexId source_;
  vgx_bfs* b_ = nullptr;
  bool stale_ = false;
};

// sssp.hpp:69
using IncrementalSSSP = BasicIncrementalSSSP<DynamicGraph>;

}  // namespace velographx
