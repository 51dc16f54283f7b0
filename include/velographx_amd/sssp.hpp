// Unweighted incremental SSSP -- the interface of
// include/velographx/incremental/sssp.hpp (BasicIncrementalSSSP<DynamicGraph>).
//
// Every edge has weight 1, so the reference's distances are hop distances
// (uint64, kDijkstraInf = 2^64 / 4 - 1 where unreached).  They are kept here
// by the device BFS engine: apply() runs the BFS repair on the batch and
// recompute() a full BFS.  One corner is kept from the reference: a source
// outside the graph gives all-infinite distances, and only a recompute -- a
// batch with a deletion triggers one (sssp.hpp:22-28) -- picks the source up
// once the graph has grown to include it; insert-only batches keep the
// distances infinite (they only relax from already reached endpoints).
#pragma once

#include <cstdint>
#include <limits>
#include <vector>

#include "velographx_amd/velographx.hpp"

namespace velographx {

class IncrementalSSSP {
 public:
  // incremental/dijkstra.hpp:14
  static constexpr std::uint64_t infinity = std::numeric_limits<std::uint64_t>::max() / 4;

  IncrementalSSSP(DynamicGraph& g, VertexId source) : g_(g), source_(source), bfs_(g, source, 0.35) {
    stale_ = source_ >= g_.vertex_count();
  }

  // sssp.hpp:19: hop distances, `infinity` where unreached.
  [[nodiscard]] std::vector<std::uint64_t> distances() const {
    const auto n = g_.vertex_count();
    std::vector<std::uint64_t> out(n, infinity);
    if (stale_) return out;
    const auto d = bfs_.distances();
    for (std::size_t i = 0; i < d.size() && i < n; ++i)
      if (d[i] != IncrementalBFS::unreachable) out[i] = d[i];
    return out;
  }

  // sssp.hpp:21-28: any removal op in the batch means a full recompute.
  void apply(const UpdateBatch& batch) {
    bool deletion = false;
    for (const auto& e : batch.updates) deletion |= !e.add;
    bfs_.apply(batch);
    if (deletion) recompute_after_update();
  }

  void apply_arrays(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k) {
    bool deletion = false;
    for (std::size_t i = 0; i < k; ++i) deletion |= add[i] == 0;
    vgx_check(vgx_bfs_apply(bfs_.handle(), src, dst, add, k));
    if (deletion) recompute_after_update();
  }

  // sssp.hpp:30-36
  void recompute() {
    bfs_.recompute();
    stale_ = source_ >= g_.vertex_count();
  }

 private:
  void recompute_after_update() {
    // the repair already left exact distances for a source inside the graph
    // since the last recompute; a source that has just come into range needs
    // a traversal of its own
    if (stale_ && source_ < g_.vertex_count()) bfs_.recompute();
    stale_ = source_ >= g_.vertex_count();
  }

  DynamicGraph& g_;
  VertexId source_;
  IncrementalBFS bfs_;
  bool stale_ = false;
};

}  // namespace velographx
