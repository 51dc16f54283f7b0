// Static CSR graph on the device -- the interface of
// include/velographx/csr_graph.hpp (src/csr_graph.cpp).
//
// A CsrGraph is a device graph built once by the bulk load (self-loops
// dropped, duplicates removed, both directions kept, vertex count = largest
// id + 1) and never updated; the algorithms run on its rows with the
// kernels of the dynamic graph (algorithms.hpp, and the BasicIncremental*
// templates bound to a CsrGraph).  Row accessors return copies (the rows
// live in HBM).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <utility>
#include <vector>

#include "velographx_amd/dynamic_graph.hpp"

namespace velographx {

class CsrGraph {
 public:
  using Edge = std::pair<VertexId, VertexId>;

  CsrGraph() = default;
  explicit CsrGraph(const std::vector<Edge>& edges, bool directed = false) : g_(0, directed) {
    g_.bulk_load_edges(edges);
  }
  CsrGraph(const std::uint32_t* src, const std::uint32_t* dst, std::size_t m, bool directed = false)
      : g_(0, directed) {
    g_.bulk_load_edges(src, dst, m);
  }
  // From one array of m (src, dst) pairs (pairs[2i], pairs[2i + 1]).
  struct Pairs {};
  CsrGraph(Pairs, const std::uint32_t* pairs, std::size_t m, bool directed = false) : g_(0, directed) {
    g_.bulk_load_pairs(pairs, m);
  }

  [[nodiscard]] std::size_t vertex_count() const { return g_.vertex_count(); }
  [[nodiscard]] std::size_t edge_entry_count() const { return g_.edge_count_directed(); }
  [[nodiscard]] bool directed() const { return g_.directed(); }

  // csr_graph.cpp:44-56: ids at or past vertex_count() throw.
  [[nodiscard]] std::vector<VertexId> neighbors(VertexId v) const {
    check(v);
    return g_.neighbors(v);
  }
  [[nodiscard]] std::vector<VertexId> in_neighbors(VertexId v) const {
    check(v);
    return g_.in_neighbors(v);
  }
  [[nodiscard]] std::size_t degree(VertexId v) const { return neighbors(v).size(); }
  // csr_graph.cpp:60-64: false for ids outside the graph.
  [[nodiscard]] bool has_edge(VertexId u, VertexId v) const {
    if (u >= vertex_count() || v >= vertex_count()) return false;
    return g_.has_edge(u, v);
  }

  // The device graph the algorithms run on (its caches are the only state
  // they touch).
  [[nodiscard]] DynamicGraph& device_graph() const noexcept { return g_; }

  // Hop distances from `source` (< vertex_count()) into out[0, vertex_count()),
  // for algorithms.hpp's bfs_distances.  The graph keeps one traversal
  // object, re-targeted per call, so a call costs a traversal and the
  // copy-out rather than a new workspace and a first, uncaptured run; a
  // concurrent call that finds it busy uses an object of its own.
  void bfs_distances_into(VertexId source, std::uint32_t* out) const {
    const std::uint64_t n = vertex_count();
    std::uint64_t count = 0;
    std::unique_lock<std::mutex> lock(bfs_->mu, std::try_to_lock);
    if (!lock.owns_lock()) {
      BfsHandle own;
      vgx_check(vgx_bfs_create(g_.handle(), source, 0.35, &own.b));
      vgx_check(vgx_bfs_distances(own.b, out, n, &count));
      return;
    }
    if (!bfs_->h.b)
      vgx_check(vgx_bfs_create(g_.handle(), source, 0.35, &bfs_->h.b));
    else
      vgx_check(vgx_bfs_set_source(bfs_->h.b, source));
    vgx_check(vgx_bfs_distances(bfs_->h.b, out, n, &count));
  }

 private:
  void check(VertexId v) const {
    if (v >= vertex_count()) throw std::out_of_range("vertex id outside graph");
  }
  struct BfsHandle {
    vgx_bfs* b = nullptr;
    BfsHandle() = default;
    BfsHandle(const BfsHandle&) = delete;
    BfsHandle& operator=(const BfsHandle&) = delete;
    ~BfsHandle() {
      if (b) vgx_bfs_destroy(b);
    }
  };
  struct BfsCache {
    std::mutex mu;
    BfsHandle h;
  };
  mutable DynamicGraph g_{0, false};
  // declared after g_, so the traversal object goes before the graph it reads
  std::unique_ptr<BfsCache> bfs_ = std::make_unique<BfsCache>();
};

}  // namespace velographx
