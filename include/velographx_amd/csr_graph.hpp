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

 private:
  void check(VertexId v) const {
    if (v >= vertex_count()) throw std::out_of_range("vertex id outside graph");
  }
  mutable DynamicGraph g_{0, false};
};

}  // namespace velographx
