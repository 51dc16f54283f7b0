// Static CSR graph and its one-shot algorithms on the device -- the
// interface of include/velographx/csr_graph.hpp and the bfs_distances /
// pagerank entries of include/velographx/algorithms.hpp:8-10
// (src/csr_graph.cpp, src/algorithms.cpp:11-27, 56-81).
//
// A CsrGraph is a device graph built once by the bulk load (self-loops
// dropped, duplicates removed, both directions kept, vertex count = largest
// id + 1) and never updated; bfs_distances runs the direction-optimizing
// BFS kernels and pagerank the PageRank sweep kernels of the dynamic graph
// over its rows.  Row accessors return copies (the rows live in HBM).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

#include "velographx_amd/velographx.hpp"

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

// algorithms.cpp:11-27: hop distances from source (kUnreachable where not
// reached); a source outside the graph throws std::out_of_range.
inline std::vector<std::uint32_t> bfs_distances(const CsrGraph& graph, VertexId source) {
  if (source >= graph.vertex_count()) throw std::out_of_range("source outside graph");
  IncrementalBFS bfs(graph.device_graph(), source);
  return bfs.distances();
}

// algorithms.cpp:56-81: power iteration from 1/n with uniform teleport and
// the dangling mass spread evenly, stopping once the L1 change is below
// `tolerance`; damping outside [0, 1] throws std::invalid_argument.
inline std::vector<double> pagerank(const CsrGraph& graph, double damping = 0.85, std::size_t max_iterations = 100,
                                    double tolerance = 1e-10) {
  if (!(damping >= 0.0 && damping <= 1.0)) throw std::invalid_argument("PageRank damping must be in [0, 1]");
  std::vector<double> ranks(graph.vertex_count());
  std::uint64_t count = 0;
  vgx_check(vgx_pagerank_static(graph.device_graph().handle(), damping, max_iterations, tolerance, ranks.data(),
                                ranks.size(), &count, nullptr));
  ranks.resize(count);
  return ranks;
}

}  // namespace velographx
