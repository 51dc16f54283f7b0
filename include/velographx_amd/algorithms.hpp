// One-shot algorithms on a CsrGraph -- the bfs_distances, pagerank and
// connected_components entries of include/velographx/algorithms.hpp:8-10
// (src/algorithms.cpp:11-81), run by the device kernels over the CsrGraph's
// rows.  The reference's triangle and similarity entries are not part of
// the traversal path and are not provided.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "velographx_amd/bfs.hpp"
#include "velographx_amd/components.hpp"
#include "velographx_amd/csr_graph.hpp"

namespace velographx {

// algorithms.cpp:11-27: hop distances from source (kUnreachable where not
// reached); a source outside the graph throws std::out_of_range.
inline std::vector<std::uint32_t> bfs_distances(const CsrGraph& graph, VertexId source) {
  if (source >= graph.vertex_count()) throw std::out_of_range("source outside graph");
  std::vector<std::uint32_t> d(graph.vertex_count());
  graph.bfs_distances_into(source, d.data());
  return d;
}

// algorithms.cpp:28-54: component ids 0, 1, ... numbered in the order of
// each component's smallest vertex; a directed graph throws
// std::invalid_argument.  The device labels are those smallest vertices,
// so the id is the label's rank among the distinct labels.
inline std::vector<VertexId> connected_components(const CsrGraph& graph) {
  if (graph.directed())
    throw std::invalid_argument(
        "connected_components requires an undirected graph; choose weak or strong connectivity explicitly for "
        "directed graphs");
  const IncrementalComponents cc(graph.device_graph());
  std::vector<VertexId> id = cc.labels();
  std::vector<VertexId> reps = id;
  std::sort(reps.begin(), reps.end());
  reps.erase(std::unique(reps.begin(), reps.end()), reps.end());
  for (auto& x : id) x = static_cast<VertexId>(std::lower_bound(reps.begin(), reps.end(), x) - reps.begin());
  return id;
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
