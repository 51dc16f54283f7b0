// How an incremental algorithm reaches the device graph its kernels run on.
//
// The reference's algorithm templates (incremental/bfs.hpp:13-383 and its
// siblings) read any graph type through graph_access.hpp.  Here every
// algorithm runs on a DynamicGraph in HBM, so GraphBinding<G> supplies one:
//   - DynamicGraph: the caller's graph itself;
//   - CsrGraph: its device graph (a CsrGraph takes no updates, so, as in
//     the reference, apply() on an algorithm bound to one does not compile);
//   - any other ReadableGraph: a device copy, loaded through
//     for_each_neighbor, that apply() keeps in step with the caller's graph.
//     The caller's graph is updated with apply_updates (its own semantics),
//     then the out-rows of the batch's endpoints are read back and compared
//     with the copy's; the differences are the update batch the device
//     algorithm applies.  recompute() reloads the copy whenever the graph's
//     version moved since the copy last matched it (always, for a graph
//     without a version).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "velographx_amd/csr_graph.hpp"
#include "velographx_amd/dynamic_graph.hpp"
#include "velographx_amd/graph_access.hpp"

namespace velographx::detail {

template <class G>
class GraphBinding {
  static_assert(ReadableGraph<G>, "an algorithm's graph must model ReadableGraph (graph_access.hpp)");

 public:
  static constexpr bool native = false;

  explicit GraphBinding(G& g) : g_(g), dev_(0, is_directed(g)) { load(); }

  [[nodiscard]] DynamicGraph& device() noexcept { return dev_; }
  [[nodiscard]] std::size_t vertex_count() const { return velographx::vertex_count(g_); }

  // Updates the caller's graph and returns the ops that bring the device
  // copy to the same arc set (and vertex count).
  const UpdateBatch& forward(const UpdateBatch& batch) {
    std::vector<VertexId> rows;
    rows.reserve(2 * batch.updates.size());
    for (const auto& e : batch.updates) {
      rows.push_back(e.src);
      rows.push_back(e.dst);
    }
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    apply_updates(g_, batch);
    const std::size_t n = velographx::vertex_count(g_);
    delta_.updates.clear();
    std::vector<VertexId> now;
    for (const VertexId u : rows) {
      if (u >= n) continue;
      now.clear();
      velographx::for_each_neighbor(g_, u, [&](VertexId v) {
        if (v != u) now.push_back(v);  // the device graph keeps no self-loops
      });
      std::sort(now.begin(), now.end());
      now.erase(std::unique(now.begin(), now.end()), now.end());
      const std::vector<VertexId> was = dev_.neighbors(u);
      std::size_t i = 0, j = 0;
      while (i < was.size() || j < now.size()) {
        if (j == now.size() || (i < was.size() && was[i] < now[j])) {
          delta_.remove(u, was[i++]);
        } else if (i == was.size() || now[j] < was[i]) {
          delta_.add(u, now[j++]);
        } else {
          ++i;
          ++j;
        }
      }
    }
    grow_to(n);
    remember();
    return delta_;
  }

  // Reloads the copy unless it is known to match the caller's graph.
  void refresh() {
    if constexpr (MutableGraph<G>) {
      if (synced_ && graph_version(g_) == seen_) return;
    }
    load();
  }

 private:
  void load() {
    const std::size_t n = velographx::vertex_count(g_);
    std::vector<std::uint32_t> s, d;
    for (std::size_t u = 0; u < n; ++u) {
      const auto src = static_cast<VertexId>(u);
      velographx::for_each_neighbor(g_, src, [&](VertexId v) {
        s.push_back(src);
        d.push_back(v);
      });
    }
    dev_.bulk_load_edges(s.data(), d.data(), s.size());
    if (n > dev_.vertex_count()) dev_.ensure_vertex(static_cast<VertexId>(n - 1));
    remember();
  }

  // A vertex count the ops do not reach yet comes with one more op: the
  // removal of an absent arc ending at the last vertex, which grows the
  // device graph inside the algorithm's own batch (as the reference's
  // graph grows inside apply_updates) and changes no arc.
  void grow_to(std::size_t n) {
    std::size_t reach = dev_.vertex_count();
    for (const auto& e : delta_.updates) reach = std::max<std::size_t>(reach, std::max(e.src, e.dst) + std::size_t{1});
    if (n <= reach) return;
    const auto last = static_cast<VertexId>(n - 1);
    if (last == 0) {
      dev_.ensure_vertex(0);
      return;
    }
    // (0, last) is absent: vertex `last` is not in the device graph yet
    delta_.remove(0, last);
  }

  void remember() {
    if constexpr (MutableGraph<G>) {
      seen_ = graph_version(g_);
      synced_ = true;
    }
  }

  G& g_;
  DynamicGraph dev_;
  UpdateBatch delta_;
  std::uint64_t seen_ = 0;
  bool synced_ = false;
};

template <>
class GraphBinding<DynamicGraph> {
 public:
  static constexpr bool native = true;
  explicit GraphBinding(DynamicGraph& g) noexcept : g_(g) {}
  [[nodiscard]] DynamicGraph& device() const noexcept { return g_; }
  [[nodiscard]] std::size_t vertex_count() const { return g_.vertex_count(); }
  const UpdateBatch& forward(const UpdateBatch& batch) const noexcept { return batch; }
  void refresh() const noexcept {}

 private:
  DynamicGraph& g_;
};

template <>
class GraphBinding<CsrGraph> {
 public:
  static constexpr bool native = true;
  explicit GraphBinding(const CsrGraph& g) noexcept : g_(g) {}
  [[nodiscard]] DynamicGraph& device() const noexcept { return g_.device_graph(); }
  [[nodiscard]] std::size_t vertex_count() const { return g_.vertex_count(); }
  void refresh() const noexcept {}

 private:
  const CsrGraph& g_;
};

}  // namespace velographx::detail
