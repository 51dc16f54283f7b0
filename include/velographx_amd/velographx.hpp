// velographx.hpp -- C++20 facade with the reference engine's class names
// (include/velographx/storage/dynamic_graph.hpp, incremental/bfs.hpp,
// incremental/pagerank.hpp) over the gfx950 C ABI in vgx.h.
//
// DynamicGraph owns an HBM-resident graph; IncrementalBFS and
// IncrementalPageRank borrow it (non-owning, like the reference's Graph&
// members) and, as in the reference, their apply() mutates the graph.
// C ABI status codes are rethrown as the std exception types the reference
// throws, so pybind11 maps them to the same Python exceptions.
#pragma once

#include <cstdint>
#include <limits>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vgx.h"

namespace velographx {

using VertexId = std::uint32_t;

inline void vgx_check(vgx_status s) {
  if (s == VGX_OK) return;
  std::string msg = vgx_last_error();
  switch (s) {
    case VGX_INVALID_ARGUMENT:
      throw std::invalid_argument(msg);
    case VGX_OUT_OF_RANGE:
      throw std::out_of_range(msg);
    case VGX_OUT_OF_MEMORY:
      throw std::bad_alloc();
    default:
      throw std::runtime_error(msg);
  }
}

// dynamic_graph.hpp:17-22
struct EdgeUpdate {
  VertexId src{};
  VertexId dst{};
  bool add{true};
  std::uint64_t timestamp{0};
};

// dynamic_graph.hpp:24-29
struct UpdateBatch {
  std::vector<EdgeUpdate> updates;
  void add(VertexId u, VertexId v, std::uint64_t ts = 0) { updates.push_back({u, v, true, ts}); }
  void remove(VertexId u, VertexId v, std::uint64_t ts = 0) { updates.push_back({u, v, false, ts}); }
  [[nodiscard]] bool empty() const noexcept { return updates.empty(); }

  // SoA view handed to the C ABI (timestamps are not used by the path).
  struct Columns {
    std::vector<std::uint32_t> src, dst;
    std::vector<std::uint8_t> add;
  };
  [[nodiscard]] Columns columns() const {
    Columns c;
    c.src.reserve(updates.size());
    c.dst.reserve(updates.size());
    c.add.reserve(updates.size());
    for (const auto& u : updates) {
      c.src.push_back(u.src);
      c.dst.push_back(u.dst);
      c.add.push_back(u.add ? 1 : 0);
    }
    return c;
  }
};

class DynamicGraph {
 public:
  explicit DynamicGraph(std::size_t vertices = 0, bool directed = false) {
    vgx_check(vgx_graph_create(vertices, directed ? 1 : 0, &g_));
  }
  ~DynamicGraph() {
    if (g_) vgx_graph_destroy(g_);
  }
  DynamicGraph(const DynamicGraph&) = delete;
  DynamicGraph& operator=(const DynamicGraph&) = delete;
  DynamicGraph(DynamicGraph&& o) noexcept : g_(std::exchange(o.g_, nullptr)) {}
  DynamicGraph& operator=(DynamicGraph&& o) noexcept {
    if (this != &o) {
      if (g_) vgx_graph_destroy(g_);
      g_ = std::exchange(o.g_, nullptr);
    }
    return *this;
  }

  [[nodiscard]] std::size_t vertex_count() const { return info().n; }
  [[nodiscard]] std::uint64_t version() const { return info().version; }
  [[nodiscard]] bool directed() const { return info().directed; }
  [[nodiscard]] std::size_t edge_count_directed() const { return info().arcs; }

  void bulk_load_edges(const std::vector<std::pair<VertexId, VertexId>>& edges) {
    std::vector<std::uint32_t> s(edges.size()), d(edges.size());
    for (std::size_t i = 0; i < edges.size(); ++i) {
      s[i] = edges[i].first;
      d[i] = edges[i].second;
    }
    vgx_check(vgx_graph_bulk_load(g_, s.data(), d.data(), s.size()));
  }

  void bulk_load_edges(const std::uint32_t* src, const std::uint32_t* dst, std::size_t m) {
    vgx_check(vgx_graph_bulk_load(g_, src, dst, m));
  }

  void apply(const UpdateBatch& batch) {
    if (batch.empty()) return;
    auto c = batch.columns();
    vgx_check(vgx_graph_apply(g_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
  }

  void add_edge(VertexId u, VertexId v) {
    UpdateBatch b;
    b.add(u, v);
    apply(b);
  }
  void remove_edge(VertexId u, VertexId v) {
    UpdateBatch b;
    b.remove(u, v);
    apply(b);
  }

  [[nodiscard]] std::vector<VertexId> neighbors(VertexId u) const { return row(u, 0); }
  [[nodiscard]] std::vector<VertexId> in_neighbors(VertexId v) const { return row(v, 1); }

  [[nodiscard]] bool has_edge(VertexId u, VertexId v) const {
    int p = 0;
    vgx_check(vgx_graph_has_edge(g_, u, v, &p));
    return p != 0;
  }

  [[nodiscard]] std::vector<std::uint32_t> degrees(bool in_direction = false) const {
    std::vector<std::uint32_t> d(vertex_count());
    vgx_check(vgx_graph_degrees(g_, in_direction ? 1 : 0, d.data()));
    return d;
  }

  void compact() { vgx_check(vgx_graph_compact(g_)); }

  [[nodiscard]] vgx_graph_storage storage() const {
    vgx_graph_storage s{};
    vgx_check(vgx_graph_storage_info(g_, &s));
    return s;
  }

  // dynamic_graph.hpp:599-606: device bytes held by the graph.
  [[nodiscard]] std::size_t storage_bytes() const { return storage().storage_bytes; }

  // dynamic_graph.hpp:569-571.  Rows are updated in place here, so the
  // graph is compact when no row holds a tombstone and no row has moved
  // since the last bulk load, compaction or snapshot.
  [[nodiscard]] bool is_compact() const {
    const auto s = storage();
    return s.tombstones == 0 && s.relocated_rows == 0;
  }

  [[nodiscard]] vgx_graph* handle() const noexcept { return g_; }

  // Takes ownership of a graph handle made by the C ABI (vgx_graph_snapshot).
  [[nodiscard]] static DynamicGraph adopt(vgx_graph* g) noexcept { return DynamicGraph(g, Adopt{}); }

 private:
  struct Adopt {};
  DynamicGraph(vgx_graph* g, Adopt) noexcept : g_(g) {}
  struct Info {
    std::uint64_t n, version, arcs;
    bool directed;
  };
  [[nodiscard]] Info info() const {
    Info i{};
    int d = 0;
    vgx_check(vgx_graph_info(g_, &i.n, &i.version, &d, &i.arcs));
    i.directed = d != 0;
    return i;
  }
  [[nodiscard]] std::vector<VertexId> row(VertexId u, int dir) const {
    std::uint64_t count = 0;
    vgx_check(vgx_graph_neighbors(g_, u, dir, nullptr, 0, &count));
    std::vector<VertexId> out(count);
    if (count) vgx_check(vgx_graph_neighbors(g_, u, dir, out.data(), count, &count));
    return out;
  }
  vgx_graph* g_ = nullptr;
};

// bfs.hpp:13-27, 383
class IncrementalBFS {
 public:
  static constexpr std::uint32_t unreachable = std::numeric_limits<std::uint32_t>::max();

  IncrementalBFS(DynamicGraph& g, VertexId source, double deletion_fallback_fraction = 0.35) {
    vgx_check(vgx_bfs_create(g.handle(), source, deletion_fallback_fraction, &b_));
  }
  ~IncrementalBFS() {
    if (b_) vgx_bfs_destroy(b_);
  }
  IncrementalBFS(const IncrementalBFS&) = delete;
  IncrementalBFS& operator=(const IncrementalBFS&) = delete;

  [[nodiscard]] std::vector<std::uint32_t> distances() const {
    std::uint64_t count = 0;
    vgx_check(vgx_bfs_distances(b_, nullptr, 0, &count));
    std::vector<std::uint32_t> d(count);
    if (count) vgx_check(vgx_bfs_distances(b_, d.data(), count, &count));
    return d;
  }
  [[nodiscard]] std::size_t reachable_count() const { return stats().reachable_count; }
  [[nodiscard]] std::size_t last_deletion_candidates() const { return stats().last_deletion_candidates; }
  [[nodiscard]] std::size_t last_affected_vertices() const { return stats().last_affected_vertices; }
  [[nodiscard]] bool last_used_full_recompute() const { return stats().last_used_full_recompute != 0; }

  void apply(const UpdateBatch& batch) {
    auto c = batch.columns();
    vgx_check(vgx_bfs_apply(b_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
  }
  void recompute() { vgx_check(vgx_bfs_recompute(b_)); }

  [[nodiscard]] vgx_bfs_stats stats() const {
    vgx_bfs_stats s{};
    vgx_check(vgx_bfs_stats_get(b_, &s));
    return s;
  }
  [[nodiscard]] vgx_bfs* handle() const noexcept { return b_; }

 private:
  vgx_bfs* b_ = nullptr;
};

// Addition without a reference counterpart: up to 64 roots traversed in one
// pass (the reference runs independent roots on threads,
// benchmarks/thread_scaling_bfs.cpp:78-108).
class MultiSourceBFS {
 public:
  MultiSourceBFS(DynamicGraph& g, const std::vector<VertexId>& roots) : n_(g.vertex_count()), k_(roots.size()) {
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
  vgx_msbfs* m_ = nullptr;
  std::size_t n_ = 0, k_ = 0;
};

// pagerank.hpp:16-27
struct PageRankValidation {
  double l1_error{0.0};
  double linf_error{0.0};
  double local_residual_l1{0.0};
  double local_residual_linf{0.0};
  double reference_residual_l1{0.0};
  double reference_residual_linf{0.0};
  std::size_t reference_iterations{0};
  bool reference_converged{false};
  bool within_tolerance{false};
  bool fallback_applied{false};
};

}  // namespace velographx
