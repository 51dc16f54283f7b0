// DynamicGraph, EdgeUpdate and UpdateBatch -- the interface of
// include/velographx/storage/dynamic_graph.hpp over the gfx950 C ABI in
// vgx.h.  The graph's rows live in HBM; every accessor that returns ids
// copies them to the host.  C ABI status codes are rethrown as the std
// exception types the reference throws, so pybind11 maps them to the same
// Python exceptions.
#pragma once

#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <type_traits>
#include <vector>

#include "velographx_amd/types.hpp"
#include "vgx.h"

namespace velographx {

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

  // dynamic_graph.hpp:489-491
  [[nodiscard]] std::size_t vertex_count() const { return info().n; }
  [[nodiscard]] std::uint64_t version() const { return info().version; }
  [[nodiscard]] bool directed() const { return info().directed; }

  // dynamic_graph.hpp:493-500
  void ensure_vertex(VertexId v) { vgx_check(vgx_graph_ensure_vertex(g_, v)); }

  // dynamic_graph.hpp:502-529
  void bulk_load_edges(const std::vector<std::pair<VertexId, VertexId>>& edges) {
    // a pair of two 32-bit ids is the (src, dst) layout the pair entry point reads
    static_assert(sizeof(std::pair<VertexId, VertexId>) == 2 * sizeof(std::uint32_t) &&
                  std::is_standard_layout_v<std::pair<VertexId, VertexId>>);
    vgx_check(vgx_graph_bulk_load_pairs(g_, reinterpret_cast<const std::uint32_t*>(edges.data()), edges.size()));
  }
  void bulk_load_edges(const std::uint32_t* src, const std::uint32_t* dst, std::size_t m) {
    vgx_check(vgx_graph_bulk_load(g_, src, dst, m));
  }
  // The same from one array of m (src, dst) pairs (pairs[2i], pairs[2i + 1]).
  void bulk_load_pairs(const std::uint32_t* pairs, std::size_t m) {
    vgx_check(vgx_graph_bulk_load_pairs(g_, pairs, m));
  }

  // dynamic_graph.hpp:531-549
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
  void apply(const UpdateBatch& batch) {
    if (batch.empty()) return;
    auto c = batch.columns();
    vgx_check(vgx_graph_apply(g_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
  }

  // dynamic_graph.hpp:551-567: ids ascending, copied from the device row.
  [[nodiscard]] std::vector<VertexId> neighbors(VertexId u) const { return row(u, 0); }
  [[nodiscard]] std::vector<VertexId> in_neighbors(VertexId v) const { return row(v, 1); }
  template <class Fn>
  void for_each_neighbor(VertexId u, Fn&& fn) const {
    for (const VertexId v : row(u, 0)) fn(v);
  }
  template <class Fn>
  void for_each_in_neighbor(VertexId v, Fn&& fn) const {
    for (const VertexId u : row(v, 1)) fn(u);
  }

  // dynamic_graph.hpp:583-587
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

  // dynamic_graph.hpp:589-597
  [[nodiscard]] std::size_t edge_count_directed() const { return info().arcs; }
  [[nodiscard]] std::size_t base_edge_count_directed() const { return delta().base_arcs; }
  // Rows are updated in place, so a row's delta is its entries inserted
  // since its last bulk load or compaction plus its tombstones (an arc
  // inserted and deleted again in between counts twice).
  [[nodiscard]] std::size_t delta_edge_count() const { return delta().delta_arcs; }

  // dynamic_graph.hpp:599-606: device bytes held by the graph.
  [[nodiscard]] std::size_t storage_bytes() const { return storage().storage_bytes; }

  // dynamic_graph.hpp:607-618
  [[nodiscard]] double delta_ratio() const {
    const auto d = delta();
    return static_cast<double>(d.delta_arcs) / static_cast<double>(d.base_arcs > 0 ? d.base_arcs : 1);
  }
  [[nodiscard]] std::size_t dirty_out_segment_count() const { return delta().dirty_out_rows; }
  [[nodiscard]] std::size_t dirty_in_segment_count() const { return delta().dirty_in_rows; }

  // dynamic_graph.hpp:569-571: no row differs from its last bulk load or
  // compaction.
  [[nodiscard]] bool is_compact() const {
    const auto d = delta();
    return d.delta_arcs == 0 && d.delta_in_arcs == 0;
  }

  // dynamic_graph.hpp:620-631: rows whose delta reaches `threshold` of
  // max(8, base length) drop their tombstones in place; true when any did.
  bool maybe_compact(double threshold = 0.25) {
    int done = 0;
    vgx_check(vgx_graph_maybe_compact(g_, threshold, &done));
    return done != 0;
  }

  // dynamic_graph.hpp:633-639: every row rewritten contiguously.
  void compact() { vgx_check(vgx_graph_compact(g_)); }

  [[nodiscard]] vgx_graph_storage storage() const {
    vgx_graph_storage s{};
    vgx_check(vgx_graph_storage_info(g_, &s));
    return s;
  }
  [[nodiscard]] vgx_graph_delta delta() const {
    vgx_graph_delta d{};
    vgx_check(vgx_graph_delta_info(g_, &d));
    return d;
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
  // One call for rows of up to 64 ids, a second one for longer rows.
  [[nodiscard]] std::vector<VertexId> row(VertexId u, int dir) const {
    std::vector<VertexId> out(64);
    std::uint64_t count = 0;
    vgx_check(vgx_graph_neighbors(g_, u, dir, out.data(), out.size(), &count));
    if (count > out.size()) {
      out.resize(count);
      vgx_check(vgx_graph_neighbors(g_, u, dir, out.data(), count, &count));
    }
    out.resize(count);
    return out;
  }
  vgx_graph* g_ = nullptr;
};

}  // namespace velographx
