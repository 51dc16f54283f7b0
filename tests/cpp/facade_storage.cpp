// DynamicGraph's storage introspection and row compaction through the
// facade, and algorithm objects that outlive a compaction of their graph.
//
//   - the getter checks of the reference's tests/test_dynamic.cpp:25-47 and
//     :148-184 (delta_edge_count, is_compact, dirty_out/in_segment_count),
//     restated here since that program also needs the triangle counter and
//     the execution planner, which are not part of this library;
//   - maybe_compact(threshold): which rows compact, what they keep, and that
//     the logical graph and version do not change;
//   - BFS, multi-source BFS and PageRank objects created before compact() or
//     maybe_compact() of their graph, recomputed after it, against fresh
//     objects (rows move and shorten under them).
// Exit status 0 when every check holds.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <utility>
#include <vector>

#include "velographx/incremental/bfs.hpp"
#include "velographx/incremental/pagerank.hpp"
#include "velographx/storage/dynamic_graph.hpp"

using namespace velographx;

#define CHECK(c)                                                              \
  do {                                                                        \
    if (!(c)) {                                                               \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      std::exit(1);                                                           \
    }                                                                         \
  } while (0)

static void reference_getters() {
  // test_dynamic.cpp:25-35
  DynamicGraph g(4, false);
  UpdateBatch b;
  b.add(0, 1);
  b.add(1, 2);
  b.add(2, 0);
  g.apply(b);
  CHECK(g.version() == 1);
  CHECK(g.has_edge(0, 1));
  CHECK(g.edge_count_directed() == 6);
  CHECK(!g.is_compact());
  CHECK(g.delta_edge_count() == 6);
  CHECK(g.base_edge_count_directed() == 0);
  // test_dynamic.cpp:37-47: reading the graph through an algorithm leaves
  // its storage as it was
  const auto base0 = g.base_edge_count_directed();
  const auto delta0 = g.delta_edge_count();
  const auto v0 = g.version();
  IncrementalBFS bfs(g, 0);
  CHECK((bfs.distances() == std::vector<std::uint32_t>{0, 1, 1, kUnreachable}));
  CHECK(!g.is_compact() && g.base_edge_count_directed() == base0 && g.delta_edge_count() == delta0 &&
        g.version() == v0);

  // test_dynamic.cpp:148-184
  std::vector<std::pair<VertexId, VertexId>> seed;
  for (VertexId i = 0; i < 200; ++i) {
    seed.emplace_back(i, 1000 + i);
    seed.emplace_back(65536 + i, 67000 + i);
  }
  DynamicGraph seg(131072, true);
  seg.bulk_load_edges(seed);
  CHECK(seg.is_compact());
  CHECK(seg.dirty_out_segment_count() == 0 && seg.dirty_in_segment_count() == 0);
  CHECK(seg.base_edge_count_directed() == 400 && seg.delta_edge_count() == 0);
  UpdateBatch sparse;
  sparse.add(10, 50000);
  seg.apply(sparse);
  CHECK(!seg.is_compact());
  CHECK(seg.dirty_out_segment_count() == 1 && seg.dirty_in_segment_count() == 1);
  CHECK(seg.has_edge(10, 50000));
  sparse = {};
  sparse.add(65540, 120000);
  seg.apply(sparse);
  CHECK(seg.dirty_out_segment_count() == 2 && seg.dirty_in_segment_count() == 2);
  const auto count = seg.edge_count_directed();
  seg.compact();
  CHECK(seg.is_compact());
  CHECK(seg.dirty_out_segment_count() == 0 && seg.dirty_in_segment_count() == 0);
  CHECK(seg.edge_count_directed() == count && seg.base_edge_count_directed() == count);
  CHECK(seg.has_edge(10, 50000) && seg.has_edge(65540, 120000));

  // a base arc deleted and re-added is no change; one inserted and deleted
  // again counts as an inserted entry plus a tombstone
  DynamicGraph t(8, true);
  t.bulk_load_edges({{0, 1}, {0, 2}});
  t.remove_edge(0, 1);
  CHECK(t.delta_edge_count() == 1 && t.dirty_out_segment_count() == 1);
  t.add_edge(0, 1);
  CHECK(t.delta_edge_count() == 0 && t.is_compact());
  t.add_edge(0, 5);
  t.remove_edge(0, 5);
  CHECK(t.delta_edge_count() == 2 && t.edge_count_directed() == 2);
  CHECK(std::abs(t.delta_ratio() - 1.0) < 1e-15);
  t.ensure_vertex(20);
  CHECK(t.vertex_count() == 21 && t.neighbors(20).empty());
  t.ensure_vertex(3);
  CHECK(t.vertex_count() == 21);
}

static void maybe_compact_rows() {
  // row 0: 40 base arcs, 30 deleted (delta 30 / 40); row 1: 40 base arcs, 2
  // deleted (2 / 40); row 2: empty base, 3 inserted (3 / max(8, 0))
  DynamicGraph g(200, true);
  std::vector<std::pair<VertexId, VertexId>> e;
  for (VertexId v = 0; v < 40; ++v) {
    e.emplace_back(0, 100 + v);
    e.emplace_back(1, 100 + v);
  }
  g.bulk_load_edges(e);
  UpdateBatch b;
  for (VertexId v = 0; v < 30; ++v) b.remove(0, 100 + v);
  b.remove(1, 100);
  b.remove(1, 101);
  for (VertexId v = 0; v < 3; ++v) b.add(2, 150 + v);
  g.apply(b);
  const auto version = g.version();
  const auto arcs = g.edge_count_directed();
  const auto n0 = g.neighbors(0), n1 = g.neighbors(1), n2 = g.neighbors(2);
  CHECK(g.dirty_out_segment_count() == 3);
  CHECK(g.delta_edge_count() == 30 + 2 + 3);
  // out-rows 0 (30 / 40) and 2 (3 / 8) reach 0.3, row 1 (2 / 40) does not.
  // The in-rows changed are 100..129 (one or, for 100 and 101, two
  // tombstones against a base of 2, so at most 2 / 8) and 150..152 (1 / 8):
  // none of them reaches 0.3
  CHECK(g.dirty_in_segment_count() == 33);
  CHECK(g.maybe_compact(0.3));
  CHECK(g.dirty_out_segment_count() == 1);
  CHECK(g.dirty_in_segment_count() == 33);
  CHECK(g.delta_edge_count() == 2);
  CHECK(g.version() == version && g.edge_count_directed() == arcs);
  CHECK(g.neighbors(0) == n0 && g.neighbors(1) == n1 && g.neighbors(2) == n2);
  CHECK(g.storage().tombstones == 2 + 32);  // out-row 1, in-rows 100..129
  CHECK(!g.maybe_compact(0.3));  // nothing left at that density
  CHECK(g.maybe_compact(0.0));   // every row with a delta
  CHECK(g.is_compact() && g.storage().tombstones == 0 && g.neighbors(1) == n1);
}

static std::vector<std::pair<VertexId, VertexId>> random_edges(std::mt19937_64& rng, VertexId n, int m) {
  std::uniform_int_distribution<VertexId> id(0, n - 1);
  std::vector<std::pair<VertexId, VertexId>> e;
  for (int i = 0; i < m; ++i) e.emplace_back(id(rng), id(rng));
  return e;
}

static void objects_across_compaction() {
  std::mt19937_64 rng(5);
  const VertexId n = 12000;
  for (const bool directed : {false, true}) {
    for (const int mode : {0, 1}) {  // 0: compact(), 1: maybe_compact(0)
      DynamicGraph g(n, directed);
      // vertex 0's in-row is longer than one PageRank work chunk
      auto e = random_edges(rng, n, 60000);
      for (VertexId v = 1; v < n; ++v) e.emplace_back(v, 0);
      g.bulk_load_edges(e);
      IncrementalBFS bfs(g, 1);
      IncrementalPageRank pr(g);
      const std::vector<VertexId> roots{1, 7, 99};
      MultiSourceBFS ms(g, roots);
      ms.run(g.vertex_count());
      // deletions leave tombstones at the front of many in-rows (sources of
      // low id) and all along vertex 0's
      UpdateBatch b;
      for (VertexId u = 0; u < n; u += 3) {
        const auto row = g.neighbors(u);
        for (std::size_t i = 0; i < row.size() && i < 3; ++i) b.remove(u, row[i]);
      }
      for (const auto& [u, v] : random_edges(rng, n, 500)) b.add(u, v);
      bfs.apply(b);
      pr.apply(b);
      CHECK(g.storage().tombstones > 0);
      if (mode == 0) g.compact();
      else CHECK(g.maybe_compact(0.0));
      bfs.recompute();
      pr.recompute();
      ms.run(g.vertex_count());
      IncrementalBFS fresh(g, 1);
      CHECK(bfs.distances() == fresh.distances());
      IncrementalPageRank fresh_pr(g);
      const auto a = pr.values(), c = fresh_pr.values();
      double l1 = 0;
      for (std::size_t i = 0; i < a.size(); ++i) l1 += std::abs(a[i] - c[i]);
      CHECK(a.size() == c.size() && l1 < 1e-9);
      for (std::uint32_t r = 0; r < roots.size(); ++r) {
        IncrementalBFS one(g, roots[r]);
        CHECK(ms.distances(r) == one.distances());
      }
    }
  }
}

int main() {
  reference_getters();
  maybe_compact_rows();
  objects_across_compaction();
  std::puts("facade_storage: all checks passed");
  return 0;
}
