// Foreign graphs through graph_access.hpp: the BasicIncremental* templates
// on graph types of the caller's own, whose algorithms run on a device copy
// kept in step with the caller's graph.
//
//   SetGraph   directed or undirected, adapted by members only (neighbors()
//              returning ids, its own apply()); stores self-loops and grows
//              its vertex count for them, which the device graph does not
//   RowGraph   adapted by ADL only, with no version (every recompute reloads)
//
// Each step is checked against a plain host BFS over the caller's graph, and
// PageRank / components / k-core against the same algorithms on a
// DynamicGraph loaded with the caller's arcs.  Exit status 0 when every
// check holds; a failed check prints its line and exits 1.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <random>
#include <set>
#include <utility>
#include <vector>

#include "velographx/graph_access.hpp"
#include "velographx/incremental/bfs.hpp"
#include "velographx/incremental/connected_components.hpp"
#include "velographx/incremental/kcore.hpp"
#include "velographx/incremental/pagerank.hpp"
#include "velographx/incremental/sssp.hpp"
#include "velographx/storage/dynamic_graph.hpp"

using namespace velographx;

#define CHECK(c)                                                              \
  do {                                                                        \
    if (!(c)) {                                                               \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      std::exit(1);                                                           \
    }                                                                         \
  } while (0)

namespace mine {

struct SetGraph {
  std::vector<std::set<VertexId>> out;
  std::uint64_t ver = 0;
  bool dir = true;

  std::size_t vertex_count() const { return out.size(); }
  bool directed() const { return dir; }
  std::uint64_t version() const { return ver; }
  std::vector<VertexId> neighbors(VertexId u) const {
    if (u >= out.size()) return {};
    return {out[u].begin(), out[u].end()};
  }
  void apply(const UpdateBatch& b) {
    for (const auto& e : b.updates) {
      const std::size_t need = static_cast<std::size_t>(std::max(e.src, e.dst)) + 1;
      if (out.size() < need) out.resize(need);
      if (e.add) out[e.src].insert(e.dst);
      else out[e.src].erase(e.dst);
      if (!dir && e.src != e.dst) {
        if (e.add) out[e.dst].insert(e.src);
        else out[e.dst].erase(e.src);
      }
    }
    ++ver;
  }
};

struct RowGraph {
  std::vector<std::vector<VertexId>> rows;
};
std::size_t vx_vertex_count(const RowGraph& g) { return g.rows.size(); }
bool vx_is_directed(const RowGraph&) { return true; }
template <class Fn>
void vx_for_each_neighbor(const RowGraph& g, VertexId u, Fn&& fn) {
  if (u < g.rows.size())
    for (auto v : g.rows[u]) fn(v);
}

}  // namespace mine

static_assert(MutableGraph<mine::SetGraph>);
static_assert(ReadableGraph<mine::RowGraph> && !MutableGraph<mine::RowGraph>);

template <class G>
static std::vector<std::uint32_t> host_bfs(const G& g, VertexId s) {
  const std::size_t n = vertex_count(g);
  std::vector<std::uint32_t> d(n, kUnreachable);
  if (s >= n) return d;
  std::deque<VertexId> q{s};
  d[s] = 0;
  while (!q.empty()) {
    const VertexId u = q.front();
    q.pop_front();
    for_each_neighbor(g, u, [&](VertexId v) {
      if (v < n && d[v] == kUnreachable) {
        d[v] = d[u] + 1;
        q.push_back(v);
      }
    });
  }
  return d;
}

// A DynamicGraph with the caller's arcs (loops dropped) and vertex count.
template <class G>
static DynamicGraph device_twin(const G& g) {
  DynamicGraph d(vertex_count(g), is_directed(g));
  std::vector<std::uint32_t> s, t;
  for (std::size_t u = 0; u < vertex_count(g); ++u)
    for_each_neighbor(g, static_cast<VertexId>(u), [&](VertexId v) {
      s.push_back(static_cast<std::uint32_t>(u));
      t.push_back(v);
    });
  d.bulk_load_edges(s.data(), t.data(), s.size());
  if (vertex_count(g) > d.vertex_count()) d.ensure_vertex(static_cast<VertexId>(vertex_count(g) - 1));
  return d;
}

static UpdateBatch random_batch(std::mt19937_64& rng, std::uint32_t id_bound, int ops, double add_p) {
  UpdateBatch b;
  std::uniform_int_distribution<std::uint32_t> id(0, id_bound - 1);
  std::bernoulli_distribution add(add_p), loop(0.05);
  for (int i = 0; i < ops; ++i) {
    const auto u = id(rng);
    const auto v = loop(rng) ? u : id(rng);
    if (add(rng)) b.add(u, v);
    else b.remove(u, v);
  }
  return b;
}

static double l1(const std::vector<double>& a, const std::vector<double>& b) {
  CHECK(a.size() == b.size());
  double s = 0;
  for (std::size_t i = 0; i < a.size(); ++i) s += std::abs(a[i] - b[i]);
  return s;
}

static void directed_set_graph() {
  std::mt19937_64 rng(7);
  mine::SetGraph g;
  g.dir = true;
  g.apply(random_batch(rng, 60, 300, 1.0));
  BasicIncrementalBFS<mine::SetGraph> bfs(g, 0);
  BasicIncrementalSSSP<mine::SetGraph> sssp(g, 0);
  BasicIncrementalPageRank<mine::SetGraph> pr(g);
  CHECK(bfs.distances() == host_bfs(g, 0));
  for (int step = 0; step < 24; ++step) {
    // ids past the vertex count grow the graph, some only through a loop
    const auto b = random_batch(rng, 64 + 2 * step, 30, 0.6);
    // as with the reference's templates, every algorithm takes the batch and
    // applies it to the graph (the second and third time change no arc)
    bfs.apply(b);
    sssp.apply(b);
    pr.apply(b);
    const auto want = host_bfs(g, 0);
    CHECK(bfs.distances() == want);
    CHECK(bfs.distances().size() == g.vertex_count());
    {
      auto twin = device_twin(g);
      IncrementalPageRank ref(twin);
      // the localized repair is within the reference's validation bound of
      // the full solve; after a recompute the two solves agree closely
      CHECK(l1(pr.values(), ref.values()) < 1e-6);
      if (step % 6 == 5) {
        pr.recompute();
        CHECK(l1(pr.values(), ref.values()) < 1e-10);
      }
      const auto sd = sssp.distances();
      CHECK(sd.size() == want.size());
      for (std::size_t i = 0; i < want.size(); ++i)
        CHECK(sd[i] == (want[i] == kUnreachable ? incremental_detail::kDijkstraInf : want[i]));
    }
  }
  // a change made to the graph directly is picked up by recompute()
  UpdateBatch cut;
  for (auto v : g.neighbors(0)) cut.remove(0, v);
  g.apply(cut);
  bfs.recompute();
  CHECK(bfs.distances() == host_bfs(g, 0));
  CHECK(bfs.reachable_count() == 1);
}

static void undirected_set_graph() {
  std::mt19937_64 rng(11);
  mine::SetGraph g;
  g.dir = false;
  g.apply(random_batch(rng, 80, 150, 1.0));
  BasicIncrementalBFS<mine::SetGraph> bfs(g, 3);
  BasicIncrementalComponents<mine::SetGraph> cc(g);
  BasicIncrementalKCore<mine::SetGraph> kc(g);
  for (int step = 0; step < 16; ++step) {
    const auto b = random_batch(rng, 84, 25, 0.55);
    bfs.apply(b);
    cc.apply(b);
    kc.apply(b);
    CHECK(bfs.distances() == host_bfs(g, 3));
    auto twin = device_twin(g);
    IncrementalComponents tc(twin);
    IncrementalKCore tk(twin);
    for (VertexId v = 0; v < g.vertex_count(); ++v)
      for (VertexId w : {VertexId{0}, VertexId{3}, static_cast<VertexId>(v / 2)})
        CHECK((cc.component(v) == cc.component(w)) == (tc.component(v) == tc.component(w)));
    CHECK(kc.core() == tk.core());
  }
}

static void unversioned_row_graph() {
  mine::RowGraph g{{{1}, {2}, {3}, {}, {0}}};
  BasicIncrementalBFS<mine::RowGraph> bfs(g, 0);
  CHECK((bfs.distances() == std::vector<std::uint32_t>{0, 1, 2, 3, kUnreachable}));
  g.rows[3].push_back(4);
  g.rows.push_back({});
  bfs.recompute();  // no version: the copy is reloaded
  CHECK((bfs.distances() == std::vector<std::uint32_t>{0, 1, 2, 3, 4, kUnreachable}));
}

int main() {
  directed_set_graph();
  undirected_set_graph();
  unversioned_row_graph();
  std::puts("facade_foreign: all checks passed");
  return 0;
}
