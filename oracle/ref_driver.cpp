// Driver around the reference engine's own headers -- TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile from the reference sources where they lie
// (include/velographx/..., src/*.cpp) into oracle/_ref/vgx_ref.  It runs
// scripted scenarios through the reference's DynamicGraph / IncrementalBFS /
// IncrementalPageRank and prints the results, so tests can cross-check the
// C restatement in vgx_oracle.c; its bench modes time the reference engine
// as the CPU baseline.
//
// Scenario language (one command per line, whitespace separated):
//   graph <n> <directed>            new DynamicGraph(n, directed)
//   bulk <m>  + m lines "u v"       bulk_load_edges
//   batch <k> + k lines "u v add"   stage an UpdateBatch
//   graph_apply                     graph.apply(batch)
//   bfs_new <source> <fallback>     IncrementalBFS(graph, source, fallback)
//   bfs_apply                       bfs.apply(batch)
//   bfs_recompute                   bfs.recompute()
//   pr_new <damping>                IncrementalPageRank(graph, damping)
//   pr_apply <iters> <tol> <frac>   pr.apply(batch, iters, tol, frac)
//   pr_recompute <iters> <tol>      pr.recompute(iters, tol)
//   pr_validate <it> <tol> <l1> <linf>
//   neighbors <u>
//   sssp_new <source> / sssp_apply  IncrementalSSSP(graph, source) / apply(batch)
//   kcore_new / kcore_apply         IncrementalKCore(graph) / apply(batch)
//   cc_new / cc_apply               IncrementalComponents(graph) / apply(batch)
//   cons_policy <arcs> <latency>    scale_aware_consolidation_policy
//   cons_eval <cur> <canon> <cur_lat> <canon_lat> <max_s> <max_l>
//                                   evaluate_consolidation
//   cons_ctl <min_epochs> <samples> <growth>   new ConsolidationController
//   cons_observe <sr> <lr> <se> <le> <sc> <epoch>   controller.observe
//   cons_mark <epoch>               controller.mark_consolidated
// Bench modes: bench-bfs, bench-bfs-mt, bench-pr, bench-repair,
// bench-snapshot, bench-static, bench-kcore, bench-cc (see their comments).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "velographx/algorithms.hpp"
#include "velographx/csr_graph.hpp"
#include "velographx/incremental/bfs.hpp"
#include "velographx/incremental/connected_components.hpp"
#include "velographx/incremental/kcore.hpp"
#include "velographx/incremental/pagerank.hpp"
#include "velographx/incremental/sssp.hpp"
#include "velographx/storage/consolidation.hpp"
#include "velographx/storage/dynamic_graph.hpp"

using velographx::DynamicGraph;
using velographx::IncrementalBFS;
using velographx::IncrementalPageRank;
using velographx::UpdateBatch;
using velographx::VertexId;

namespace {

void print_dist(const IncrementalBFS& bfs) {
  const auto& d = bfs.distances();
  std::printf("dist %zu", d.size());
  for (auto x : d) std::printf(" %u", x);
  std::printf("\n");
  std::printf("bfsdiag %zu %zu %zu %d\n", bfs.reachable_count(), bfs.last_deletion_candidates(),
              bfs.last_affected_vertices(), bfs.last_used_full_recompute() ? 1 : 0);
}

void print_pr(const IncrementalPageRank& pr) {
  const auto& v = pr.values();
  std::printf("values %zu", v.size());
  for (auto x : v) std::printf(" %.17g", x);
  std::printf("\n");
  std::printf("prdiag %zu %zu %.17g %.17g %d\n", pr.last_repaired_vertices(), pr.last_repair_iterations(),
              pr.last_residual_l1(), pr.last_residual_linf(), pr.last_full_recompute_converged() ? 1 : 0);
}

void print_sssp(const velographx::IncrementalSSSP& s) {
  const auto& d = s.distances();
  std::printf("sdist %zu", d.size());
  for (auto x : d) std::printf(" %llu", static_cast<unsigned long long>(x));
  std::printf("\n");
}

void print_cc(velographx::IncrementalComponents& c, std::size_t n) {
  std::printf("cclabels %zu", n);
  for (std::size_t v = 0; v < n; ++v) std::printf(" %u", c.component(static_cast<VertexId>(v)));
  std::printf("\n");
  std::printf("ccdiag %zu\n", c.last_repaired_vertices());
}

void print_kcore(const velographx::IncrementalKCore& k) {
  const auto& c = k.core();
  std::printf("kcore %zu", c.size());
  for (auto x : c) std::printf(" %u", x);
  std::printf("\n");
  std::printf("kcorediag %zu\n", k.last_repaired_vertices());
}

int run_scenario(std::istream& in) {
  std::unique_ptr<DynamicGraph> g;
  std::unique_ptr<IncrementalBFS> bfs;
  std::unique_ptr<IncrementalPageRank> pr;
  std::unique_ptr<velographx::ConsolidationController> ctl;
  std::unique_ptr<velographx::IncrementalSSSP> sssp;
  std::unique_ptr<velographx::IncrementalKCore> kcore;
  std::unique_ptr<velographx::IncrementalComponents> cc;
  UpdateBatch batch;
  std::string cmd;
  while (in >> cmd) {
    if (cmd == "graph") {
      std::size_t n;
      int directed;
      in >> n >> directed;
      bfs.reset();
      pr.reset();
      sssp.reset();
      kcore.reset();
      cc.reset();
      g = std::make_unique<DynamicGraph>(n, directed != 0);
    } else if (cmd == "bulk") {
      std::size_t m;
      in >> m;
      std::vector<std::pair<VertexId, VertexId>> e(m);
      for (auto& p : e) in >> p.first >> p.second;
      g->bulk_load_edges(e);
    } else if (cmd == "batch") {
      std::size_t k;
      in >> k;
      batch = UpdateBatch{};
      for (std::size_t i = 0; i < k; ++i) {
        VertexId u, v;
        int a;
        in >> u >> v >> a;
        if (a) batch.add(u, v);
        else batch.remove(u, v);
      }
    } else if (cmd == "graph_apply") {
      g->apply(batch);
    } else if (cmd == "bfs_new") {
      VertexId s;
      double f;
      in >> s >> f;
      bfs = std::make_unique<IncrementalBFS>(*g, s, f);
      print_dist(*bfs);
    } else if (cmd == "bfs_apply") {
      bfs->apply(batch);
      print_dist(*bfs);
    } else if (cmd == "bfs_recompute") {
      bfs->recompute();
      print_dist(*bfs);
    } else if (cmd == "pr_new") {
      double d;
      in >> d;
      pr = std::make_unique<IncrementalPageRank>(*g, d);
      print_pr(*pr);
    } else if (cmd == "pr_apply") {
      std::size_t it;
      double tol, frac;
      in >> it >> tol >> frac;
      pr->apply(batch, it, tol, frac);
      print_pr(*pr);
    } else if (cmd == "pr_recompute") {
      std::size_t it;
      double tol;
      in >> it >> tol;
      pr->recompute(it, tol);
      print_pr(*pr);
    } else if (cmd == "pr_validate") {
      std::size_t it;
      double tol, l1, linf;
      in >> it >> tol >> l1 >> linf;
      const auto v = pr->validate_against_full(it, tol, l1, linf);
      std::printf("prvalid %.17g %.17g %zu %d %d\n", v.l1_error, v.linf_error, v.reference_iterations,
                  v.reference_converged ? 1 : 0, v.within_tolerance ? 1 : 0);
    } else if (cmd == "cc_new") {
      cc = std::make_unique<velographx::IncrementalComponents>(*g);
      print_cc(*cc, g->vertex_count());
    } else if (cmd == "cc_apply") {
      cc->apply(batch);
      print_cc(*cc, g->vertex_count());
    } else if (cmd == "kcore_new") {
      kcore = std::make_unique<velographx::IncrementalKCore>(*g);
      print_kcore(*kcore);
    } else if (cmd == "kcore_apply") {
      kcore->apply(batch);
      print_kcore(*kcore);
    } else if (cmd == "sssp_new") {
      VertexId s;
      in >> s;
      sssp = std::make_unique<velographx::IncrementalSSSP>(*g, s);
      print_sssp(*sssp);
    } else if (cmd == "sssp_apply") {
      sssp->apply(batch);
      print_sssp(*sssp);
    } else if (cmd == "cons_policy") {
      std::size_t arcs;
      double latency;
      in >> arcs >> latency;
      const auto p = velographx::scale_aware_consolidation_policy(arcs, latency);
      std::printf("cpolicy %.17g %.17g\n", p.max_storage_growth_ratio, p.max_neighbor_latency_ratio);
    } else if (cmd == "cons_eval") {
      std::size_t cur, canon;
      double cur_lat, canon_lat, max_s, max_l;
      in >> cur >> canon >> cur_lat >> canon_lat >> max_s >> max_l;
      const auto sg = velographx::evaluate_consolidation(cur, canon, cur_lat, canon_lat, {max_s, max_l});
      std::printf("csignal %.17g %.17g %d %d %d\n", sg.storage_growth_ratio, sg.neighbor_latency_ratio,
                  sg.storage_limit_exceeded ? 1 : 0, sg.latency_limit_exceeded ? 1 : 0,
                  sg.should_consolidate ? 1 : 0);
    } else if (cmd == "cons_ctl") {
      velographx::ConsolidationControllerConfig c;
      in >> c.min_epochs_between_consolidations >> c.latency_breach_samples >> c.min_storage_growth_for_latency_trigger;
      ctl = std::make_unique<velographx::ConsolidationController>(c);
    } else if (cmd == "cons_observe") {
      velographx::ConsolidationSignal sg;
      int se, le, sc;
      std::size_t epoch;
      in >> sg.storage_growth_ratio >> sg.neighbor_latency_ratio >> se >> le >> sc >> epoch;
      sg.storage_limit_exceeded = se != 0;
      sg.latency_limit_exceeded = le != 0;
      sg.should_consolidate = sc != 0;
      const bool fire = ctl->observe(sg, epoch);
      std::printf("cobserve %d %zu %zu\n", fire ? 1 : 0, ctl->latency_breach_streak(), ctl->last_consolidation_epoch());
    } else if (cmd == "cons_mark") {
      std::size_t epoch;
      in >> epoch;
      ctl->mark_consolidated(epoch);
    } else if (cmd == "neighbors") {
      VertexId u;
      in >> u;
      const auto nb = g->neighbors(u);
      std::printf("neighbors %zu", nb.size());
      for (auto x : nb) std::printf(" %u", x);
      std::printf("\n");
    } else {
      std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
      return 2;
    }
  }
  return 0;
}

std::vector<std::pair<VertexId, VertexId>> read_edges_bin(const char* path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open edge file");
  f.seekg(0, std::ios::end);
  const auto bytes = static_cast<std::size_t>(f.tellg());
  f.seekg(0);
  std::vector<std::uint32_t> raw(bytes / 4);
  f.read(reinterpret_cast<char*>(raw.data()), static_cast<std::streamsize>(raw.size() * 4));
  std::vector<std::pair<VertexId, VertexId>> e(raw.size() / 2);
  for (std::size_t i = 0; i < e.size(); ++i) e[i] = {raw[2 * i], raw[2 * i + 1]};
  return e;
}

using Clock = std::chrono::steady_clock;

// bench-bfs <edges.bin> <n> <directed> <time_budget_s> <root...>
// Builds the graph, then runs IncrementalBFS::recompute from each root in
// turn (one thread), stopping once the time budget is spent.  Prints one
// line per root: "root <r> <seconds> <reached> <arcs_in_component>".
int bench_bfs(int argc, char** argv) {
  const auto edges = read_edges_bin(argv[2]);
  const std::size_t n = std::stoull(argv[3]);
  const bool directed = std::stoi(argv[4]) != 0;
  const double budget = std::stod(argv[5]);
  DynamicGraph g(n, directed);
  g.bulk_load_edges(edges);
  const auto t_all = Clock::now();
  for (int i = 6; i < argc; ++i) {
    const VertexId r = static_cast<VertexId>(std::stoul(argv[i]));
    const auto t0 = Clock::now();
    IncrementalBFS bfs(g, r, 2.0);
    const auto t1 = Clock::now();
    std::uint64_t arcs = 0;
    const auto& d = bfs.distances();
    for (std::size_t u = 0; u < d.size(); ++u)
      if (d[u] != IncrementalBFS::unreachable) g.for_each_neighbor(static_cast<VertexId>(u), [&](VertexId) { ++arcs; });
    std::printf("root %u %.9f %zu %llu\n", r, std::chrono::duration<double>(t1 - t0).count(),
                bfs.reachable_count(), static_cast<unsigned long long>(arcs));
    std::fflush(stdout);
    if (std::chrono::duration<double>(Clock::now() - t_all).count() > budget) break;
  }
  return 0;
}

// bench-bfs-mt <edges.bin> <threads> <passes> <root...>
// The root-parallel BFS of the reference's thread-scaling benchmark
// (benchmarks/thread_scaling_bfs.cpp:78-108): an undirected CsrGraph, then
// `threads` workers that take the next query off a shared counter and run
// bfs_distances from its root; the queries are the roots repeated `passes`
// times.  Prints "mt <threads> <queries> <seconds> <sum of reached>".
int bench_bfs_mt(int argc, char** argv) {
  const auto edges = read_edges_bin(argv[2]);
  const unsigned threads = static_cast<unsigned>(std::stoul(argv[3]));
  const std::size_t passes = std::stoull(argv[4]);
  std::vector<VertexId> roots;
  for (int i = 5; i < argc; ++i) roots.push_back(static_cast<VertexId>(std::stoul(argv[i])));
  const velographx::CsrGraph graph(edges, false);
  std::vector<VertexId> queries;
  for (std::size_t p = 0; p < passes; ++p) queries.insert(queries.end(), roots.begin(), roots.end());
  std::vector<std::uint64_t> reached(queries.size(), 0);
  std::atomic<std::size_t> next{0};
  const auto t0 = Clock::now();
  std::vector<std::thread> workers;
  for (unsigned t = 0; t < threads; ++t) {
    workers.emplace_back([&] {
      for (;;) {
        const auto i = next.fetch_add(1, std::memory_order_relaxed);
        if (i >= queries.size()) break;
        const auto d = velographx::bfs_distances(graph, queries[i]);
        std::uint64_t c = 0;
        for (auto x : d) c += x != velographx::kUnreachable ? 1 : 0;
        reached[i] = c;
      }
    });
  }
  for (auto& w : workers) w.join();
  const double sec = std::chrono::duration<double>(Clock::now() - t0).count();
  std::uint64_t total = 0;
  for (auto c : reached) total += c;
  std::printf("mt %u %zu %.9f %llu\n", threads, queries.size(), sec, static_cast<unsigned long long>(total));
  return 0;
}

// bench-snapshot <edges.bin> <n> <directed> <repeats>
// Builds the graph, then times consolidate_to_csr_snapshot (one thread)
// <repeats> times, and the neighbour probe of the steady-state storage
// benchmark (4096 rows, steady_state_storage.cpp:61-72's row sequence).
// Prints "snapshot <seconds> <arcs> <source_bytes> <consolidated_bytes>"
// per repeat and "probe <ns_per_probe> <checksum>".
int bench_snapshot(char** argv) {
  const auto edges = read_edges_bin(argv[2]);
  const std::size_t n = std::stoull(argv[3]);
  const bool directed = std::stoi(argv[4]) != 0;
  const int repeats = std::stoi(argv[5]);
  DynamicGraph g(n, directed);
  g.bulk_load_edges(edges);
  for (int r = 0; r < repeats; ++r) {
    const auto t0 = Clock::now();
    auto snap = velographx::consolidate_to_csr_snapshot(g);
    const auto t1 = Clock::now();
    std::printf("snapshot %.9f %zu %zu %zu\n", std::chrono::duration<double>(t1 - t0).count(), snap.directed_edges,
                snap.source_storage_bytes, snap.consolidated_storage_bytes);
    std::fflush(stdout);
  }
  constexpr std::size_t kProbes = 4096;
  std::uint64_t sum = 0;
  const auto t0 = Clock::now();
  for (std::size_t i = 0; i < kProbes; ++i) {
    const auto u = static_cast<VertexId>((i * 11400714819323198485ULL) % g.vertex_count());
    const auto row = g.neighbors(u);
    sum += row.size();
    for (const auto v : row) sum += v;
  }
  const double ns = std::chrono::duration<double, std::nano>(Clock::now() - t0).count() / kProbes;
  std::printf("probe %.3f %llu\n", ns, static_cast<unsigned long long>(sum));
  return 0;
}

// bench-static <edges.bin> <directed> <root> <pr_iterations>
// Times the CsrGraph construction, bfs_distances from <root> and
// pagerank(graph, 0.85, <pr_iterations>, 0.0) (one thread).  Prints
// "static <build_s> <bfs_s> <pagerank_s> <vertices> <entries>".
int bench_static(char** argv) {
  const auto edges = read_edges_bin(argv[2]);
  const bool directed = std::stoi(argv[3]) != 0;
  const VertexId root = static_cast<VertexId>(std::stoul(argv[4]));
  const std::size_t iters = std::stoull(argv[5]);
  const auto t0 = Clock::now();
  velographx::CsrGraph g(edges, directed);
  const auto t1 = Clock::now();
  const auto d = velographx::bfs_distances(g, root);
  const auto t2 = Clock::now();
  const auto pr = velographx::pagerank(g, 0.85, iters, 0.0);
  const auto t3 = Clock::now();
  std::printf("static %.9f %.9f %.9f %zu %zu\n", std::chrono::duration<double>(t1 - t0).count(),
              std::chrono::duration<double>(t2 - t1).count(), std::chrono::duration<double>(t3 - t2).count(),
              g.vertex_count(), g.edge_entry_count());
  return d.empty() || pr.empty() ? 1 : 0;
}

// Order-free digest of a per-vertex array: the sum mod 2^64 of splitmix64's
// output function (Steele, Lea and Flood, OOPSLA 2014) applied to
// (v << 32 | value[v]), equal for two arrays exactly when (with
// overwhelming probability) every entry is.
std::uint64_t vertex_digest(const std::vector<std::uint32_t>& value) {
  std::uint64_t h = 0;
  for (std::size_t v = 0; v < value.size(); ++v) {
    std::uint64_t x = (static_cast<std::uint64_t>(v) << 32) | value[v];
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    h += x ^ (x >> 31);
  }
  return h;
}

// bench-kcore <edges.bin> <n>
// Times IncrementalKCore(graph) on the undirected graph (its constructor
// peels every vertex, one thread).  Prints "kcore <seconds> <max_core>
// <digest>" with the digest of the core numbers.
int bench_kcore(char** argv) {
  const auto edges = read_edges_bin(argv[2]);
  const std::size_t n = std::stoull(argv[3]);
  DynamicGraph g(n, false);
  g.bulk_load_edges(edges);
  const auto t0 = Clock::now();
  velographx::IncrementalKCore kc(g);
  const auto t1 = Clock::now();
  std::uint32_t mx = 0;
  for (auto c : kc.core()) mx = std::max(mx, c);
  std::printf("kcore %.9f %u %llu\n", std::chrono::duration<double>(t1 - t0).count(), mx,
              static_cast<unsigned long long>(vertex_digest(kc.core())));
  return 0;
}

// bench-cc <edges.bin> <n>
// Times IncrementalComponents(graph) on the undirected graph (its
// constructor unions every arc, one thread).  Prints "cc <seconds> <count>
// <digest>" with the number of components and the digest of the partition
// labelled by each component's smallest vertex.
int bench_cc(char** argv) {
  const auto edges = read_edges_bin(argv[2]);
  const std::size_t n = std::stoull(argv[3]);
  DynamicGraph g(n, false);
  g.bulk_load_edges(edges);
  const auto t0 = Clock::now();
  velographx::IncrementalComponents cc(g);
  const auto t1 = Clock::now();
  std::size_t roots = 0;
  const std::size_t nv = g.vertex_count();
  // the partition, labelled by each component's smallest vertex
  std::vector<std::uint32_t> root(nv), smallest(nv, ~0u), canon(nv);
  for (std::size_t v = 0; v < nv; ++v) {
    root[v] = cc.component(static_cast<VertexId>(v));
    if (root[v] == v) ++roots;
    smallest[root[v]] = std::min<std::uint32_t>(smallest[root[v]], static_cast<std::uint32_t>(v));
  }
  for (std::size_t v = 0; v < nv; ++v) canon[v] = smallest[root[v]];
  std::printf("cc %.9f %zu %llu\n", std::chrono::duration<double>(t1 - t0).count(), roots,
              static_cast<unsigned long long>(vertex_digest(canon)));
  return 0;
}

// bench-pr <edges.bin> <n> <directed> <iterations>
// Times IncrementalPageRank(graph, 0.85) followed by recompute(iterations,
// 0.0): exactly <iterations> pull sweeps.  Prints
// "pr <seconds> <iterations>".
int bench_pr(char** argv) {
  const auto edges = read_edges_bin(argv[2]);
  const std::size_t n = std::stoull(argv[3]);
  const bool directed = std::stoi(argv[4]) != 0;
  const std::size_t iters = std::stoull(argv[5]);
  DynamicGraph g(n, directed);
  g.bulk_load_edges(edges);
  IncrementalPageRank pr(g, 0.85);
  const auto t0 = Clock::now();
  pr.recompute(iters, 0.0);
  const auto t1 = Clock::now();
  std::printf("pr %.9f %zu\n", std::chrono::duration<double>(t1 - t0).count(), pr.last_repair_iterations());
  return 0;
}

// bench-repair <edges.bin> <n> <directed> <root> <batches.bin> <budget_s>
// batches.bin: repeated records {u32 k; k x (u32 src, u32 dst, u32 add)}.
// Applies each batch through IncrementalBFS::apply (one thread) until the
// budget is spent.  Prints "batch <seconds> <k> <full_recompute>" per batch.
int bench_repair(char** argv) {
  const auto edges = read_edges_bin(argv[2]);
  const std::size_t n = std::stoull(argv[3]);
  const bool directed = std::stoi(argv[4]) != 0;
  const VertexId root = static_cast<VertexId>(std::stoul(argv[5]));
  const double budget = std::stod(argv[7]);
  std::ifstream f(argv[6], std::ios::binary);
  if (!f) throw std::runtime_error("cannot open batch file");
  DynamicGraph g(n, directed);
  g.bulk_load_edges(edges);
  IncrementalBFS bfs(g, root);
  const auto t_all = Clock::now();
  std::uint32_t k = 0;
  while (f.read(reinterpret_cast<char*>(&k), 4)) {
    std::vector<std::uint32_t> raw(static_cast<std::size_t>(k) * 3);
    f.read(reinterpret_cast<char*>(raw.data()), static_cast<std::streamsize>(raw.size() * 4));
    UpdateBatch b;
    for (std::uint32_t i = 0; i < k; ++i) {
      if (raw[3 * i + 2]) b.add(raw[3 * i], raw[3 * i + 1]);
      else b.remove(raw[3 * i], raw[3 * i + 1]);
    }
    const auto t0 = Clock::now();
    bfs.apply(b);
    const auto t1 = Clock::now();
    std::printf("batch %.9f %u %d\n", std::chrono::duration<double>(t1 - t0).count(), k,
                bfs.last_used_full_recompute() ? 1 : 0);
    std::fflush(stdout);
    if (std::chrono::duration<double>(Clock::now() - t_all).count() > budget) break;
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 2 && std::strcmp(argv[1], "bench-bfs") == 0 && argc >= 7) return bench_bfs(argc, argv);
  if (argc >= 6 && std::strcmp(argv[1], "bench-bfs-mt") == 0) return bench_bfs_mt(argc, argv);
  if (argc == 6 && std::strcmp(argv[1], "bench-pr") == 0) return bench_pr(argv);
  if (argc == 6 && std::strcmp(argv[1], "bench-snapshot") == 0) return bench_snapshot(argv);
  if (argc == 6 && std::strcmp(argv[1], "bench-static") == 0) return bench_static(argv);
  if (argc == 4 && std::strcmp(argv[1], "bench-kcore") == 0) return bench_kcore(argv);
  if (argc == 4 && std::strcmp(argv[1], "bench-cc") == 0) return bench_cc(argv);
  if (argc == 8 && std::strcmp(argv[1], "bench-repair") == 0) return bench_repair(argv);
  if (argc >= 3 && std::strcmp(argv[1], "scenario") == 0) {
    std::ifstream f(argv[2]);
    return run_scenario(f);
  }
  return run_scenario(std::cin);
}
