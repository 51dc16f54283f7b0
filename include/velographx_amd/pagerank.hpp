// BasicIncrementalPageRank<Graph> -- the interface of
// include/velographx/incremental/pagerank.hpp:29-336 (IncrementalPageRank
// is BasicIncrementalPageRank<DynamicGraph>, pagerank.hpp:334) over the
// gfx950 C ABI.  precision 64 keeps the reference's FP64 arithmetic and
// stopping rules; precision 32 runs the same iteration in float.  The
// sweeps run on the device graph GraphBinding<Graph> supplies (bfs.hpp
// says which).
#pragma once

#include <concepts>
#include <cstdint>
#include <vector>

#include "velographx_amd/dynamic_graph.hpp"
#include "velographx_amd/graph_binding.hpp"
#include "vgx.h"

namespace velographx {

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

template <class Graph>
class BasicIncrementalPageRank {
 public:
  // pagerank.hpp:31-37
  explicit BasicIncrementalPageRank(Graph& g, double damping = 0.85, int precision = 64)
      : bind_(g), precision_(precision) {
    vgx_check(vgx_pagerank_create(bind_.device().handle(), damping, precision, &p_));
  }
  ~BasicIncrementalPageRank() {
    if (p_) vgx_pagerank_destroy(p_);
  }
  BasicIncrementalPageRank(const BasicIncrementalPageRank&) = delete;
  BasicIncrementalPageRank& operator=(const BasicIncrementalPageRank&) = delete;

  // pagerank.hpp:39 (a copy: the values live in HBM)

  [[nodiscard]] std::vector<double> values() const {
    std::uint64_t n = 0;
    vgx_check(vgx_pagerank_values(p_, nullptr, 0, &n));
    std::vector<double> v(n);
    if (n) vgx_check(vgx_pagerank_values(p_, v.data(), n, &n));
    return v;
  }
  [[nodiscard]] std::size_t last_repaired_vertices() const { return stats().last_repaired_vertices; }
  [[nodiscard]] std::size_t last_repair_iterations() const { return stats().last_repair_iterations; }
  [[nodiscard]] double last_residual_l1() const { return stats().last_residual_l1; }
  [[nodiscard]] double last_residual_linf() const { return stats().last_residual_linf; }
  [[nodiscard]] bool last_full_recompute_converged() const { return stats().last_full_recompute_converged != 0; }
  [[nodiscard]] bool last_used_full_recompute() const { return stats().last_used_full_recompute != 0; }
  [[nodiscard]] int precision() const noexcept { return precision_; }

  // pagerank.hpp:47-191
  void apply(const UpdateBatch& batch, std::size_t local_iterations = 24, double tol = 1e-9,
             double full_fallback_fraction = 0.60) {
    auto c = bind_.forward(batch).columns();
    vgx_check(vgx_pagerank_apply(p_, c.src.data(), c.dst.data(), c.add.data(), c.src.size(), local_iterations, tol,
                                 full_fallback_fraction));
  }

  void recompute(std::size_t max_iterations = 200, double tol = 1e-12) {
    bind_.refresh();
    vgx_check(vgx_pagerank_recompute(p_, max_iterations, tol));
  }

  // addition: full solve to an L1 residual (src/algorithms.cpp:56-81 rule)
  void solve_l1(std::size_t max_iterations = 200, double l1_tol = 1e-7, bool warm_start = false) {
    vgx_check(vgx_pagerank_solve_l1(p_, max_iterations, l1_tol, warm_start ? 1 : 0));
  }

  // addition: localized residual-push repair with a full-sweep fallback
  vgx_pagerank_push_stats apply_push(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add,
                                     std::size_t k, double l1_tol = 1e-7, double threshold = 0.0,
                                     double max_push_fraction = 0.05, std::size_t max_sweeps = 200)
    requires std::same_as<Graph, DynamicGraph>
  {
    vgx_pagerank_push_stats st{};
    vgx_check(vgx_pagerank_apply_push(p_, src, dst, add, k, l1_tol, threshold, max_push_fraction, max_sweeps, &st));
    return st;
  }
  vgx_pagerank_push_stats apply_push(const UpdateBatch& batch, double l1_tol = 1e-7, double threshold = 0.0,
                                     double max_push_fraction = 0.05, std::size_t max_sweeps = 200) {
    auto c = bind_.forward(batch).columns();
    vgx_pagerank_push_stats st{};
    vgx_check(vgx_pagerank_apply_push(p_, c.src.data(), c.dst.data(), c.add.data(), c.src.size(), l1_tol, threshold,
                                      max_push_fraction, max_sweeps, &st));
    return st;
  }

  [[nodiscard]] PageRankValidation validate_against_full(std::size_t reference_max_iterations = 500,
                                                         double reference_tol = 1e-12, double l1_tolerance = 1e-6,
                                                         double linf_tolerance = 1e-7) const {
    double e[4] = {0, 0, 0, 0};
    std::uint64_t it = 0;
    int conv = 0;
    vgx_check(vgx_pagerank_validate(p_, reference_max_iterations, reference_tol, e, &it, &conv));
    PageRankValidation v;
    v.l1_error = e[0];
    v.linf_error = e[1];
    v.reference_residual_l1 = e[2];
    v.reference_residual_linf = e[3];
    v.reference_iterations = it;
    v.reference_converged = conv != 0;
    const auto s = stats();
    v.local_residual_l1 = s.last_residual_l1;
    v.local_residual_linf = s.last_residual_linf;
    v.within_tolerance = v.reference_converged && v.l1_error <= l1_tolerance && v.linf_error <= linf_tolerance;
    return v;
  }

  [[nodiscard]] PageRankValidation apply_validated(const UpdateBatch& batch, std::size_t local_iterations = 64,
                                                   double local_tol = 1e-10, double full_fallback_fraction = 0.95,
                                                   std::size_t reference_max_iterations = 500,
                                                   double reference_tol = 1e-12, double l1_tolerance = 1e-6,
                                                   double linf_tolerance = 1e-7) {
    apply(batch, local_iterations, local_tol, full_fallback_fraction);
    auto v = validate_against_full(reference_max_iterations, reference_tol, l1_tolerance, linf_tolerance);
    if (!v.within_tolerance) {
      vgx_check(vgx_pagerank_install_reference(p_, reference_max_iterations, reference_tol));
      v.fallback_applied = true;
    }
    return v;
  }

  [[nodiscard]] vgx_pagerank_stats stats() const {
    vgx_pagerank_stats s{};
    vgx_check(vgx_pagerank_stats_get(p_, &s));
    return s;
  }
  [[nodiscard]] vgx_pagerank* handle() const noexcept { return p_; }

 private:
  detail::GraphBinding<Graph> bind_;
  vgx_pagerank* p_ = nullptr;
  int precision_ = 64;
};

// pagerank.hpp:334
using IncrementalPageRank = BasicIncrementalPageRank<DynamicGraph>;

}  // namespace velographx
