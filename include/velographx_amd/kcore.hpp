// Incremental k-core -- the interface of
// include/velographx/incremental/kcore.hpp:17-127 (BasicIncrementalKCore<Graph>;
// IncrementalKCore is the DynamicGraph instance, :127) over the device
// engine's k-core (vgx_kcore_*): core numbers by level-synchronous peeling
// on the live rows of the device graph GraphBinding<Graph> supplies.
#pragma once

#include <concepts>
#include <cstdint>
#include <vector>

#include "velographx_amd/dynamic_graph.hpp"
#include "velographx_amd/graph_binding.hpp"

namespace velographx {

template <class Graph>
class BasicIncrementalKCore {
 public:
  explicit BasicIncrementalKCore(Graph& g) : bind_(g) { vgx_check(vgx_kcore_create(bind_.device().handle(), &k_)); }
  ~BasicIncrementalKCore() {
    if (k_) vgx_kcore_destroy(k_);
  }
  BasicIncrementalKCore(const BasicIncrementalKCore&) = delete;
  BasicIncrementalKCore& operator=(const BasicIncrementalKCore&) = delete;

  // kcore.hpp:28 (a copy: the values live in HBM)
  [[nodiscard]] std::vector<std::uint32_t> core() const {
    std::uint64_t n = 0;
    vgx_check(vgx_kcore_core(k_, nullptr, 0, &n));
    std::vector<std::uint32_t> c(n);
    if (n) vgx_check(vgx_kcore_core(k_, c.data(), n, &n));
    return c;
  }
  [[nodiscard]] std::size_t last_repaired_vertices() const {
    std::uint64_t r = 0;
    vgx_check(vgx_kcore_stats(k_, &r, nullptr, nullptr));
    return r;
  }

  // kcore.hpp:31-64
  void apply(const UpdateBatch& batch) {
    auto c = bind_.forward(batch).columns();
    vgx_check(vgx_kcore_apply(k_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
  }
  void apply_arrays(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k)
    requires std::same_as<Graph, DynamicGraph>
  {
    vgx_check(vgx_kcore_apply(k_, src, dst, add, k));
  }
  void recompute() {
    bind_.refresh();
    vgx_check(vgx_kcore_recompute(k_));
  }

  [[nodiscard]] vgx_kcore* handle() const noexcept { return k_; }

 private:
  detail::GraphBinding<Graph> bind_;
  vgx_kcore* k_ = nullptr;
};

// kcore.hpp:127
using IncrementalKCore = BasicIncrementalKCore<DynamicGraph>;

}  // namespace velographx
