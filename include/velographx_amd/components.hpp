// Incremental connected components -- the interface of
// include/velographx/incremental/connected_components.hpp
// (BasicIncrementalComponents<DynamicGraph>) over the device engine
// (vgx_components_*).  component(v) is the smallest vertex id of v's
// component: the reference's partition, with an order-free representative
// instead of its union-find root.
#pragma once

#include <cstdint>
#include <vector>

#include "velographx_amd/velographx.hpp"

namespace velographx {

class IncrementalComponents {
 public:
  explicit IncrementalComponents(DynamicGraph& g) { vgx_check(vgx_components_create(g.handle(), &c_)); }
  ~IncrementalComponents() {
    if (c_) vgx_components_destroy(c_);
  }
  IncrementalComponents(const IncrementalComponents&) = delete;
  IncrementalComponents& operator=(const IncrementalComponents&) = delete;

  // connected_components.hpp:27-30
  [[nodiscard]] std::uint32_t component(VertexId v) const {
    std::uint32_t label = 0;
    vgx_check(vgx_components_component(c_, v, &label));
    return label;
  }
  [[nodiscard]] std::vector<std::uint32_t> labels() const {
    std::uint64_t n = 0;
    vgx_check(vgx_components_labels(c_, nullptr, 0, &n));
    std::vector<std::uint32_t> l(n);
    if (n) vgx_check(vgx_components_labels(c_, l.data(), n, &n));
    return l;
  }
  [[nodiscard]] std::size_t last_repaired_vertices() const {
    std::uint64_t r = 0;
    vgx_check(vgx_components_stats(c_, &r, nullptr));
    return r;
  }

  void apply(const UpdateBatch& batch) {
    auto c = batch.columns();
    vgx_check(vgx_components_apply(c_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
  }
  void apply_arrays(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k) {
    vgx_check(vgx_components_apply(c_, src, dst, add, k));
  }
  void recompute() { vgx_check(vgx_components_recompute(c_)); }

  [[nodiscard]] vgx_components* handle() const noexcept { return c_; }

 private:
  vgx_components* c_ = nullptr;
};

}  // namespace velographx
