// Incremental connected components -- the interface of
// include/velographx/incremental/connected_components.hpp:16-157
// (BasicIncrementalComponents<Graph>; IncrementalComponents is the
// DynamicGraph instance, :157) over the device engine (vgx_components_*),
// on the device graph GraphBinding<Graph> supplies.  component(v) is the
// smallest vertex id of v's component: the reference's partition, with an
// order-free representative instead of its union-find root.
#pragma once

#include <concepts>
#include <cstdint>
#include <vector>

#include "velographx_amd/dynamic_graph.hpp"
#include "velographx_amd/graph_binding.hpp"

namespace velographx {

template <class Graph>
class BasicIncrementalComponents {
 public:
  explicit BasicIncrementalComponents(Graph& g) : bind_(g) {
    vgx_check(vgx_components_create(bind_.device().handle(), &c_));
  }
  ~BasicIncrementalComponents() {
    if (c_) vgx_components_destroy(c_);
  }
  BasicIncrementalComponents(const BasicIncrementalComponents&) = delete;
  BasicIncrementalComponents& operator=(const BasicIncrementalComponents&) = delete;

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

  // connected_components.hpp:35-119
  void apply(const UpdateBatch& batch) {
    auto c = bind_.forward(batch).columns();
    vgx_check(vgx_components_apply(c_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
  }
  void apply_arrays(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k)
    requires std::same_as<Graph, DynamicGraph>
  {
    vgx_check(vgx_components_apply(c_, src, dst, add, k));
  }
  void recompute() {
    bind_.refresh();
    vgx_check(vgx_components_recompute(c_));
  }

  [[nodiscard]] vgx_components* handle() const noexcept { return c_; }

 private:
  detail::GraphBinding<Graph> bind_;
  vgx_components* c_ = nullptr;
};

// connected_components.hpp:157
using IncrementalComponents = BasicIncrementalComponents<DynamicGraph>;

}  // namespace velographx
