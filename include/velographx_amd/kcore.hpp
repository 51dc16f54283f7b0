// Incremental k-core -- the interface of include/velographx/incremental/kcore.hpp
// (BasicIncrementalKCore<DynamicGraph>) over the device engine's k-core
// (vgx_kcore_*): core numbers by level-synchronous peeling on the live rows.
#pragma once

#include <cstdint>
#include <vector>

#include "velographx_amd/velographx.hpp"

namespace velographx {

class IncrementalKCore {
 public:
  explicit IncrementalKCore(DynamicGraph& g) { vgx_check(vgx_kcore_create(g.handle(), &k_)); }
  ~IncrementalKCore() {
    if (k_) vgx_kcore_destroy(k_);
  }
  IncrementalKCore(const IncrementalKCore&) = delete;
  IncrementalKCore& operator=(const IncrementalKCore&) = delete;

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

  void apply(const UpdateBatch& batch) {
    auto c = batch.columns();
    vgx_check(vgx_kcore_apply(k_, c.src.data(), c.dst.data(), c.add.data(), c.src.size()));
  }
  void apply_arrays(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k) {
    vgx_check(vgx_kcore_apply(k_, src, dst, add, k));
  }
  void recompute() { vgx_check(vgx_kcore_recompute(k_)); }

  [[nodiscard]] vgx_kcore* handle() const noexcept { return k_; }

 private:
  vgx_kcore* k_ = nullptr;
};

}  // namespace velographx
