// selector.hpp -- C++20 facade over the adaptive repair-vs-recompute
// selector of vgx.h (the reference's adaptive publication policy,
// benchmarks/publication_policy_bfs.cpp:162-233, :297-330).
#pragma once

#include <cstdint>
#include <string>

#include "velographx_amd/bfs.hpp"

namespace velographx {

using SelectorTrace = vgx_selector_trace;

inline std::string policy_reason_name(int reason) { return vgx_policy_reason_name(reason); }

// The decision model on recorded features (no device work).
class AdaptivePolicy {
 public:
  AdaptivePolicy(std::uint64_t vertices, double initial_full_us, double initial_reachable_fraction) {
    vgx_check(vgx_policy_create(vertices, initial_full_us, initial_reachable_fraction, &p_));
  }
  ~AdaptivePolicy() {
    if (p_) vgx_policy_destroy(p_);
  }
  AdaptivePolicy(const AdaptivePolicy&) = delete;
  AdaptivePolicy& operator=(const AdaptivePolicy&) = delete;

  // Fills predicted_*, chose_full and reason of t from its features.
  void decide(SelectorTrace& t) const { vgx_check(vgx_policy_decide(p_, &t)); }
  // t carries the decision and its outcome (execution_us, fallback, affected).
  void observe(const SelectorTrace& t) { vgx_check(vgx_policy_observe(p_, &t)); }

 private:
  vgx_policy* p_ = nullptr;
};

// The selector bound to an IncrementalBFS: features on the GPU, either arm
// executed and timed answer-ready.
class AdaptiveSelector {
 public:
  explicit AdaptiveSelector(IncrementalBFS& bfs) { vgx_check(vgx_selector_create(bfs.handle(), &s_)); }
  ~AdaptiveSelector() {
    if (s_) vgx_selector_destroy(s_);
  }
  AdaptiveSelector(const AdaptiveSelector&) = delete;
  AdaptiveSelector& operator=(const AdaptiveSelector&) = delete;

  SelectorTrace apply(const UpdateBatch& batch) {
    auto c = batch.columns();
    return apply(c.src.data(), c.dst.data(), c.add.data(), c.src.size());
  }
  SelectorTrace apply(const std::uint32_t* src, const std::uint32_t* dst, const std::uint8_t* add, std::size_t k) {
    SelectorTrace t{};
    vgx_check(vgx_selector_apply(s_, src, dst, add, k, &t));
    last_ = t;
    return t;
  }
  [[nodiscard]] const SelectorTrace& last_trace() const noexcept { return last_; }

 private:
  vgx_selector* s_ = nullptr;
  SelectorTrace last_{};
};

}  // namespace velographx
