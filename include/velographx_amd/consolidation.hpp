// Consolidation policy, steady-state controller and snapshot cutover over the
// device-resident graph -- the interface of storage/consolidation.hpp:10-163.
//
// The signal's inputs come from the device: storage bytes from the graph's
// arena and row tables (vgx_graph_storage_info), neighbour latency from the
// on-device probe (vgx_graph_probe_neighbors).  The policy and the
// controller are a few comparisons per epoch and run on the host; the
// snapshot is built on the device (vgx_graph_snapshot) without touching the
// source graph, so a caller can validate it before cutting over.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>

#include "velographx_amd/dynamic_graph.hpp"

namespace velographx {

// consolidation.hpp:11-16
struct ConsolidationSnapshot {
  DynamicGraph graph;
  std::size_t source_storage_bytes{0};
  std::size_t consolidated_storage_bytes{0};
  std::size_t directed_edges{0};
};

// consolidation.hpp:18-21: ratios at which storage growth or neighbour
// latency (current over canonical) call for consolidation.
struct ConsolidationPolicy {
  double max_storage_growth_ratio{1.25};
  double max_neighbor_latency_ratio{1.25};
};

// consolidation.hpp:23-33: graphs of at least 100,000,000 directed arcs get
// a storage bound of 1.5 instead of 1.25; the latency bound is the caller's.
inline ConsolidationPolicy scale_aware_consolidation_policy(std::size_t directed_edges,
                                                            double latency_ratio = 1.25) noexcept {
  const bool large = directed_edges >= std::size_t{100'000'000};
  return ConsolidationPolicy{large ? 1.50 : 1.25, latency_ratio};
}

// consolidation.hpp:35-41
struct ConsolidationSignal {
  double storage_growth_ratio{1.0};
  double neighbor_latency_ratio{1.0};
  bool storage_limit_exceeded{false};
  bool latency_limit_exceeded{false};
  bool should_consolidate{false};
};

// consolidation.hpp:43-61: a zero canonical size or a non-positive
// canonical latency reads as ratio 1; a limit is exceeded at ratio >= bound.
inline ConsolidationSignal evaluate_consolidation(std::size_t current_storage_bytes,
                                                  std::size_t canonical_storage_bytes,
                                                  double current_neighbor_latency,
                                                  double canonical_neighbor_latency,
                                                  ConsolidationPolicy policy = {}) noexcept {
  ConsolidationSignal s;
  if (canonical_storage_bytes != 0)
    s.storage_growth_ratio =
        static_cast<double>(current_storage_bytes) / static_cast<double>(canonical_storage_bytes);
  if (canonical_neighbor_latency > 0.0) s.neighbor_latency_ratio = current_neighbor_latency / canonical_neighbor_latency;
  s.storage_limit_exceeded = s.storage_growth_ratio >= policy.max_storage_growth_ratio;
  s.latency_limit_exceeded = s.neighbor_latency_ratio >= policy.max_neighbor_latency_ratio;
  s.should_consolidate = s.storage_limit_exceeded || s.latency_limit_exceeded;
  return s;
}

// consolidation.hpp:67-71
struct ConsolidationControllerConfig {
  std::size_t min_epochs_between_consolidations{10};
  std::size_t latency_breach_samples{5};
  double min_storage_growth_for_latency_trigger{1.10};
};

// consolidation.hpp:73-119.  A storage breach triggers at once.  A latency
// breach counts only with storage growth of at least
// min_storage_growth_for_latency_trigger, only after the cooldown since the
// last consolidation (epochs inside the cooldown clear the count), and
// triggers on the latency_breach_samples-th consecutive counted epoch.
class ConsolidationController {
 public:
  explicit ConsolidationController(ConsolidationControllerConfig config = {}) : config_(config) {}

  bool observe(const ConsolidationSignal& signal, std::size_t epoch) noexcept {
    if (signal.storage_limit_exceeded) {
      streak_ = 0;
      return true;
    }
    if (last_ != 0 && epoch < last_ + config_.min_epochs_between_consolidations) {
      streak_ = 0;
      return false;
    }
    const bool counts = signal.latency_limit_exceeded &&
                        signal.storage_growth_ratio >= config_.min_storage_growth_for_latency_trigger;
    streak_ = counts ? streak_ + 1 : 0;
    return counts && streak_ >= config_.latency_breach_samples;
  }

  void mark_consolidated(std::size_t epoch) noexcept {
    last_ = epoch;
    streak_ = 0;
  }

  [[nodiscard]] std::size_t latency_breach_streak() const noexcept { return streak_; }
  [[nodiscard]] std::size_t last_consolidation_epoch() const noexcept { return last_; }
  [[nodiscard]] const ConsolidationControllerConfig& config() const noexcept { return config_; }

 private:
  ConsolidationControllerConfig config_{};
  std::size_t streak_{0};
  std::size_t last_{0};
};

// consolidation.hpp:132-163: the current logical graph rebuilt, on the
// device, into compacted rows in a new graph; the source is unchanged.
inline ConsolidationSnapshot consolidate_to_csr_snapshot(const DynamicGraph& source) {
  vgx_graph* h = nullptr;
  vgx_check(vgx_graph_snapshot(source.handle(), &h));
  DynamicGraph g = DynamicGraph::adopt(h);
  const std::size_t consolidated = g.storage_bytes();
  return ConsolidationSnapshot{std::move(g), source.storage_bytes(), consolidated, source.edge_count_directed()};
}

// benchmarks/steady_state_storage.cpp:61-86: median nanoseconds per probe of
// `count` out-rows over `repeats` device launches, and the rows' digest.
struct NeighborProbe {
  double ns_per_probe{0.0};
  std::uint64_t digest{0};
};

inline NeighborProbe probe_neighbors(const DynamicGraph& g, std::size_t count = 4096, std::uint32_t repeats = 5) {
  NeighborProbe p;
  vgx_check(vgx_graph_probe_neighbors(g.handle(), count, repeats, &p.ns_per_probe, &p.digest));
  return p;
}

}  // namespace velographx
