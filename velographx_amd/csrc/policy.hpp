// Adaptive repair-vs-recompute policy for incremental BFS (host code).
//
// Restates the reference's "adaptive" publication policy
// (benchmarks/publication_policy_bfs.cpp:162-233 for the decision,
// :297-330 for the model update; constants from
// benchmarks/adaptive_policy_bfs.cpp:21-31).  The policy sees per-batch
// features and the measured cost of the arm that ran; it keeps an
// exponential moving average of each arm's answer-ready cost and chooses
// the full recompute when the incremental arm is confidently the slower one.
//
// Large graphs (>= 200,000 vertices): the incremental cost is predicted as
// EMA_inc * clamp(sqrt(f / f_last_inc), 0.5, 2.5) * (1 + 3 a_prev), with f
// the update fraction and a_prev the affected fraction of the last
// incremental batch; uncertainty bands are EMAs of the relative prediction
// error clamped to [0.05, 0.35] (0.35 / 0.20 before the first sample).
// Small graphs use reach guards and a learned margin of 1.25 on a model no
// older than 4 batches.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace vgx {

enum PolicyReason : int {
  kLargePreflightFull = 0,
  kLargeShallowColdStart = 1,
  kLargeInitialIncrementalProbe = 2,
  kLargeConfidentFull = 3,
  kLargeOverlapIncremental = 4,
  kScalePreflightFull = 5,
  kScaleVerySparseReach = 6,
  kScaleSparseReachGuard = 7,
  kScaleFreshCostModel = 8,
  kScaleConfidenceIncremental = 9,
  kScaleInsufficientEvidence = 10,
  kPolicyReasonCount = 11,
};

inline const char* policy_reason_name(int r) {
  static const char* names[kPolicyReasonCount] = {
      "large_preflight_full",           "large_shallow_cold_start",
      "large_initial_incremental_probe", "large_uncertainty_confident_full",
      "large_uncertainty_overlap_incremental", "scale_preflight_full",
      "scale_very_sparse_reach",        "scale_sparse_reach_guard",
      "scale_fresh_cost_model",         "scale_confidence_incremental",
      "scale_insufficient_evidence_incremental"};
  return r >= 0 && r < kPolicyReasonCount ? names[r] : "unknown";
}

struct PolicyFeatures {
  double update_fraction = 0.0;                   // ops / directed arcs before the batch
  double shallow_parent_deletion_fraction = 0.0;  // first batch only
  bool first_batch = false;
};

struct PolicyDecision {
  bool full = false;
  int reason = 0;
  double predicted_incremental_us = 0.0;
  double predicted_full_us = 0.0;
};

class AdaptivePolicy {
 public:
  static constexpr double kPreflightFullUpdate = 0.05;
  static constexpr double kVerySparseReach = 0.02;
  static constexpr double kSparseReach = 0.20;
  static constexpr double kSparseReachFullUpdate = 0.0025;
  static constexpr double kEmaAlpha = 0.25;
  static constexpr double kLearnedFullMargin = 1.25;
  static constexpr uint64_t kFreshAge = 4;
  static constexpr uint64_t kLargeScaleVertices = 200000;

  AdaptivePolicy(uint64_t vertices, double initial_full_us, double initial_reachable_fraction)
      : vertices_(vertices),
        reach_(initial_reachable_fraction),
        ema_full_us_(initial_full_us),
        large_(vertices >= kLargeScaleVertices) {}

  PolicyDecision decide(const PolicyFeatures& f) const {
    PolicyDecision d;
    if (large_) {
      const double scale = have_inc_ ? std::sqrt((f.update_fraction + 1e-12) / (last_inc_update_fraction_ + 1e-12))
                                     : 1.0;
      const double bounded = std::clamp(scale, 0.50, 2.50);
      const double pred_inc = have_inc_ ? ema_inc_us_ * bounded * (1.0 + 3.0 * prev_affected_fraction_) : 0.0;
      const double pred_full = ema_full_us_;
      const double inc_unc = have_inc_err_ ? std::clamp(ema_inc_err_, 0.05, 0.35) : 0.35;
      const double full_unc = have_full_err_ ? std::clamp(ema_full_err_, 0.05, 0.35) : 0.20;
      const double inc_lower = pred_inc * (1.0 - inc_unc);
      const double full_upper = pred_full * (1.0 + full_unc);
      d.predicted_incremental_us = pred_inc;
      d.predicted_full_us = pred_full;
      if (f.update_fraction >= kPreflightFullUpdate) {
        d.full = true;
        d.reason = kLargePreflightFull;
      } else if (f.first_batch && f.shallow_parent_deletion_fraction > 0.0) {
        d.full = true;
        d.reason = kLargeShallowColdStart;
      } else if (!have_inc_) {
        d.full = false;
        d.reason = kLargeInitialIncrementalProbe;
      } else if (inc_lower > full_upper) {
        d.full = true;
        d.reason = kLargeConfidentFull;
      } else {
        d.full = false;
        d.reason = kLargeOverlapIncremental;
      }
      return d;
    }
    const double pred_inc = have_inc_ ? ema_inc_us_ * (1.0 + 2.0 * prev_affected_fraction_) : 0.0;
    const double pred_full = have_full_ ? ema_full_us_ : 0.0;
    d.predicted_incremental_us = pred_inc;
    d.predicted_full_us = pred_full;
    const bool fresh = have_inc_ && have_full_ && inc_age_ <= kFreshAge && full_age_ <= kFreshAge;
    if (f.update_fraction >= kPreflightFullUpdate) {
      d.full = true;
      d.reason = kScalePreflightFull;
    } else if (reach_ <= kVerySparseReach) {
      d.full = true;
      d.reason = kScaleVerySparseReach;
    } else if (reach_ < kSparseReach && f.update_fraction >= kSparseReachFullUpdate) {
      d.full = true;
      d.reason = kScaleSparseReachGuard;
    } else if (fresh && pred_inc > pred_full * kLearnedFullMargin) {
      d.full = true;
      d.reason = kScaleFreshCostModel;
    } else {
      d.full = false;
      d.reason = fresh ? kScaleConfidenceIncremental : kScaleInsufficientEvidence;
    }
    return d;
  }

  // execution_us: answer-ready cost of the arm that ran; affected: vertices
  // the incremental arm invalidated (ignored after a full recompute).
  void observe(const PolicyFeatures& f, const PolicyDecision& d, bool internal_fallback, double execution_us,
               uint64_t affected) {
    ++inc_age_;
    ++full_age_;
    const bool observed_full = d.full || internal_fallback;
    if (large_ && d.predicted_incremental_us > 0.0 && d.predicted_full_us > 0.0) {
      if (observed_full) {
        const double rel = std::abs(execution_us - d.predicted_full_us) / std::max(1.0, d.predicted_full_us);
        ema_full_err_ = have_full_err_ ? (1.0 - kEmaAlpha) * ema_full_err_ + kEmaAlpha * rel : rel;
        have_full_err_ = true;
      } else {
        const double rel =
            std::abs(execution_us - d.predicted_incremental_us) / std::max(1.0, d.predicted_incremental_us);
        ema_inc_err_ = have_inc_err_ ? (1.0 - kEmaAlpha) * ema_inc_err_ + kEmaAlpha * rel : rel;
        have_inc_err_ = true;
      }
    }
    if (observed_full) {
      ema_full_us_ = have_full_ ? (1.0 - kEmaAlpha) * ema_full_us_ + kEmaAlpha * execution_us : execution_us;
      have_full_ = true;
      full_age_ = 0;
      prev_affected_fraction_ = 0.0;
    } else {
      ema_inc_us_ = have_inc_ ? (1.0 - kEmaAlpha) * ema_inc_us_ + kEmaAlpha * execution_us : execution_us;
      have_inc_ = true;
      inc_age_ = 0;
      last_inc_update_fraction_ = f.update_fraction;
      prev_affected_fraction_ =
          static_cast<double>(affected) / static_cast<double>(std::max<uint64_t>(1, vertices_));
    }
  }

  bool large() const { return large_; }

 private:
  uint64_t vertices_;
  double reach_;
  double ema_inc_us_ = 0.0;
  double ema_full_us_;
  double last_inc_update_fraction_ = 0.0;
  double ema_inc_err_ = 0.0;
  double ema_full_err_ = 0.0;
  double prev_affected_fraction_ = 0.0;
  bool have_inc_err_ = false;
  bool have_full_err_ = false;
  bool have_inc_ = false;
  bool have_full_ = true;  // the initial BFS measured the full arm
  uint64_t inc_age_ = kFreshAge + 1;
  uint64_t full_age_ = 0;
  bool large_;
};

}  // namespace vgx
