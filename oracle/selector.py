"""Restatement of the reference's adaptive repair-vs-recompute policy --
TEST INFRASTRUCTURE ONLY (only ``tests/`` imports it, as the checker).

Follows benchmarks/publication_policy_bfs.cpp: the "adaptive" branch of the
decision (:162-233) and the model update after each batch (:297-330), with
the constants of benchmarks/adaptive_policy_bfs.cpp:21-31 and the state
initialisation of :86-104 (the initial BFS is the full arm's first sample;
the incremental age starts past the freshness limit).  Pinned in
tests/test_selector.py by hand cases that walk each branch.
"""
from __future__ import annotations

import math

PREFLIGHT_FULL_UPDATE = 0.05
VERY_SPARSE_REACH = 0.02
SPARSE_REACH = 0.20
SPARSE_REACH_FULL_UPDATE = 0.0025
EMA_ALPHA = 0.25
LEARNED_FULL_MARGIN = 1.25
FRESH_AGE = 4
LARGE_SCALE_VERTICES = 200000


def _clamp(x, lo, hi):
    return min(max(x, lo), hi)


class Policy:
    def __init__(self, vertices: int, initial_full_us: float, initial_reachable_fraction: float):
        self.vertices = vertices
        self.reach = initial_reachable_fraction
        self.large = vertices >= LARGE_SCALE_VERTICES
        self.prev_affected_fraction = 0.0
        self.ema_inc = 0.0
        self.ema_full = initial_full_us
        self.last_inc_update_fraction = 0.0
        self.ema_inc_err = 0.0
        self.ema_full_err = 0.0
        self.have_inc_err = False
        self.have_full_err = False
        self.have_inc = False
        self.have_full = True
        self.inc_age = FRESH_AGE + 1
        self.full_age = 0

    def decide(self, update_fraction: float, shallow: float = 0.0, first_batch: bool = False) -> dict:
        """(full, reason, predicted_incremental_us, predicted_full_us)."""
        if self.large:
            scale = (math.sqrt((update_fraction + 1e-12) / (self.last_inc_update_fraction + 1e-12))
                     if self.have_inc else 1.0)
            pred_inc = (self.ema_inc * _clamp(scale, 0.5, 2.5) * (1.0 + 3.0 * self.prev_affected_fraction)
                        if self.have_inc else 0.0)
            pred_full = self.ema_full
            inc_unc = _clamp(self.ema_inc_err, 0.05, 0.35) if self.have_inc_err else 0.35
            full_unc = _clamp(self.ema_full_err, 0.05, 0.35) if self.have_full_err else 0.20
            if update_fraction >= PREFLIGHT_FULL_UPDATE:
                full, reason = True, "large_preflight_full"
            elif first_batch and shallow > 0.0:
                full, reason = True, "large_shallow_cold_start"
            elif not self.have_inc:
                full, reason = False, "large_initial_incremental_probe"
            elif pred_inc * (1.0 - inc_unc) > pred_full * (1.0 + full_unc):
                full, reason = True, "large_uncertainty_confident_full"
            else:
                full, reason = False, "large_uncertainty_overlap_incremental"
        else:
            pred_inc = self.ema_inc * (1.0 + 2.0 * self.prev_affected_fraction) if self.have_inc else 0.0
            pred_full = self.ema_full if self.have_full else 0.0
            fresh = self.have_inc and self.have_full and self.inc_age <= FRESH_AGE and self.full_age <= FRESH_AGE
            if update_fraction >= PREFLIGHT_FULL_UPDATE:
                full, reason = True, "scale_preflight_full"
            elif self.reach <= VERY_SPARSE_REACH:
                full, reason = True, "scale_very_sparse_reach"
            elif self.reach < SPARSE_REACH and update_fraction >= SPARSE_REACH_FULL_UPDATE:
                full, reason = True, "scale_sparse_reach_guard"
            elif fresh and pred_inc > pred_full * LEARNED_FULL_MARGIN:
                full, reason = True, "scale_fresh_cost_model"
            else:
                full = False
                reason = "scale_confidence_incremental" if fresh else "scale_insufficient_evidence_incremental"
        return {"chose_full": full, "reason": reason, "predicted_incremental_us": pred_inc,
                "predicted_full_us": pred_full, "update_fraction": update_fraction}

    def observe(self, decision: dict, internal_fallback: bool, execution_us: float, affected: int) -> None:
        self.inc_age += 1
        self.full_age += 1
        observed_full = decision["chose_full"] or internal_fallback
        pi, pf = decision["predicted_incremental_us"], decision["predicted_full_us"]
        if self.large and pi > 0.0 and pf > 0.0:
            if observed_full:
                rel = abs(execution_us - pf) / max(1.0, pf)
                self.ema_full_err = (1 - EMA_ALPHA) * self.ema_full_err + EMA_ALPHA * rel if self.have_full_err else rel
                self.have_full_err = True
            else:
                rel = abs(execution_us - pi) / max(1.0, pi)
                self.ema_inc_err = (1 - EMA_ALPHA) * self.ema_inc_err + EMA_ALPHA * rel if self.have_inc_err else rel
                self.have_inc_err = True
        if observed_full:
            self.ema_full = (1 - EMA_ALPHA) * self.ema_full + EMA_ALPHA * execution_us if self.have_full else execution_us
            self.have_full = True
            self.full_age = 0
            self.prev_affected_fraction = 0.0
        else:
            self.ema_inc = (1 - EMA_ALPHA) * self.ema_inc + EMA_ALPHA * execution_us if self.have_inc else execution_us
            self.have_inc = True
            self.inc_age = 0
            self.last_inc_update_fraction = decision["update_fraction"]
            self.prev_affected_fraction = affected / max(1, self.vertices)


def shallow_parent_deletion_fraction(src, dst, add, dist, directed: bool) -> float:
    """benchmarks/adaptive_policy_bfs.cpp:126-148 on host arrays."""
    inf = 0xFFFFFFFF
    n = len(dist)
    shallow = 0
    for u, v, a in zip(src, dst, add):
        if a:
            continue
        if u < n and v < n and dist[u] != inf and dist[v] != inf:
            if dist[u] + 1 == dist[v] and dist[v] <= 3:
                shallow += 1
            if not directed and dist[v] + 1 == dist[u] and dist[u] <= 3:
                shallow += 1
    return shallow / max(1, len(src))
