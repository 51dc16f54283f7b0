"""Restatement of the reference's consolidation policy and steady-state
controller, and of its neighbour-probe row sequence -- TEST INFRASTRUCTURE
ONLY (only ``tests/`` imports it, as the checker).

* ``scale_aware_policy``, ``evaluate``, ``Controller`` follow
  include/velographx/storage/consolidation.hpp:18-119.  Pinned in
  tests/test_consolidation.py against the reference's own test vectors
  (tests/test_consolidation_controller.cpp:10-54, tests/test_dynamic.cpp:132-139)
  and, where oracle/_ref builds, against the reference header compiled into
  oracle/_ref/vgx_ref on random signal streams.
* ``probe_rows`` is the row sequence of benchmarks/steady_state_storage.cpp:61-72
  (u_i = (i * 11400714819323198485 mod 2^64) mod n); ``probe_digest`` sums,
  mod 2^64, mix(row length) + sum of mix(v) over each probed row, with
  mix = splitmix64's output function.  The reference folds its rows into a
  chained hash instead; the digest here is the engine's own order-free
  observable, pinned by splitmix64's published outputs for seed 0 and by
  exact Python-integer arithmetic on small graphs.
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
PROBE_MULT = 11400714819323198485  # == GOLDEN
LARGE_GRAPH_DIRECTED_ARCS = 100_000_000


def scale_aware_policy(directed_edges: int, latency_ratio: float = 1.25):
    """consolidation.hpp:27-33 -> (max_storage_growth_ratio, max_neighbor_latency_ratio)."""
    return (1.50 if directed_edges >= LARGE_GRAPH_DIRECTED_ARCS else 1.25, latency_ratio)


def evaluate(current_bytes, canonical_bytes, current_latency, canonical_latency, policy=(1.25, 1.25)):
    """consolidation.hpp:43-61 -> (storage_ratio, latency_ratio, storage_exceeded,
    latency_exceeded, should_consolidate)."""
    sr = 1.0 if canonical_bytes == 0 else float(current_bytes) / float(canonical_bytes)
    lr = 1.0 if canonical_latency <= 0.0 else current_latency / canonical_latency
    se = sr >= policy[0]
    le = lr >= policy[1]
    return (sr, lr, se, le, se or le)


class Controller:
    """consolidation.hpp:67-119."""

    def __init__(self, min_epochs_between=10, breach_samples=5, min_growth_for_latency=1.10):
        self.min_epochs_between = min_epochs_between
        self.breach_samples = breach_samples
        self.min_growth_for_latency = min_growth_for_latency
        self.streak = 0
        self.last = 0

    def observe(self, signal, epoch: int) -> bool:
        sr, _lr, se, le, _sc = signal
        if se:  # :77-81 the storage bound bypasses hysteresis and cooldown
            self.streak = 0
            return True
        cooldown_complete = self.last == 0 or epoch >= self.last + self.min_epochs_between  # :83-84
        if not cooldown_complete:  # :85-92
            self.streak = 0
            return False
        growth = sr >= self.min_growth_for_latency  # :94-95
        if le and growth:  # :96-100
            self.streak += 1
        else:
            self.streak = 0
        return le and growth and self.streak >= self.breach_samples  # :102-103

    def mark_consolidated(self, epoch: int) -> None:  # :106-109
        self.last = epoch
        self.streak = 0


def probe_rows(n: int, count: int) -> np.ndarray:
    """steady_state_storage.cpp:64-65: the vertex of probe i, with the
    multiplication wrapping at 2^64 as the reference's unsigned 64-bit one."""
    i = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return (i * np.uint64(PROBE_MULT)) % np.uint64(n)


def mix(x: np.ndarray) -> np.ndarray:
    """splitmix64's output function applied to x + 0x9E3779B97F4A7C15
    (Steele, Lea, Flood, OOPSLA 2014), elementwise over uint64."""
    z = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def probe_digest(off: np.ndarray, adj: np.ndarray, n: int, count: int) -> int:
    """Digest of `count` probes over a CSR (off[n+1], adj): the sum mod 2^64
    of mix(len) + sum of mix(v) over each probed row."""
    if n == 0:
        return 0
    off = np.asarray(off, dtype=np.int64)
    row_sum = np.zeros(n, dtype=np.uint64)
    if len(adj):
        m = mix(np.asarray(adj, dtype=np.uint64))
        # per-row sums mod 2^64 (uint64 addition wraps)
        owner = np.repeat(np.arange(n), np.diff(off))
        np.add.at(row_sum, owner, m)
    with np.errstate(over="ignore"):
        row_h = row_sum + mix(np.diff(off).astype(np.uint64))
    u = probe_rows(n, count)
    total = 0
    for h in row_h[u.astype(np.int64)]:
        total = (total + int(h)) & MASK64
    return total
