"""Restatement of the reference's k-core -- TEST INFRASTRUCTURE ONLY (only
``tests/`` imports it, as the checker).

``core_numbers`` follows recompute_region over every vertex
(include/velographx/incremental/kcore.hpp:83-121): a min-heap of remaining
degrees, stale entries skipped, each popped vertex's core is the running
maximum of popped degrees.  ``affected_count`` is the reference's repair
region for a batch (kcore.hpp:31-56: the components of the batch's in-range
endpoints before and after the update).  Pinned in tests/test_kcore.py
against the reference's own vector (tests/test_incremental.cpp:51-56),
closed forms (complete graphs, cycles, trees, isolated vertices) and, where
oracle/_ref builds, the reference header compiled into oracle/_ref/vgx_ref.
"""
from __future__ import annotations

import heapq

import numpy as np


def core_numbers(n: int, off, adj) -> np.ndarray:
    off = np.asarray(off, dtype=np.int64)
    adj = np.asarray(adj, dtype=np.int64)
    degree = np.diff(off).astype(np.int64)[:n].copy()
    core = np.zeros(n, dtype=np.uint32)
    removed = np.zeros(n, dtype=bool)
    heap = [(int(degree[u]), u) for u in range(n)]
    heapq.heapify(heap)
    current = 0
    while heap:
        d, u = heapq.heappop(heap)
        if removed[u] or d != degree[u]:
            continue
        removed[u] = True
        current = max(current, d)
        core[u] = current
        for v in adj[off[u]:off[u + 1]]:
            if v >= n or removed[v]:
                continue
            if degree[v] != 0:
                degree[v] -= 1
            heapq.heappush(heap, (int(degree[v]), int(v)))
    return core


def _component_marks(n, off, adj, seeds, marked):
    stack = [s for s in seeds if s < n and not marked[s]]
    for s in stack:
        marked[s] = True
    while stack:
        u = stack.pop()
        for v in adj[off[u]:off[u + 1]]:
            if v < n and not marked[v]:
                marked[v] = True
                stack.append(int(v))


def affected_count(before, after, src, dst) -> int:
    """Vertices in the components (before and after the batch) of the batch's
    endpoints; `before`/`after` are (n, off, adj)."""
    n1, off1, adj1 = after
    marked = np.zeros(n1, dtype=bool)
    n0, off0, adj0 = before
    seeds = [int(x) for pair in zip(src, dst) for x in pair]
    m0 = np.zeros(n0, dtype=bool)
    _component_marks(n0, off0, adj0, [s for s in seeds if s < n0], m0)
    marked[:n0] |= m0
    _component_marks(n1, off1, adj1, seeds, marked)
    return int(marked.sum())
