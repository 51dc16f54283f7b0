"""Connected components for the checker -- TEST INFRASTRUCTURE ONLY (only
``tests/`` imports it).

``min_labels`` gives each vertex the smallest vertex id of its component
(union-find over the undirected CSR), the representative the engine uses;
``canonical`` turns any labelling into that form, so the reference's
union-find roots (include/velographx/incremental/connected_components.hpp)
compare as a partition.  ``repaired_count`` restates the reference's repair
count (:35-62): the vertices of the pre-batch components holding an arc
whose last op in the batch removes it.  Pinned in tests/test_components.py
against the reference's own vectors
(tests/test_incremental_components_deletion.cpp) and, where oracle/_ref
builds, the reference header compiled into oracle/_ref/vgx_ref.
"""
from __future__ import annotations

import numpy as np


def _find(p, x):
    while p[x] != x:
        p[x] = p[p[x]]
        x = p[x]
    return x


def min_labels(n: int, off, adj) -> np.ndarray:
    off = np.asarray(off, dtype=np.int64)
    adj = np.asarray(adj, dtype=np.int64)
    p = list(range(n))
    for u in range(n):
        for v in adj[off[u]:off[u + 1]]:
            a, b = _find(p, u), _find(p, int(v))
            if a != b:
                p[max(a, b)] = min(a, b)
    return np.array([_find(p, v) for v in range(n)], dtype=np.uint32)


def canonical(labels) -> np.ndarray:
    """Each vertex's label replaced by the smallest vertex id sharing it."""
    labels = np.asarray(labels)
    first = {}
    out = np.empty(len(labels), dtype=np.uint32)
    for v, l in enumerate(labels.tolist()):
        out[v] = first.setdefault(l, v)
    return out


def repaired_count(labels_before, has_edge_before, src, dst, add) -> int:
    last = {}
    for u, v, a in zip(src, dst, add):
        u, v = int(u), int(v)
        if u == v:
            continue
        last[(min(u, v), max(u, v))] = int(a)
    n = len(labels_before)
    hit = set()
    for (u, v), a in last.items():
        if not a and u < n and v < n and has_edge_before(u, v):
            hit.add(int(labels_before[u]))
            hit.add(int(labels_before[v]))
    return int(sum(1 for l in labels_before if int(l) in hit)) if hit else 0
