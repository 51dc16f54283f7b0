"""Restatement of the reference's unweighted incremental SSSP -- TEST
INFRASTRUCTURE ONLY (only ``tests/`` imports it, as the checker).

Follows include/velographx/incremental/sssp.hpp:14-66 over
``oracle.ModelGraph``: recompute() is Dijkstra with unit weights from the
source (all infinite when the source is outside the graph,
dijkstra.hpp:37-49); apply() mutates the graph, then recomputes when the
batch holds any removal op and otherwise relaxes from the inserted arcs'
endpoints (sssp.hpp:38-60, propagate_dijkstra dijkstra.hpp:16-35).  Pinned in
tests/test_sssp.py against the reference's own vector
(tests/test_incremental.cpp:31-32) and, where oracle/_ref builds, against the
reference header compiled into oracle/_ref/vgx_ref on random streams.
"""
from __future__ import annotations

import heapq

import numpy as np

INF = (2**64 - 1) // 4  # dijkstra.hpp:14


class ModelSSSP:
    def __init__(self, graph, source: int):
        self.g = graph
        self.source = int(source)
        self.recompute()

    def _propagate(self, heap, off, adj):
        dist = self.dist
        while heap:
            d, u = heapq.heappop(heap)
            if u >= len(dist) or d != dist[u]:
                continue
            for v in adj[off[u]:off[u + 1]]:
                v = int(v)
                if v < len(dist) and d + 1 < dist[v]:
                    dist[v] = d + 1
                    heapq.heappush(heap, (d + 1, v))

    def recompute(self):
        n = self.g.n
        self.dist = [INF] * n
        if self.source >= n:
            return
        off, adj = self.g.csr()
        self.dist[self.source] = 0
        self._propagate([(0, self.source)], off, adj)

    def apply(self, src, dst, add):
        deletion = bool(np.any(np.asarray(add) == 0))
        self.g.apply(src, dst, add)
        if deletion:
            self.recompute()
            return
        n = self.g.n
        if len(self.dist) < n:
            self.dist.extend([INF] * (n - len(self.dist)))
        dist = self.dist
        heap = []
        for u, v, a in zip(src, dst, add):
            u, v = int(u), int(v)
            if not a or u >= len(dist) or v >= len(dist):
                continue
            if dist[u] != INF and dist[u] + 1 < dist[v]:
                dist[v] = dist[u] + 1
                heapq.heappush(heap, (dist[v], v))
            if not self.g.directed and dist[v] != INF and dist[v] + 1 < dist[u]:
                dist[u] = dist[v] + 1
                heapq.heappush(heap, (dist[u], u))
        off, adj = self.g.csr()
        self._propagate(heap, off, adj)

    def distances(self):
        return np.array(self.dist, dtype=np.uint64)
