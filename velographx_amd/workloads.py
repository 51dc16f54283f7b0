"""Seeded synthetic workloads for the traversal path.

* R-MAT tuples of scripts/prepare_publication_datasets.py:80-129 (seed 1 by
  default), generated natively by ``rmat_edges``;
* Erdos-Renyi G(n, m) pairs;
* update-batch streams: ``k`` ops per batch, a fraction inserts of uniform
  random non-edges u != v and the rest deletes drawn uniformly from the
  current edges (the batches of the reference harness' make_batch,
  benchmarks/adaptive_policy_bfs.cpp:110-124, are sliding windows over a
  fixed edge list; these are the random-update streams the benchmark
  configurations describe);
* roots: ``splitmix64(seed ^ i) mod |LCC|`` over the sorted ids of the
  largest connected component.

All randomness is splitmix64 (Steele, Lea, Flood 2014), counter based and
vectorised with numpy, so every stream is reproducible from its seed.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """splitmix64 finaliser of the counter x (numpy uint64, wraps mod 2^64)."""
    with np.errstate(over="ignore"):
        z = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based splitmix64 stream: draw(k) returns k fresh uint64."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.ctr = 0

    def draw(self, k: int):
        c = np.arange(self.ctr, self.ctr + k, dtype=np.uint64)
        self.ctr += k
        with np.errstate(over="ignore"):
            return splitmix64(c * np.uint64(0x9E3779B97F4A7C15) ^ self.seed)

    def below(self, k: int, bound: int):
        return (self.draw(k) % np.uint64(max(bound, 1))).astype(np.int64)

    def uniform(self, k: int):
        return (self.draw(k) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def erdos_renyi(n: int, m: int, seed: int = 1):
    """m uniform vertex pairs over n vertices (loops/duplicates removed at load)."""
    s = Stream(seed * 0x5851F42D4C957F2D + 7)
    u = s.below(m, n).astype(np.uint32)
    v = s.below(m, n).astype(np.uint32)
    return u, v


def canonical_edges(src, dst, directed: bool):
    """Unique loop-free edges as (u, v) uint32 arrays; undirected pairs have u < v."""
    src = np.asarray(src, dtype=np.uint64)
    dst = np.asarray(dst, dtype=np.uint64)
    keep = src != dst
    s, d = src[keep], dst[keep]
    if not directed:
        s, d = np.minimum(s, d), np.maximum(s, d)
    keys = np.unique((s << np.uint64(32)) | d)
    return (keys >> np.uint64(32)).astype(np.uint32), (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)


class BatchStream:
    """Random update batches over a live edge set kept on the host.

    Membership is tested against the sorted initial key array minus the
    sorted array of deleted keys, plus the sorted array of keys inserted
    since.  Inserts are non-edges of the pre-batch graph, deletions are live
    edges of the pre-batch graph (swap-remove on a dense edge array), so the
    two sets are disjoint and the shuffled batch order cannot change the
    final state of any edge.
    """

    def __init__(self, u, v, n: int, directed: bool, seed: int = 1):
        self.n = int(n)
        self.directed = bool(directed)
        self.rng = Stream(0xB5AD4ECEDA1CE2A9 ^ int(seed))
        self.eu = np.array(u, dtype=np.uint32)
        self.ev = np.array(v, dtype=np.uint32)
        self.size = self.eu.size
        self.base = np.sort((self.eu.astype(np.uint64) << np.uint64(32)) | self.ev.astype(np.uint64))
        self.deleted = np.zeros(0, np.uint64)
        self.inserted = np.zeros(0, np.uint64)

    def _keys(self, a, b):
        a = a.astype(np.uint64)
        b = b.astype(np.uint64)
        if not self.directed:
            a, b = np.minimum(a, b), np.maximum(a, b)
        return (a << np.uint64(32)) | b

    def _present(self, keys):
        i = np.minimum(np.searchsorted(self.base, keys), max(self.base.size - 1, 0))
        in_base = self.base[i] == keys if self.base.size else np.zeros(keys.size, bool)
        return np.isin(keys, self.inserted) | (in_base & ~np.isin(keys, self.deleted))

    def next(self, k: int, insert_fraction: float = 0.8):
        n_ins = int(round(k * insert_fraction))
        n_del = min(k - n_ins, self.size)
        # inserts: uniform random non-edges u != v of the pre-batch graph
        ins_u = np.zeros(0, np.uint32)
        ins_v = np.zeros(0, np.uint32)
        ins_k = np.zeros(0, np.uint64)
        while ins_u.size < n_ins:
            need = (n_ins - ins_u.size) * 2 + 16
            cu = self.rng.below(need, self.n).astype(np.uint32)
            cv = self.rng.below(need, self.n).astype(np.uint32)
            ok = cu != cv
            cu, cv = cu[ok], cv[ok]
            ck = self._keys(cu, cv)
            ok = ~self._present(ck) & ~np.isin(ck, ins_k)
            cu, cv, ck = cu[ok], cv[ok], ck[ok]
            _, first = np.unique(ck, return_index=True)
            first = np.sort(first)[: n_ins - ins_u.size]
            ins_u = np.concatenate([ins_u, cu[first]])
            ins_v = np.concatenate([ins_v, cv[first]])
            ins_k = np.concatenate([ins_k, ck[first]])
        # deletes: uniform over the live edges of the pre-batch graph
        del_u = np.zeros(n_del, np.uint32)
        del_v = np.zeros(n_del, np.uint32)
        picks = self.rng.below(n_del, 1 << 62).tolist()
        for j, r in enumerate(picks):
            i = r % self.size
            del_u[j], del_v[j] = self.eu[i], self.ev[i]
            self.size -= 1
            self.eu[i] = self.eu[self.size]
            self.ev[i] = self.ev[self.size]
        del_k = self._keys(del_u, del_v)
        # commit: live array, inserted / deleted key sets
        grow = self.size + ins_u.size - self.eu.size
        if grow > 0:
            extra = max(grow, 1 << 16)
            self.eu = np.concatenate([self.eu, np.zeros(extra, np.uint32)])
            self.ev = np.concatenate([self.ev, np.zeros(extra, np.uint32)])
        lo = np.minimum(ins_u, ins_v) if not self.directed else ins_u
        hi = np.maximum(ins_u, ins_v) if not self.directed else ins_v
        self.eu[self.size:self.size + ins_u.size] = lo
        self.ev[self.size:self.size + ins_u.size] = hi
        self.size += ins_u.size
        was_inserted = np.isin(del_k, self.inserted)
        self.inserted = np.union1d(np.setdiff1d(self.inserted, del_k[was_inserted]), ins_k)
        self.deleted = np.union1d(np.setdiff1d(self.deleted, ins_k), del_k[~was_inserted])
        src = np.concatenate([ins_u, del_u])
        dst = np.concatenate([ins_v, del_v])
        add = np.concatenate([np.ones(ins_u.size, np.uint8), np.zeros(n_del, np.uint8)])
        # interleave inserts and deletes deterministically
        order = np.argsort(self.rng.draw(src.size), kind="stable")
        return src[order], dst[order], add[order]


def log_uniform_sizes(count: int, lo: float = 1.0, hi: float = 5.0, seed: int = 1):
    """Batch sizes round(10 ** U(lo, hi)) (10 .. 100,000 for the defaults)."""
    s = Stream(0x2545F4914F6CDD1D ^ int(seed))
    return np.rint(10.0 ** (lo + (hi - lo) * s.uniform(count))).astype(np.int64)


def roots_in_component(component_ids, count: int, seed: int = 1):
    """splitmix64(seed ^ i) mod |C| over the sorted component ids."""
    comp = np.sort(np.asarray(component_ids, dtype=np.uint32))
    if comp.size == 0:
        return np.zeros(0, dtype=np.uint32)
    with np.errstate(over="ignore"):
        idx = splitmix64(np.uint64(seed) ^ np.arange(count, dtype=np.uint64)) % np.uint64(comp.size)
    return comp[idx.astype(np.int64)]


def largest_component_from_hub(vx_module, graph):
    """Ids of the component of the highest-degree vertex (the largest
    component of an R-MAT / ER graph), found with one device BFS."""
    deg = graph.degrees()
    hub = int(np.argmax(deg))
    bfs = vx_module.IncrementalBFS(graph, hub)
    d = bfs.distances_numpy()
    return np.nonzero(d != 0xFFFFFFFF)[0].astype(np.uint32)
