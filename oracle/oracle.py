"""CPU oracle for the traversal hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The shipped
library never calls it.

* ``ModelGraph`` -- the logical semantics of ``DynamicGraph`` (reference
  ``include/velographx/storage/dynamic_graph.hpp``): simple graph, self-loops
  dropped before vertex growth (:717-724), vertex count grows to
  ``max(src, dst) + 1`` for every non-loop op including removals (:721),
  undirected ops mirror both arcs (:723), the last op on an arc wins because
  ``apply`` is sequential (:543-549), ``version`` += 1 per non-empty batch.
  Held as a sorted array of 64-bit arc keys ``(u << 32) | v``.
* ctypes wrappers over ``liboracle.so`` (``vgx_oracle.c``), which restate
  the BFS / incremental BFS / PageRank algorithms line by line.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
UNREACHABLE = 0xFFFFFFFF

_lib = None


def build(force: bool = False) -> str:
    """Compile vgx_oracle.c into oracle/liboracle.so (gcc, host only)."""
    src = os.path.join(HERE, "vgx_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-std=c11", "-fopenmp", "-shared", "-fPIC", src, "-o", LIB_PATH,
                               "-lm"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


class IncBfsDiag(ctypes.Structure):
    _fields_ = [("reachable_count", ctypes.c_uint64), ("deletion_candidates", ctypes.c_uint64),
                ("affected_vertices", ctypes.c_uint64), ("used_full_recompute", ctypes.c_int)]


class PrInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("residual_l1", ctypes.c_double),
                ("residual_linf", ctypes.c_double), ("converged", ctypes.c_int)]


class PrApplyDiag(ctypes.Structure):
    _fields_ = [("last_repaired_vertices", ctypes.c_uint64), ("last_repair_iterations", ctypes.c_uint64),
                ("last_residual_l1", ctypes.c_double), ("last_residual_linf", ctypes.c_double),
                ("last_full_recompute_converged", ctypes.c_int), ("used_full_recompute", ctypes.c_int),
                ("dangling_vertices", ctypes.c_uint64)]


_P = ctypes.c_void_p


def _setup(L):
    L.orc_bfs.restype = ctypes.c_uint64
    L.orc_bfs.argtypes = [ctypes.c_uint32, _P, _P, ctypes.c_uint32, _P]
    L.orc_incbfs_apply.restype = ctypes.c_int
    L.orc_incbfs_apply.argtypes = [ctypes.c_uint32, _P, _P, _P, _P, ctypes.c_uint32, _P, _P, _P, _P,
                                   ctypes.c_int, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64,
                                   _P, _P, _P, _P, ctypes.c_uint64, _P, ctypes.POINTER(IncBfsDiag)]
    L.orc_pagerank_full.restype = ctypes.c_int
    L.orc_pagerank_full.argtypes = [ctypes.c_uint32, _P, _P, _P, ctypes.c_double, ctypes.c_uint64,
                                    ctypes.c_double, _P, ctypes.POINTER(PrInfo)]
    L.orc_pagerank_apply.restype = ctypes.c_int
    L.orc_pagerank_apply.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, _P, _P, _P, _P,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _P, _P,
                                     ctypes.c_uint64, ctypes.c_double, ctypes.c_double, _P,
                                     ctypes.POINTER(PrApplyDiag)]
    L.orc_pagerank_static.restype = ctypes.c_int
    L.orc_pagerank_static.argtypes = [ctypes.c_uint32, _P, _P, ctypes.c_double, ctypes.c_uint64,
                                      ctypes.c_double, _P]
    L.orc_csr_build.restype = ctypes.c_uint64
    L.orc_csr_build.argtypes = [ctypes.c_uint32, ctypes.c_uint64, _P, _P, ctypes.c_int, _P, _P]
    L.orc_csr_apply.restype = ctypes.c_uint64
    L.orc_csr_apply.argtypes = [ctypes.c_uint32, _P, _P, ctypes.c_int, ctypes.c_uint64, _P, _P, _P,
                                ctypes.c_uint32, _P, _P]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# --------------------------------------------------------------------------
# logical graph model


class ModelGraph:
    """Sorted arc-key model of DynamicGraph's logical state (see module doc)."""

    def __init__(self, vertices: int = 0, directed: bool = False):
        self.n = int(vertices)
        self.directed = bool(directed)
        self.keys = np.zeros(0, dtype=np.uint64)
        self.version = 0

    # dynamic_graph.hpp:502-529
    def bulk_load_edges(self, edges):
        e = np.asarray(edges, dtype=np.uint64).reshape(-1, 2)
        e = e[e[:, 0] != e[:, 1]]
        if e.shape[0]:
            self.n = max(self.n, int(e.max()) + 1)
        k = (e[:, 0] << np.uint64(32)) | e[:, 1]
        if not self.directed:
            k = np.concatenate([k, (e[:, 1] << np.uint64(32)) | e[:, 0]])
        self.keys = np.unique(k)
        self.version += 1

    # dynamic_graph.hpp:543-549, 717-733
    def apply(self, src, dst, add):
        src = np.asarray(src, dtype=np.uint64)
        dst = np.asarray(dst, dtype=np.uint64)
        add = np.asarray(add, dtype=bool)
        if src.size == 0:
            return
        keep = src != dst
        s, d, a = src[keep], dst[keep], add[keep]
        if s.size:
            self.n = max(self.n, int(max(s.max(), d.max())) + 1)
            order = np.arange(s.size)
            k = (s << np.uint64(32)) | d
            if not self.directed:
                k = np.concatenate([k, (d << np.uint64(32)) | s])
                a = np.concatenate([a, a])
                order = np.concatenate([order, order])
            # last op per arc wins: sort by (key, order) and take run ends
            idx = np.lexsort((order, k))
            ks, as_ = k[idx], a[idx]
            last = np.ones(ks.size, dtype=bool)
            last[:-1] = ks[1:] != ks[:-1]
            fk, fa = ks[last], as_[last]
            cur = np.setdiff1d(self.keys, fk[~fa], assume_unique=True)
            self.keys = np.union1d(cur, fk[fa])
        self.version += 1

    def has_edge(self, u, v):
        key = np.uint64((int(u) << 32) | int(v))
        i = np.searchsorted(self.keys, key)
        return bool(i < self.keys.size and self.keys[i] == key)

    def edge_count_directed(self):
        return int(self.keys.size)

    def csr(self, reverse: bool = False):
        """(off uint64[n+1], adj uint32[m]) with ascending rows."""
        u = (self.keys >> np.uint64(32)).astype(np.uint32)
        v = (self.keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        if reverse:
            order = np.lexsort((u, v))
            u, v = v[order], u[order]
        cnt = np.bincount(u, minlength=self.n).astype(np.uint64)
        off = np.zeros(self.n + 1, dtype=np.uint64)
        np.cumsum(cnt, out=off[1:])
        return off, np.ascontiguousarray(v)

    def neighbors(self, u):
        off, adj = self.csr()
        return adj[off[u]:off[u + 1]].tolist()

    def out_degree(self):
        u = (self.keys >> np.uint64(32)).astype(np.int64)
        return np.bincount(u, minlength=self.n).astype(np.uint32)


class CsrModel:
    """The same logical semantics as ModelGraph, held as a canonical CSR and
    updated by the C oracle (orc_csr_build, orc_csr_apply), for the checks at
    2^20-2^22 vertices where sorting the numpy key array per batch is too
    slow.  Pinned against ModelGraph in tests/test_oracle.py."""

    def __init__(self, vertices: int = 0, directed: bool = False):
        self.n = int(vertices)
        self.directed = bool(directed)
        self.off = np.zeros(self.n + 1, dtype=np.uint64)
        self.adj = np.zeros(0, dtype=np.uint32)
        self.version = 0
        self._rev = None

    # dynamic_graph.hpp:502-529
    def bulk_load_edges(self, src, dst):
        src = _c(src, np.uint32)
        dst = _c(dst, np.uint32)
        keep = src != dst
        if keep.any():
            self.n = max(self.n, int(max(src[keep].max(), dst[keep].max())) + 1)
        off = np.empty(self.n + 1, dtype=np.uint64)
        adj = np.empty(src.size * (1 if self.directed else 2), dtype=np.uint32)
        m = lib().orc_csr_build(self.n, src.size, _ptr(src), _ptr(dst), int(self.directed), _ptr(off), _ptr(adj))
        self.off, self.adj = off, adj[:m].copy()
        self.version += 1
        self._rev = None

    # dynamic_graph.hpp:543-549, 717-733
    def apply(self, src, dst, add):
        src = _c(src, np.uint32)
        dst = _c(dst, np.uint32)
        add = _c(add, np.uint8)
        if src.size == 0:
            return
        keep = src != dst
        n_new = self.n
        if keep.any():
            n_new = max(self.n, int(max(src[keep].max(), dst[keep].max())) + 1)
        off = np.empty(n_new + 1, dtype=np.uint64)
        adj = np.empty(self.adj.size + 2 * src.size, dtype=np.uint32)
        m = lib().orc_csr_apply(self.n, _ptr(self.off), _ptr(self.adj), int(self.directed), src.size, _ptr(src),
                                _ptr(dst), _ptr(add), n_new, _ptr(off), _ptr(adj))
        self.n = n_new
        self.off, self.adj = off, adj[:m].copy()
        self.version += 1
        self._rev = None

    def edge_count_directed(self):
        return int(self.adj.size)

    def csr(self, reverse: bool = False):
        """(off uint64[n+1], adj uint32[m]) with ascending rows."""
        if not reverse:
            return self.off, self.adj
        if not self.directed:
            return self.off, self.adj
        if self._rev is None:
            rows = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(self.off).astype(np.int64))
            off = np.empty(self.n + 1, dtype=np.uint64)
            adj = np.empty(self.adj.size, dtype=np.uint32)
            lib().orc_csr_build(self.n, self.adj.size, _ptr(self.adj), _ptr(rows), 1, _ptr(off), _ptr(adj))
            self._rev = (off, adj)
        return self._rev

    def out_degree(self):
        return np.diff(self.off).astype(np.uint32)

    def undirected_edges(self):
        """Each undirected edge once as (u, v) with u < v (undirected graphs)."""
        rows = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(self.off).astype(np.int64))
        keep = rows < self.adj
        return rows[keep], self.adj[keep]


# --------------------------------------------------------------------------
# algorithm wrappers


def bfs(n, off, adj, source):
    """Queue BFS distances (uint32, UNREACHABLE) and the reached count."""
    off = _c(off, np.uint64)
    adj = _c(adj, np.uint32)
    dist = np.empty(n, dtype=np.uint32)
    reached = lib().orc_bfs(n, _ptr(off), _ptr(adj), ctypes.c_uint32(source & 0xFFFFFFFF), _ptr(dist))
    return dist, int(reached)


def model_bfs(g, source):
    off, adj = g.csr()
    return bfs(g.n, off, adj, source)


def incbfs_apply(pre: ModelGraph, post: ModelGraph, source, fallback, src, dst, add, dist_pre, reachable_pre):
    """One BasicIncrementalBFS::apply step; returns (dist_post, diag dict)."""
    po, pa = pre.csr()
    pio, pia = pre.csr(reverse=True)
    qo, qa = post.csr()
    qio, qia = post.csr(reverse=True)
    src = _c(src, np.uint32)
    dst = _c(dst, np.uint32)
    add = _c(add, np.uint8)
    dp = np.full(pre.n, UNREACHABLE, dtype=np.uint32)
    dp[:min(pre.n, len(dist_pre))] = np.asarray(dist_pre, dtype=np.uint32)[:pre.n]
    out = np.empty(max(post.n, pre.n), dtype=np.uint32)
    d = IncBfsDiag()
    lib().orc_incbfs_apply(pre.n, _ptr(po), _ptr(pa), _ptr(pio), _ptr(pia), post.n, _ptr(qo), _ptr(qa),
                           _ptr(qio), _ptr(qia), int(pre.directed), ctypes.c_uint32(source & 0xFFFFFFFF),
                           float(fallback), src.size, _ptr(src), _ptr(dst), _ptr(add), _ptr(dp),
                           int(reachable_pre), _ptr(out), ctypes.byref(d))
    n_out = post.n if src.size else pre.n
    return out[:n_out].copy(), {"reachable_count": d.reachable_count,
                                "deletion_candidates": d.deletion_candidates,
                                "affected_vertices": d.affected_vertices,
                                "used_full_recompute": bool(d.used_full_recompute)}


def pagerank_full(g, damping=0.85, max_iterations=200, tol=1e-12):
    io, ia = g.csr(reverse=True)
    od = g.out_degree()
    vals = np.empty(g.n, dtype=np.float64)
    info = PrInfo()
    lib().orc_pagerank_full(g.n, _ptr(io), _ptr(ia), _ptr(od), float(damping), int(max_iterations),
                            float(tol), _ptr(vals), ctypes.byref(info))
    return vals, {"iterations": info.iterations, "residual_l1": info.residual_l1,
                  "residual_linf": info.residual_linf, "converged": bool(info.converged)}


class ModelPageRank:
    """BasicIncrementalPageRank state machine over a ModelGraph (pagerank.hpp:29-332)."""

    def __init__(self, g: ModelGraph, damping=0.85):
        if not (0.0 <= damping <= 1.0):
            raise ValueError("PageRank damping must be in [0, 1]")
        self.g = g
        self.damping = damping
        self.diag = PrApplyDiag()
        self.recompute()

    def recompute(self, max_iterations=200, tol=1e-12):
        vals, info = pagerank_full(self.g, self.damping, max_iterations, tol)
        self.values = vals
        self.diag.dangling_vertices = int((self.g.out_degree() == 0).sum())
        self.diag.last_repaired_vertices = self.g.n
        self.diag.last_repair_iterations = info["iterations"]
        self.diag.last_residual_l1 = info["residual_l1"]
        self.diag.last_residual_linf = info["residual_linf"]
        self.diag.last_full_recompute_converged = int(info["converged"])

    def apply(self, src, dst, add, local_iterations=24, tol=1e-9, full_fallback_fraction=0.60):
        n_prev = self.values.size
        dangling_before = self.diag.dangling_vertices
        self.g.apply(src, dst, add)
        n = self.g.n
        oo, oa = self.g.csr()
        io, ia = self.g.csr(reverse=True)
        rank = np.zeros(max(n, n_prev, 1), dtype=np.float64)
        rank[:n_prev] = self.values
        s = _c(src, np.uint32)
        d = _c(dst, np.uint32)
        lib().orc_pagerank_apply(n_prev, int(dangling_before), n, _ptr(oo), _ptr(oa), _ptr(io), _ptr(ia),
                                 int(self.g.directed), float(self.damping), s.size, _ptr(s), _ptr(d),
                                 int(local_iterations), float(tol), float(full_fallback_fraction),
                                 _ptr(rank), ctypes.byref(self.diag))
        if s.size:
            self.values = rank[:n].copy()

    def validate_against_full(self, reference_max_iterations=500, reference_tol=1e-12,
                              l1_tolerance=1e-6, linf_tolerance=1e-7):
        ref, info = pagerank_full(self.g, self.damping, reference_max_iterations, reference_tol)
        err = np.abs(self.values - ref) if ref.size == self.values.size else None
        l1 = float(err.sum()) if err is not None else float("inf")
        linf = float(err.max(initial=0.0)) if err is not None else float("inf")
        return {"l1_error": l1, "linf_error": linf, "reference_converged": info["converged"],
                "reference_iterations": info["iterations"],
                "within_tolerance": info["converged"] and l1 <= l1_tolerance and linf <= linf_tolerance}


def pagerank_static(g: ModelGraph, damping=0.85, max_iterations=100, tolerance=1e-10):
    oo, oa = g.csr()
    vals = np.empty(g.n, dtype=np.float64)
    lib().orc_pagerank_static(g.n, _ptr(oo), _ptr(oa), float(damping), int(max_iterations),
                              float(tolerance), _ptr(vals))
    return vals


class ModelIncrementalBFS:
    """BasicIncrementalBFS over a ModelGraph: apply() mutates the graph like the reference."""

    def __init__(self, g: ModelGraph, source, fallback=0.35):
        self.g = g
        self.source = int(source)
        self.fallback = fallback
        self.recompute()
        self.diag = {"deletion_candidates": 0, "affected_vertices": 0, "used_full_recompute": False}

    def recompute(self):
        self.dist, self.reachable = model_bfs(self.g, self.source)

    def apply(self, src, dst, add):
        if len(src) == 0:
            self.diag = {"deletion_candidates": 0, "affected_vertices": 0, "used_full_recompute": False}
            return
        pre = ModelGraph(self.g.n, self.g.directed)
        pre.keys = self.g.keys.copy()
        self.g.apply(src, dst, add)
        self.dist, d = incbfs_apply(pre, self.g, self.source, self.fallback, src, dst, add, self.dist,
                                    self.reachable)
        self.reachable = d.pop("reachable_count")
        self.diag = d
