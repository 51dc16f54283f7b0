"""Pins the CPU oracle (oracle/vgx_oracle.c + oracle/oracle.py) before it is
trusted as the checker for the HIP path.

* the reference's own known-answer tests (hand cases with exact distances,
  affected counts and fallback flags);
* the compiled reference engine (oracle/_ref/vgx_ref) on randomized streams,
  where it was built;
* mathematics: scipy's BFS, BFS invariants, a dense linear solve for
  PageRank, sum-to-one.
"""
import numpy as np
import pytest

from tests import refdriver

INF = 0xFFFFFFFF


def mk(orc, n, directed, edges):
    g = orc.ModelGraph(n, directed)
    g.bulk_load_edges(edges)
    return g


def apply_ops(inc, ops):
    src = [o[0] for o in ops]
    dst = [o[1] for o in ops]
    add = [o[2] for o in ops]
    inc.apply(src, dst, add)


# ---------------------------------------------------------------- BFS KATs


def test_static_bfs_known_answer(orc):
    # tests/test_static.cpp:12-18: CsrGraph({{0,1},{1,2},{2,0},{2,3}}) undirected
    g = mk(orc, 0, False, [(0, 1), (1, 2), (2, 0), (2, 3)])
    d, r = orc.model_bfs(g, 0)
    assert d.tolist() == [0, 1, 1, 2] and r == 4
    assert g.edge_count_directed() == 8


def test_ci_smoke_bfs(orc):
    # .github/workflows/ci.yml:217-227
    g = orc.ModelGraph(4, False)
    g.apply([0, 1], [1, 2], [1, 1])
    inc = orc.ModelIncrementalBFS(g, 0)
    assert inc.dist.tolist() == [0, 1, 2, INF]
    apply_ops(inc, [(2, 3, 1)])
    assert inc.dist.tolist() == [0, 1, 2, 3]


# tests/test_incremental_bfs_deletion.cpp:57-169
HAND_CASES = [
    # name, n, directed, edges, source, fallback, ops, expect{vertex: dist}, affected, full, reachable
    ("alternate-parent", 5, True, [(0, 1), (0, 2), (1, 3), (2, 3), (3, 4)], 0, 0.75,
     [(1, 3, 0)], {3: 2, 4: 3}, 0, False, None),
    ("support-fixpoint", 5, True, [(0, 1), (0, 2), (1, 3), (2, 3), (3, 4)], 0, 0.90,
     [(0, 1, 0), (0, 2, 0)], {1: INF, 2: INF, 3: INF, 4: INF}, 4, False, 1),
    ("surviving-branch", 7, True, [(0, 1), (1, 2), (2, 3), (3, 4), (0, 5), (5, 6), (6, 3)], 0, 0.75,
     [(1, 2, 0)], {2: INF, 3: 3, 4: 4}, None, False, None),
    ("mixed", 6, True, [(0, 1), (1, 2), (2, 3), (3, 4), (0, 5)], 0, 0.75,
     [(1, 2, 0), (5, 3, 1)], {2: INF, 3: 2, 4: 3}, None, False, 5),
    ("undirected", 5, False, [(0, 1), (1, 2), (2, 3), (0, 4), (4, 3)], 0, 0.75,
     [(1, 2, 0)], {2: 3}, None, False, None),
    ("final-state", 5, True, [(0, 1), (1, 2), (2, 3), (3, 4)], 0, 0.75,
     [(0, 3, 1), (0, 3, 0)], {3: 3, 4: 4}, None, False, None),
    ("fallback", 20, True, [(i, i + 1) for i in range(19)], 0, 0.05,
     [(0, 1, 0)], {}, None, True, 1),
]


@pytest.mark.parametrize("case", HAND_CASES, ids=[c[0] for c in HAND_CASES])
def test_incremental_bfs_hand_cases(orc, case):
    name, n, directed, edges, source, fb, ops, expect, affected, full, reach = case
    g = mk(orc, n, directed, edges)
    inc = orc.ModelIncrementalBFS(g, source, fb)
    apply_ops(inc, ops)
    for v, dv in expect.items():
        assert int(inc.dist[v]) == dv, (name, v)
    if affected is not None:
        assert inc.diag["affected_vertices"] == affected
    assert inc.diag["used_full_recompute"] == full
    if reach is not None:
        assert inc.reachable == reach
    full_d, full_r = orc.model_bfs(g, source)
    assert np.array_equal(inc.dist, full_d)
    assert inc.reachable == full_r == int((inc.dist != INF).sum())


def test_incremental_bfs_out_of_range_source(orc):
    # test_incremental_bfs_deletion.cpp:163-169
    g = mk(orc, 4, True, [(0, 1), (1, 2)])
    inc = orc.ModelIncrementalBFS(g, 7, 0.75)
    assert inc.reachable == 0 and (inc.dist == INF).all()


def test_incremental_bfs_planted_mistake_is_caught(orc):
    """A shortest-parent test with the wrong level (d[u] == d[v]) must change
    the affected count of the support-fixpoint case -- checks that the hand
    cases constrain the support-count logic rather than only distances."""
    g = mk(orc, 5, True, [(0, 1), (0, 2), (1, 3), (2, 3), (3, 4)])
    inc = orc.ModelIncrementalBFS(g, 0, 0.90)
    apply_ops(inc, [(0, 1, 0)])
    # vertex 1 loses its only parent; 3 keeps parent 2: exactly one affected
    assert inc.diag["affected_vertices"] == 1
    assert inc.diag["deletion_candidates"] == 1


# -------------------------------------------------- randomized differential


def _rand_stream(rng, n, m0, epochs, per_batch, directed):
    live = set()
    while len(live) < m0:
        u, v = rng.integers(0, n, 2)
        if u != v:
            live.add((int(u), int(v)) if directed else (int(min(u, v)), int(max(u, v))))
    init = sorted(live)
    batches = []
    for _ in range(epochs):
        ops = []
        for _ in range(per_batch):
            if not live or rng.random() < 0.55:
                u, v = (int(x) for x in rng.integers(0, n + 2, 2))
                ops.append((u, v, 1))
                if u != v:
                    live.add((u, v) if directed else (min(u, v), max(u, v)))
            else:
                u, v = sorted(live)[int(rng.integers(0, len(live)))]
                live.discard((u, v))
                ops.append((u, v, 0) if rng.random() < 0.5 or directed else (v, u, 0))
        batches.append(ops)
    return init, batches


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_matches_compiled_reference_bfs(orc, directed, seed):
    if not refdriver.available():
        pytest.skip("oracle/_ref/vgx_ref not built (reference tree absent)")
    rng = np.random.default_rng(seed * 7919)
    n = 40
    init, batches = _rand_stream(rng, n, 90, 60, 7, directed)
    sc = refdriver.Scenario()
    sc.graph(n, directed)
    sc.bulk(init)
    sc.cmd("bfs_new", 0, 0.35 if seed != 3 else 0.1)
    for ops in batches:
        sc.batch(ops)
        sc.cmd("bfs_apply")
    recs = sc.run()
    g = mk(orc, n, directed, init)
    inc = orc.ModelIncrementalBFS(g, 0, 0.35 if seed != 3 else 0.1)
    assert np.array_equal(recs[0][1], inc.dist)
    saw_full = 0
    for i, ops in enumerate(batches):
        apply_ops(inc, ops)
        rd, rdiag = recs[2 + 2 * i][1], recs[3 + 2 * i][1]
        assert np.array_equal(rd, inc.dist), i
        assert rdiag["reachable_count"] == inc.reachable
        assert rdiag["deletion_candidates"] == inc.diag["deletion_candidates"]
        assert rdiag["used_full_recompute"] == inc.diag["used_full_recompute"]
        saw_full += rdiag["used_full_recompute"]
        if not rdiag["used_full_recompute"]:
            assert rdiag["affected_vertices"] == inc.diag["affected_vertices"], i
        else:
            assert inc.diag["affected_vertices"] > 0


def test_bfs_matches_scipy_and_invariants(orc):
    import scipy.sparse as sp
    from scipy.sparse.csgraph import shortest_path

    rng = np.random.default_rng(5)
    n = 300
    e = rng.integers(0, n, size=(700, 2))
    g = mk(orc, n, False, e.tolist())
    off, adj = g.csr()
    rows = np.repeat(np.arange(g.n), np.diff(off).astype(np.int64))
    A = sp.csr_matrix((np.ones(adj.size), (rows, adj)), shape=(g.n, g.n))
    for root in (0, 5, 77):
        d, r = orc.bfs(g.n, off, adj, root)
        ref = shortest_path(A, unweighted=True, indices=root)
        exp = np.where(np.isinf(ref), INF, ref).astype(np.uint32)
        assert np.array_equal(d, exp)
        # invariants: |d(u)-d(v)| <= 1 on every edge inside the component
        du, dv = d[rows], d[adj]
        fin = (du != INF) & (dv != INF)
        assert (np.abs(du[fin].astype(np.int64) - dv[fin].astype(np.int64)) <= 1).all()
        assert ((du == INF) == (dv == INF)).all()


# -------------------------------------------------------------- PageRank


def _dense_pagerank(g, d):
    n = g.n
    od = g.out_degree().astype(np.float64)
    off, adj = g.csr()
    P = np.zeros((n, n))
    for u in range(n):
        for v in adj[off[u]:off[u + 1]]:
            P[u, v] = 1.0 / od[u]
    dang = (od == 0).astype(np.float64)
    # x = (1-d)/n + d*(dang.x)/n + d P^T x  ->  (I - d P^T - d/n 1 dang^T) x = (1-d)/n
    M = np.eye(n) - d * P.T - d / n * np.outer(np.ones(n), dang)
    return np.linalg.solve(M, np.full(n, (1 - d) / n))


@pytest.mark.parametrize("directed", [True, False])
def test_pagerank_full_vs_dense_solve(orc, directed):
    rng = np.random.default_rng(11)
    n = 120
    g = mk(orc, n, directed, rng.integers(0, n, size=(400, 2)).tolist())
    vals, info = orc.pagerank_full(g, 0.85, 500, 1e-14)
    x = _dense_pagerank(g, 0.85)
    assert info["converged"]
    assert np.abs(vals - x).max() < 1e-12
    assert abs(vals.sum() - 1.0) < 1e-12
    st = orc.pagerank_static(g, 0.85, 500, 1e-15)
    assert np.abs(st - x).max() < 1e-12


def test_pagerank_planted_mistake_changes_result(orc):
    """Dropping the dangling term must move the answer away from the dense
    solve: the dense check sees the term the oracle implements."""
    rng = np.random.default_rng(3)
    g = mk(orc, 50, True, rng.integers(0, 50, size=(90, 2)).tolist())
    assert (g.out_degree() == 0).any()
    vals, _ = orc.pagerank_full(g, 0.85, 500, 1e-14)
    x = _dense_pagerank(g, 0.85)
    assert np.abs(vals - x).max() < 1e-12
    n = g.n
    od = g.out_degree().astype(np.float64)
    off, adj = g.csr()
    P = np.zeros((n, n))
    for u in range(n):
        for v in adj[off[u]:off[u + 1]]:
            P[u, v] = 1.0 / od[u]
    wrong = np.linalg.solve(np.eye(n) - 0.85 * P.T, np.full(n, 0.15 / n))
    assert np.abs(wrong - vals).max() > 1e-4


def test_pagerank_localized_reference_cases(orc):
    # tests/test_pagerank_localized.cpp:31-64
    g = orc.ModelGraph(8, True)
    g.apply([0, 1, 2, 3, 4, 5, 6, 7], [1, 2, 3, 0, 5, 6, 7, 4], [1] * 8)
    pr = orc.ModelPageRank(g)
    assert pr.diag.last_full_recompute_converged and pr.diag.last_residual_linf <= 1e-12
    pr.apply([0], [2], [1], 128, 1e-13, 0.95)
    assert pr.diag.last_repair_iterations > 1
    assert 3 <= pr.diag.last_repaired_vertices < g.n
    v = pr.validate_against_full(1000, 1e-13, 1e-7, 1e-8)
    assert v["reference_converged"] and v["l1_error"] <= 1e-7 and v["linf_error"] <= 1e-8
    # :86-101 dangling transition -> full recompute
    g2 = orc.ModelGraph(4, True)
    g2.apply([0, 1], [1, 2], [1, 1])
    pr2 = orc.ModelPageRank(g2)
    assert abs(pr2.values.sum() - 1.0) < 1e-10
    pr2.apply([2], [0], [1], 128, 1e-13, 0.95)
    assert pr2.diag.last_repaired_vertices == g2.n
    assert pr2.diag.last_full_recompute_converged
    # :75-81 active-fraction fallback
    g3 = orc.ModelGraph(8, True)
    g3.apply([0, 1, 2, 3, 4, 5, 6, 7], [1, 2, 3, 0, 5, 6, 7, 4], [1] * 8)
    pr3 = orc.ModelPageRank(g3)
    pr3.apply([0], [2], [1], 64, 1e-12, 0.20)
    assert pr3.diag.last_repaired_vertices == g3.n


@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_matches_compiled_reference_pagerank(orc, seed):
    if not refdriver.available():
        pytest.skip("oracle/_ref/vgx_ref not built (reference tree absent)")
    rng = np.random.default_rng(seed)
    n = 24
    # cycles so that no vertex dangles and localized repair really runs
    init = [(i, (i + 1) % n) for i in range(n)] + [tuple(x) for x in rng.integers(0, n, size=(20, 2))]
    batches = []
    for _ in range(12):
        ops = []
        for _ in range(3):
            u, v = (int(x) for x in rng.integers(0, n, 2))
            ops.append((u, v, int(rng.random() < 0.7)))
        batches.append(ops)
    sc = refdriver.Scenario()
    sc.graph(n, True)
    sc.bulk(init)
    sc.cmd("pr_new", 0.85)
    for ops in batches:
        sc.batch(ops)
        sc.cmd("pr_apply", 24, 1e-9, 0.6)
    recs = sc.run()
    g = mk(orc, n, True, init)
    pr = orc.ModelPageRank(g, 0.85)
    assert np.abs(recs[0][1] - pr.values).max() < 1e-14
    for i, ops in enumerate(batches):
        pr.apply([o[0] for o in ops], [o[1] for o in ops], [o[2] for o in ops], 24, 1e-9, 0.6)
        vals, diag = recs[2 + 2 * i][1], recs[3 + 2 * i][1]
        assert np.abs(vals - pr.values).max() < 1e-14, i
        assert diag["last_repaired_vertices"] == pr.diag.last_repaired_vertices
        assert diag["last_repair_iterations"] == pr.diag.last_repair_iterations
        assert diag["last_full_recompute_converged"] == bool(pr.diag.last_full_recompute_converged)


# ------------------------------------------------- at-size graph model


def _random_ops(rng, n, k, loops=True):
    s = rng.integers(0, n, k).astype(np.uint32)
    d = rng.integers(0, n, k).astype(np.uint32)
    if loops:
        d[::17] = s[::17]  # self-loops are ignored
    return s, d


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_csr_model_equals_model_graph(orc, directed, seed):
    """CsrModel (orc_csr_build + orc_csr_apply) against the numpy ModelGraph:
    bulk loads with loops and duplicates, then batch streams where the same
    arc gets several ops in one batch (the last one wins), removals of absent
    arcs, inserts of present ones, and vertex growth by an insert and by a
    removal beyond the current vertex count."""
    rng = np.random.default_rng(seed)
    n = 300
    s, d = _random_ops(rng, n, 4000)
    s = np.concatenate([s, s[:500]])  # duplicates
    d = np.concatenate([d, d[:500]])
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    cm = orc.CsrModel(n, directed)
    cm.bulk_load_edges(s, d)
    for rev in (False, True):
        assert all(np.array_equal(x, y) for x, y in zip(cm.csr(rev), mg.csr(rev)))
    for b in range(12):
        k = int(rng.integers(1, 900))
        bs, bd = _random_ops(rng, n, k)
        ba = (rng.random(k) < 0.6).astype(np.uint8)
        # repeat some arcs inside the batch with the opposite op, also mirrored
        r = rng.integers(0, k, k // 4)
        bs = np.concatenate([bs, bs[r], bd[r]])
        bd = np.concatenate([bd, bd[r], bs[r]])
        ba = np.concatenate([ba, 1 - ba[r], ba[r]]).astype(np.uint8)
        if b == 4:  # growth by an insert
            bs = np.append(bs, np.uint32(n + 5))
            bd = np.append(bd, np.uint32(3))
            ba = np.append(ba, np.uint8(1))
        if b == 7:  # growth by a removal of an absent arc
            bs = np.append(bs, np.uint32(2))
            bd = np.append(bd, np.uint32(n + 40))
            ba = np.append(ba, np.uint8(0))
        mg.apply(bs, bd, ba)
        cm.apply(bs, bd, ba)
        assert cm.n == mg.n and cm.edge_count_directed() == mg.edge_count_directed()
        for rev in (False, True):
            assert all(np.array_equal(x, y) for x, y in zip(cm.csr(rev), mg.csr(rev))), (b, rev)
        if not directed:
            u, v = cm.undirected_edges()
            assert u.size * 2 == cm.edge_count_directed() and np.all(u < v)
    dist, _ = orc.model_bfs(cm, 0)
    ref, _ = orc.model_bfs(mg, 0)
    assert np.array_equal(dist, ref)


def test_pagerank_full_threads_give_the_same_bits(orc):
    """orc_pagerank_full splits its per-vertex loop over OpenMP threads; every
    rank is one sequential sum, so one thread and many give identical bits."""
    import ctypes

    gomp = ctypes.CDLL("libgomp.so.1")
    rng = np.random.default_rng(5)
    n = 5000
    s, d = _random_ops(rng, n, 40000)
    g = orc.CsrModel(n, True)
    g.bulk_load_edges(s, d)
    gomp.omp_set_num_threads(1)
    try:
        one, i1 = orc.pagerank_full(g, 0.85, 60, 0.0)
    finally:
        gomp.omp_set_num_threads(8)
    many, i8 = orc.pagerank_full(g, 0.85, 60, 0.0)
    assert i1["iterations"] == i8["iterations"] == 60
    assert np.array_equal(one, many)
