"""GPU parity: the static CsrGraph with bfs_distances and pagerank
(include/velographx/csr_graph.hpp, src/algorithms.cpp:11-27, 56-81).  BFS is
bit-exact against the oracle's queue BFS; the FP64 PageRank is within 1e-12
(L1 and max-abs) of the oracle's push-form restatement."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = 0xFFFFFFFF


@pytest.fixture(scope="module")
def vx():
    import velographx_amd as vx

    assert vx.device_count() > 0, "no gfx950 device visible"
    return vx


@pytest.fixture(scope="module")
def orc():
    from oracle import oracle as orc

    orc.build()
    return orc


def test_reference_static_vectors(vx):
    """tests/test_static.cpp:12-41 (the CsrGraph, BFS and PageRank lines)."""
    g = vx.CsrGraph(np.array([[0, 1], [1, 2], [2, 0], [2, 3]], dtype=np.uint32), False)
    assert g.vertex_count() == 4
    assert g.edge_entry_count() == 8
    assert g.degree(2) == 3
    d = vx.bfs_distances(g, 0)
    assert d.tolist() == [0, 1, 1, 2]
    pr = vx.pagerank(g)
    assert pr.size == 4 and abs(pr.sum() - 1.0) < 1e-9
    with pytest.raises(IndexError):
        vx.bfs_distances(g, 99)


def test_construction_semantics(vx):
    # self-loops dropped before the vertex count, duplicates removed, both
    # directions kept (csr_graph.cpp:7-42)
    g = vx.CsrGraph(np.array([[0, 3], [3, 0], [0, 3], [9, 9], [5, 2]], dtype=np.uint32), True)
    assert g.vertex_count() == 6 and g.directed()
    assert g.edge_entry_count() == 3
    assert g.neighbors(0) == [3] and g.neighbors(3) == [0] and g.in_neighbors(2) == [5]
    assert g.has_edge(5, 2) and not g.has_edge(2, 5) and not g.has_edge(9, 9) and not g.has_edge(0, 100)
    with pytest.raises(IndexError):
        g.neighbors(6)
    with pytest.raises(IndexError):
        g.degree(6)
    e = vx.CsrGraph(np.zeros((0, 2), dtype=np.uint32))
    assert e.vertex_count() == 0 and e.edge_entry_count() == 0
    assert vx.pagerank(e).size == 0
    loops = vx.CsrGraph(np.array([[4, 4]], dtype=np.uint32))
    assert loops.vertex_count() == 0


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("scale", [10, 13])
def test_rmat_vs_oracle(vx, orc, directed, scale):
    s, d = vx.rmat_edges(scale, 16, 3 + scale)
    edges = np.stack([s, d], 1)
    g = vx.CsrGraph(edges, directed)
    mg = orc.ModelGraph(0, directed)
    mg.bulk_load_edges(edges)
    assert g.vertex_count() == mg.n and g.edge_entry_count() == mg.edge_count_directed()
    off, adj = mg.csr()
    rng = np.random.default_rng(scale)
    for v in rng.integers(0, mg.n, 20):
        v = int(v)
        assert g.neighbors(v) == adj[off[v]:off[v + 1]].tolist()
        assert g.degree(v) == int(off[v + 1] - off[v])
    for root in [0, mg.n - 1] + [int(x) for x in rng.integers(0, mg.n, 6)]:
        want, _ = orc.bfs(mg.n, off, adj, root)
        assert np.array_equal(vx.bfs_distances(g, root), want), root
    for damping, iters, tol in ((0.85, 100, 1e-10), (0.5, 7, 0.0), (1.0, 30, 1e-6), (0.0, 3, 0.0)):
        got = vx.pagerank(g, damping, iters, tol)
        want = orc.pagerank_static(mg, damping, iters, tol)
        assert np.abs(got - want).sum() <= 1e-12 and np.abs(got - want).max() <= 1e-12, (damping, iters, tol)


def test_pagerank_edge_cases(vx, orc):
    edges = np.array([[0, 1], [1, 2], [2, 0], [3, 0], [4, 5]], dtype=np.uint32)  # 5 is dangling
    g = vx.CsrGraph(edges, True)
    assert np.array_equal(vx.pagerank(g, 0.85, 0), np.full(6, 1.0 / 6))
    mg = orc.ModelGraph(0, True)
    mg.bulk_load_edges(edges)
    assert np.allclose(vx.pagerank(g, 0.85, 500, 1e-15), orc.pagerank_static(mg, 0.85, 500, 1e-15), rtol=0,
                       atol=1e-13)
    for bad in (-0.1, 1.5, float("nan")):
        with pytest.raises(ValueError):
            vx.pagerank(g, bad)


def test_same_distances_as_dynamic_graph(vx):
    """tests/test_storage_independence.cpp:23-28: the static and the dynamic
    graph of the same edges give the same BFS."""
    s, d = vx.rmat_edges(12, 16, 9)
    g = vx.CsrGraph(np.stack([s, d], 1), False)
    dg = vx.Graph(0, False)
    dg.bulk_load_arrays(s, d)
    for root in (0, 5, 77):
        assert np.array_equal(vx.bfs_distances(g, root), vx.IncrementalBFS(dg, root).distances_numpy())
