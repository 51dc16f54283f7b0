"""GPU parity: update batches whose arc counts sit on either side of the
staging sort's size tiers (one workgroup of 64, 256 or 1024 threads up to
256, 1024 and 4096 arcs, the device-wide sort above), with repeated ops on
one arc (the last op wins, so the sort must keep batch order), deletions of
existing arcs, self-loops and vertex growth across a power of two (the
one-workgroup sort packs both ends of an arc into 2b + 1 bits for a vertex
count of at most 2^b).  Adjacency is compared with the oracle's model graph
(dynamic_graph.hpp:502-549, 717-733) after every batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vx():
    import velographx_amd as vx

    assert vx.device_count() > 0, "no gfx950 device visible"
    return vx


def check_graph(g, model):
    assert g.vertex_count == model.n
    assert g.edge_count_directed == model.edge_count_directed()
    off, adj = g.export_csr()
    moff, madj = model.csr()
    assert np.array_equal(off, moff) and np.array_equal(adj, madj)
    if model.directed:
        off, adj = g.export_csr(True)
        moff, madj = model.csr(reverse=True)
        assert np.array_equal(off, moff) and np.array_equal(adj, madj)


def make_ops(rng, model, n_ids, k):
    """k ops: a third repeats one arc with alternating add/remove, a third
    deletes existing arcs, the rest insert random pairs (self-loops
    included)."""
    off, adj = model.csr()
    rep = k // 3
    u0, v0 = int(rng.integers(0, n_ids)), int(rng.integers(0, n_ids))
    s = [u0] * rep
    d = [v0] * rep
    a = [int(i % 2 == 0) for i in range(rep)]
    n_del = k // 3 if len(adj) else 0
    if n_del:
        pick = rng.integers(0, len(adj), n_del)
        rows = np.searchsorted(off, pick, side="right") - 1
        s += rows.tolist()
        d += adj[pick].tolist()
        a += [0] * n_del
    rest = k - len(s)
    s += rng.integers(0, n_ids, rest).tolist()
    d += rng.integers(0, n_ids, rest).tolist()
    a += [1] * rest
    p = rng.permutation(k)
    return (np.array(s, np.uint32)[p], np.array(d, np.uint32)[p], np.array(a, np.uint8)[p])


@pytest.mark.parametrize("directed", [False, True])
def test_sort_tier_boundaries(vx, orc, directed):
    rng = np.random.default_rng(21 + int(directed))
    n = 3000
    g = vx.Graph(n, directed)
    m = orc.ModelGraph(n, directed)
    e = rng.integers(0, n, size=(20000, 2)).astype(np.uint32)
    g.bulk_load_edges(e)
    m.bulk_load_edges(e.tolist())
    check_graph(g, m)
    # arcs per batch: ops for a directed graph, twice the ops otherwise
    arcs = [256, 257, 1024, 1025, 4096, 4097, 64, 65]
    for i, target in enumerate(arcs):
        k = target if directed else (target + 1) // 2
        # ids up to 4095 (a vertex count of 2^12), then past it (2^13)
        n_ids = 4096 if i < 4 else 4200
        s, d, a = make_ops(rng, m, n_ids, k)
        g.apply_arrays(s, d, a)
        m.apply(s, d, a)
        check_graph(g, m)
        assert g.version == m.version


@pytest.mark.parametrize("directed", [False, True])
def test_host_batch_over_several_staging_pieces(vx, orc, directed):
    """A host batch is staged and uploaded in pieces of 2^18 ops: a batch of
    two pieces and a part, with one arc's alternating ops spread over all of
    them (the last op wins) and the only ids past the vertex count in its
    last ops (the growth is decided over every piece), equals the oracle's
    model graph."""
    rng = np.random.default_rng(31 + int(directed))
    n = 20000
    g = vx.Graph(n, directed)
    m = orc.ModelGraph(n, directed)
    e = rng.integers(0, n, size=(100000, 2)).astype(np.uint32)
    g.bulk_load_edges(e)
    m.bulk_load_edges(e.tolist())
    k = 2 * (1 << 18) + 1000
    s, d, a = make_ops(rng, m, n, k - 16)
    gs = rng.integers(0, n + 5000, 16).astype(np.uint32)
    gd = (n + 4000 + np.arange(16)).astype(np.uint32)
    s, d = np.concatenate([s, gs]), np.concatenate([d, gd])
    a = np.concatenate([a, np.ones(16, np.uint8)])
    g.apply_arrays(s, d, a)
    m.apply(s, d, a)
    assert g.vertex_count == n + 4016
    check_graph(g, m)
    assert g.version == m.version
