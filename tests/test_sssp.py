"""Unweighted incremental SSSP: the oracle against the reference's own
vector and the compiled reference (CPU), then the engine against the oracle
(GPU, marked)."""
import numpy as np
import pytest

from oracle import sssp as osp
from tests import refdriver


@pytest.fixture(scope="module")
def orc():
    from oracle import oracle as orc

    orc.build()
    return orc


def random_stream(n, seed, count, grow=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        k = int(rng.integers(1, 60))
        bs = rng.integers(0, n + grow, k).astype(np.uint32)
        bd = rng.integers(0, n + grow, k).astype(np.uint32)
        # every other batch insert-only (the relax path), the rest mixed
        ba = np.ones(k, np.uint8) if i % 2 == 0 else (rng.random(k) < 0.6).astype(np.uint8)
        out.append((bs, bd, ba))
    return out


def test_oracle_reference_vector(orc):
    # tests/test_incremental.cpp:20-32
    g = orc.ModelGraph(5, False)
    g.apply(np.array([0, 1], np.uint32), np.array([1, 2], np.uint32), np.array([1, 1], np.uint8))
    g.apply(np.array([2], np.uint32), np.array([3], np.uint32), np.array([1], np.uint8))
    s = osp.ModelSSSP(g, 0)
    assert s.distances()[3] == 3 and s.distances()[4] == osp.INF


def test_oracle_is_hop_distance(orc):
    """With every weight 1 the restated Dijkstra equals the queue BFS."""
    rng = np.random.default_rng(4)
    e = rng.integers(0, 300, size=(1500, 2))
    for directed in (False, True):
        g = orc.ModelGraph(300, directed)
        g.bulk_load_edges(e)
        off, adj = g.csr()
        for root in (0, 7, 299):
            want, _ = orc.bfs(g.n, off, adj, root)
            got = osp.ModelSSSP(g, root).distances()
            w = want.astype(np.uint64)
            w[want == 0xFFFFFFFF] = osp.INF
            assert np.array_equal(got, w)


@pytest.mark.skipif(not refdriver.available(), reason="oracle/_ref (compiled reference) not built")
@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("source", [0, 205])  # 205: outside the graph until it grows
def test_oracle_matches_compiled_reference(orc, directed, source):
    n = 200
    rng = np.random.default_rng(9)
    e = rng.integers(0, n, size=(600, 2))
    stream = random_stream(n, 10 + int(directed), 16, grow=12)
    sc = refdriver.Scenario()
    sc.graph(n, directed)
    sc.bulk([(int(u), int(v)) for u, v in e])
    sc.cmd("sssp_new", source)
    for bs, bd, ba in stream:
        sc.batch(zip(bs.tolist(), bd.tolist(), ba.tolist()))
        sc.cmd("sssp_apply")
    recs = [r for r in sc.run() if r[0] == "sdist"]
    g = orc.ModelGraph(n, directed)
    g.bulk_load_edges(e)
    m = osp.ModelSSSP(g, source)
    assert np.array_equal(recs[0][1], m.distances())
    for (bs, bd, ba), (_, want) in zip(stream, recs[1:]):
        m.apply(bs, bd, ba)
        assert np.array_equal(m.distances(), want)


@pytest.mark.gpu
@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("source", [3, 4100])  # 4100: outside the graph until it grows
def test_engine_matches_oracle(orc, directed, source):
    import velographx_amd as vx

    assert vx.device_count() > 0
    n = 1 << 12
    s, d = vx.rmat_edges(12, 16, 12)
    g = vx.Graph(n, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    eng = vx.IncrementalSSSP(g, source)
    ref = osp.ModelSSSP(mg, source)
    assert np.array_equal(eng.distances_numpy(), ref.distances())
    for bs, bd, ba in random_stream(n, 20 + int(directed), 10, grow=10):
        eng.apply_arrays(bs, bd, ba)
        ref.apply(bs, bd, ba)
        assert np.array_equal(eng.distances_numpy(), ref.distances())
    b = vx.UpdateBatch()
    b.add(1, 2)
    b.remove(5, 6)
    eng.apply(b)
    ref.apply(np.array([1, 5], np.uint32), np.array([2, 6], np.uint32), np.array([1, 0], np.uint8))
    assert np.array_equal(eng.distances_numpy(), ref.distances())
    eng.recompute()
    ref.recompute()
    assert np.array_equal(eng.distances_numpy(), ref.distances())
    assert vx.IncrementalSSSP.infinity == osp.INF
