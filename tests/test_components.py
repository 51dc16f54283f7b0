"""Connected components: the oracle against the reference's own vectors and
the compiled reference (CPU, as a partition plus the repair count); the
device labelling against the oracle (GPU, marked)."""
import numpy as np
import pytest

from oracle import components as occ
from tests import refdriver


@pytest.fixture(scope="module")
def orc():
    from oracle import oracle as orc

    orc.build()
    return orc


def labels_of(g):
    off, adj = g.csr()
    return occ.min_labels(g.n, off, adj)


def has_edge_fn(g):
    off, adj = g.csr()
    off = np.asarray(off)
    adj = np.asarray(adj)
    return lambda u, v: bool(np.any(adj[off[u]:off[u + 1]] == v))


def test_oracle_reference_vectors(orc):
    # tests/test_incremental_components_deletion.cpp:7-40
    g = orc.ModelGraph(8, False)
    g.apply(np.array([0, 1, 2, 4, 5], np.uint32), np.array([1, 2, 3, 5, 6], np.uint32), np.ones(5, np.uint8))
    lab = labels_of(g)
    assert lab[0] == lab[3] and lab[4] == lab[6] and lab[0] != lab[4]
    for (bs, bd, ba), want in (((np.array([1], np.uint32), np.array([2], np.uint32), np.zeros(1, np.uint8)), 4),
                               ((np.array([4, 1], np.uint32), np.array([5, 2], np.uint32),
                                 np.array([0, 1], np.uint8)), 3)):
        before, he = labels_of(g), has_edge_fn(g)
        assert occ.repaired_count(before, he, bs, bd, ba) == want
        g.apply(bs, bd, ba)
    lab = labels_of(g)
    assert lab[0] == lab[3] and lab[4] != lab[5] and lab[5] == lab[6]


@pytest.mark.skipif(not refdriver.available(), reason="oracle/_ref (compiled reference) not built")
def test_oracle_matches_compiled_reference(orc):
    n = 400
    rng = np.random.default_rng(5)
    e = rng.integers(0, n, size=(380, 2))  # sparse: many components
    stream = []
    for _ in range(12):
        k = int(rng.integers(1, 60))
        bs = rng.integers(0, n + 8, k).astype(np.uint32)
        bd = rng.integers(0, n + 8, k).astype(np.uint32)
        ba = (rng.random(k) < 0.5).astype(np.uint8)
        pick = rng.integers(0, len(e), 10)  # removals of loaded edges, some effective
        stream.append((np.concatenate([bs, e[pick, 0].astype(np.uint32)]),
                       np.concatenate([bd, e[pick, 1].astype(np.uint32)]),
                       np.concatenate([ba, np.zeros(10, np.uint8)])))
    sc = refdriver.Scenario()
    sc.graph(n, False)
    sc.bulk([(int(u), int(v)) for u, v in e])
    sc.cmd("cc_new")
    for bs, bd, ba in stream:
        sc.batch(zip(bs.tolist(), bd.tolist(), ba.tolist()))
        sc.cmd("cc_apply")
    recs = sc.run()
    labs = [r[1] for r in recs if r[0] == "cclabels"]
    diags = [r[1] for r in recs if r[0] == "ccdiag"]
    g = orc.ModelGraph(n, False)
    g.bulk_load_edges(e)
    assert np.array_equal(occ.canonical(labs[0]), labels_of(g)) and diags[0] == n
    for (bs, bd, ba), want, dg in zip(stream, labs[1:], diags[1:]):
        before, he = labels_of(g), has_edge_fn(g)
        rep = occ.repaired_count(before, he, bs, bd, ba)
        g.apply(bs, bd, ba)
        assert np.array_equal(occ.canonical(want), labels_of(g))
        assert rep == dg


@pytest.mark.gpu
@pytest.mark.parametrize("scale", [10, 13])
def test_engine_matches_oracle(orc, scale):
    import velographx_amd as vx

    assert vx.device_count() > 0
    n = 1 << scale
    rng = np.random.default_rng(scale)
    # sparse R-MAT-free random graph plus an R-MAT core: many components
    s, d = vx.rmat_edges(scale, 1, 50 + scale)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    cc = vx.IncrementalComponents(g)
    assert np.array_equal(cc.labels_numpy(), labels_of(mg))
    assert cc.last_repaired_vertices == n
    for _ in range(5):
        k = int(rng.integers(1, 400))
        bs = rng.integers(0, n + 16, k).astype(np.uint32)
        bd = rng.integers(0, n + 16, k).astype(np.uint32)
        ba = (rng.random(k) < 0.5).astype(np.uint8)
        pick = rng.integers(0, s.size, 50)
        bs = np.concatenate([bs, s[pick]])
        bd = np.concatenate([bd, d[pick]])
        ba = np.concatenate([ba, np.zeros(50, np.uint8)])
        before, he = labels_of(mg), has_edge_fn(mg)
        rep = occ.repaired_count(before, he, bs, bd, ba)
        cc.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        assert np.array_equal(cc.labels_numpy(), labels_of(mg))
        assert cc.last_repaired_vertices == rep
    v = int(rng.integers(0, mg.n))
    assert cc.component(v) == labels_of(mg)[v]
    with pytest.raises(IndexError):
        cc.component(mg.n)
    cc.recompute()
    assert np.array_equal(cc.labels_numpy(), labels_of(mg))


@pytest.mark.gpu
def test_engine_reference_vectors_and_errors():
    import velographx_amd as vx

    g = vx.Graph(8, False)
    b = vx.UpdateBatch()
    for u, v in ((0, 1), (1, 2), (2, 3), (4, 5), (5, 6)):
        b.add(u, v)
    g.apply(b)
    cc = vx.IncrementalComponents(g)
    assert cc.component(0) == cc.component(3) and cc.component(4) == cc.component(6)
    assert cc.component(0) != cc.component(4)
    d = vx.UpdateBatch()
    d.remove(1, 2)
    cc.apply(d)
    assert cc.component(0) == cc.component(1) and cc.component(2) == cc.component(3)
    assert cc.component(0) != cc.component(2) and cc.component(4) == cc.component(6)
    assert cc.last_repaired_vertices == 4
    m = vx.UpdateBatch()
    m.remove(4, 5)
    m.add(1, 2)
    cc.apply(m)
    assert cc.component(0) == cc.component(3) and cc.component(4) != cc.component(5)
    assert cc.component(5) == cc.component(6)
    assert cc.last_repaired_vertices == 3
    cc.apply(vx.UpdateBatch())
    assert cc.last_repaired_vertices == 0
    with pytest.raises(ValueError):
        vx.IncrementalComponents(vx.Graph(4, True))
    assert vx.IncrementalComponents(vx.Graph(0, False)).labels_numpy().size == 0
