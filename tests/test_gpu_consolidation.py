"""GPU parity: the consolidated snapshot and the neighbour probe.  The
snapshot holds exactly the source's logical graph (both directions) in
compacted rows and leaves the source untouched; the probe digest equals the
oracle's for the same rows, for the source and the snapshot alike; a
controller-driven cutover loop keeps the graph equal to the model."""
import numpy as np
import pytest

from oracle import consolidation as oc

pytestmark = pytest.mark.gpu


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


def csr_pair(g):
    out = g.export_csr()
    inn = g.export_csr(True) if g.directed else out
    return out, inn


def same_csr(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))


def check_model(g, mg):
    assert g.vertex_count == mg.n
    assert g.edge_count_directed == mg.edge_count_directed()
    off, adj = g.export_csr()
    moff, madj = mg.csr()
    assert np.array_equal(off, moff) and np.array_equal(adj, madj)
    if mg.directed:
        off, adj = g.export_csr(True)
        moff, madj = mg.csr(reverse=True)
        assert np.array_equal(off, moff) and np.array_equal(adj, madj)


def churned(vx, orc, directed, scale=12, seed=5, sizes=(3000, 200, 9000, 40, 2500)):
    from velographx_amd import workloads as wl

    n = 1 << scale
    s, d = vx.rmat_edges(scale, 16, seed)
    g = vx.Graph(n, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    eu, ev = wl.canonical_edges(s, d, directed)
    stream = wl.BatchStream(eu, ev, n, directed, seed=seed + 1)
    for k in sizes:
        bs, bd, ba = stream.next(k, 0.6)
        g.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
    # vertex growth and a long new row
    hub = n + 7
    bs = np.full(600, hub, dtype=np.uint32)
    bd = np.arange(600, dtype=np.uint32) * 3
    ba = np.ones(600, dtype=np.uint8)
    g.apply_arrays(bs, bd, ba)
    mg.apply(bs, bd, ba)
    return g, mg


def test_reference_directed_case(vx):
    """tests/test_dynamic.cpp:59-130: the reference's directed graph through
    its delta, repatch and compaction steps, then the snapshot."""
    g = vx.Graph(6, True)
    g.bulk_load_edges(np.array([[0, 2], [0, 1], [0, 2], [3, 2], [4, 2], [2, 5]], dtype=np.uint32))
    assert g.is_compact() and g.edge_count_directed == 5
    b = vx.UpdateBatch()
    b.remove(0, 2)
    b.add(1, 2)
    b.add(5, 2)
    g.apply(b)
    g.add_edge(0, 2)
    g.compact()
    assert g.is_compact()
    g.remove_edge(0, 2)
    g.add_edge(0, 4)
    g.compact()
    assert g.is_compact()
    source_storage = g.storage_bytes()
    count = g.edge_count_directed
    n0 = g.neighbors(0)
    in4 = g.in_neighbors(4)
    snap = vx.consolidate_to_csr_snapshot(g)
    assert snap.directed_edges == count
    assert snap.source_storage_bytes == source_storage
    assert snap.consolidated_storage_bytes > 0
    h = snap.graph
    assert h.is_compact() and h.edge_count_directed == count
    assert h.neighbors(0) == n0 and h.in_neighbors(4) == in4
    assert g.neighbors(0) == n0
    assert h.version == 1


@pytest.mark.parametrize("directed", [False, True])
def test_snapshot_equals_source_and_model(vx, orc, directed):
    g, mg = churned(vx, orc, directed)
    st = g.storage()
    assert st["tombstones"] > 0 and not g.is_compact()
    before = csr_pair(g)
    version = g.version
    snap = vx.consolidate_to_csr_snapshot(g)
    h = snap.graph
    # the source is unchanged
    assert g.version == version and g.storage() == st and same_csr(csr_pair(g), before)
    # the snapshot holds the same logical graph, compacted
    assert h.version == 1 and h.directed == directed and h.vertex_count == g.vertex_count
    assert h.edge_count_directed == g.edge_count_directed == snap.directed_edges
    assert same_csr(csr_pair(h), before)
    check_model(h, mg)
    hs = h.storage()
    assert hs["tombstones"] == 0 and hs["relocated_rows"] == 0 and h.is_compact()
    assert snap.source_storage_bytes == g.storage_bytes()
    assert snap.consolidated_storage_bytes == h.storage_bytes()
    assert hs["arena_used"] <= hs["arena_entries"]
    # a BFS over the snapshot equals the oracle's
    off, adj = mg.csr()
    for root in (0, 17, mg.n - 1):
        full, _ = orc.bfs(mg.n, off, adj, root)
        assert np.array_equal(vx.IncrementalBFS(h, root).distances_numpy(), full)
    # the snapshot is a working graph: later batches keep it equal to the model
    rng = np.random.default_rng(11)
    for k in (500, 5000):
        bs = rng.integers(0, mg.n, k).astype(np.uint32)
        bd = rng.integers(0, mg.n, k).astype(np.uint32)
        ba = (rng.random(k) < 0.5).astype(np.uint8)
        h.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        check_model(h, mg)


def test_snapshot_of_empty_and_compact_graphs(vx):
    e = vx.Graph(0, False)
    s = vx.consolidate_to_csr_snapshot(e)
    assert s.graph.vertex_count == 0 and s.graph.edge_count_directed == 0 and s.directed_edges == 0
    g = vx.Graph(10, True)
    g.bulk_load_edges(np.array([[1, 2], [2, 3], [9, 0]], dtype=np.uint32))
    s = vx.consolidate_to_csr_snapshot(g)
    assert same_csr(csr_pair(s.graph), csr_pair(g)) and s.graph.vertex_count == 10


@pytest.mark.parametrize("directed", [False, True])
def test_probe_digest_matches_oracle(vx, orc, directed):
    g, mg = churned(vx, orc, directed, seed=9)
    off, adj = mg.csr()
    for count in (1, 63, 4096, 3 * mg.n):
        ns, dig = g.probe_neighbors(count, 3)
        assert ns > 0.0
        assert dig == oc.probe_digest(off, adj, mg.n, count), count
    h = vx.consolidate_to_csr_snapshot(g).graph
    assert h.probe_neighbors(4096, 3)[1] == g.probe_neighbors(4096, 3)[1]
    with pytest.raises(ValueError):
        g.probe_neighbors(0, 3)
    assert vx.Graph(0, False).probe_neighbors(16, 1) == (0.0, 0)


def test_controller_driven_cutover(vx, orc):
    """The steady-state loop of benchmarks/steady_state_storage.cpp:170-212:
    per epoch a mutation batch, compact-free storage and latency signals, the
    controller's decision, and on a trigger a snapshot cut over in place of
    the graph; the live graph equals the model after every epoch and at
    least one cutover happens."""
    from velographx_amd import workloads as wl

    n = 1 << 12
    s, d = vx.rmat_edges(12, 16, 21)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    eu, ev = wl.canonical_edges(s, d, False)
    stream = wl.BatchStream(eu, ev, n, False, seed=22)
    canonical_bytes = g.storage_bytes()
    canonical_ns, _ = g.probe_neighbors(4096, 3)
    policy = vx.scale_aware_consolidation_policy(g.edge_count_directed)
    ctl = vx.ConsolidationController()
    cutovers = 0
    for epoch in range(1, 31):
        bs, bd, ba = stream.next(8000, 0.6)
        g.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        ns, _ = g.probe_neighbors(4096, 3)
        sig = vx.evaluate_consolidation(g.storage_bytes(), canonical_bytes, ns, canonical_ns, policy)
        if ctl.observe(sig, epoch):
            snap = vx.consolidate_to_csr_snapshot(g)
            assert snap.graph.probe_neighbors(4096, 1)[1] == g.probe_neighbors(4096, 1)[1]
            g = snap.graph
            canonical_bytes = g.storage_bytes()
            canonical_ns, _ = g.probe_neighbors(4096, 3)
            ctl.mark_consolidated(epoch)
            cutovers += 1
        check_model(g, mg)
    assert cutovers >= 1


@pytest.mark.parametrize("directed", [False, True])
def test_row_compaction_keeps_graph_and_objects(vx, orc, directed):
    """maybe_compact (dynamic_graph.hpp:620-631) at falling thresholds on a
    churned R-MAT graph: the logical graph and version stay, the delta and
    tombstones only shrink, threshold 0 leaves the graph compact, and BFS and
    PageRank objects made before it recompute to the oracle's answer."""
    g, mg = churned(vx, orc, directed)
    bfs = vx.IncrementalBFS(g, 0)
    pr = vx.IncrementalPageRank(g)
    version = g.version
    delta, dirty = g.delta_edge_count, g.dirty_out_segment_count()
    tomb = g.storage()["tombstones"]
    assert delta > 0 and dirty > 0 and not g.is_compact()
    assert g.delta_ratio == pytest.approx(delta / max(1, g.base_edge_count_directed))
    for threshold in (0.5, 0.1, 0.0):
        g.maybe_compact(threshold)
        check_model(g, mg)
        assert g.version == version
        assert g.delta_edge_count <= delta and g.dirty_out_segment_count() <= dirty
        assert g.storage()["tombstones"] <= tomb
        delta, dirty, tomb = g.delta_edge_count, g.dirty_out_segment_count(), g.storage()["tombstones"]
    assert g.is_compact() and tomb == 0 and dirty == 0 and g.dirty_in_segment_count() == 0
    assert g.base_edge_count_directed == g.edge_count_directed
    assert not g.maybe_compact(0.0)
    bfs.recompute()
    full, _ = orc.model_bfs(mg, 0)
    assert np.array_equal(bfs.distances_numpy(), full)
    pr.recompute()
    ref, _ = orc.pagerank_full(mg)
    assert np.abs(np.asarray(pr.values) - ref).max() < 1e-10
