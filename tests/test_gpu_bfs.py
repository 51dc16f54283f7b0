"""GPU parity: graph storage, full BFS and incremental BFS repair against the
CPU oracle (bit-exact distances; diagnostics as the reference defines them)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = 0xFFFFFFFF


@pytest.fixture(scope="module")
def vx():
    import velographx_amd as vx

    assert vx.device_count() > 0, "no gfx950 device visible"
    return vx


def ops_arrays(ops):
    return (np.array([o[0] for o in ops], np.uint32), np.array([o[1] for o in ops], np.uint32),
            np.array([o[2] for o in ops], np.uint8))


def batch_of(vx, ops):
    b = vx.UpdateBatch()
    for u, v, a in ops:
        (b.add if a else b.remove)(u, v)
    return b


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


# ------------------------------------------------------------ reference smoke


def test_ci_smoke(vx):
    # .github/workflows/ci.yml:188-227
    g = vx.from_numpy_edges(np.array([[0, 1], [1, 2], [2, 0]], dtype=np.uint32))
    assert g.vertex_count == 3
    assert g.neighbors(0) == [1, 2]
    import scipy.sparse as sp

    csr = sp.csr_matrix((np.ones(3), ([0, 1, 2], [1, 2, 0])), shape=(3, 3))
    g2 = vx.from_scipy_csr(csr, directed=True)
    assert g2.vertex_count == 3 and g2.neighbors(1) == [2]
    g4 = vx.Graph(4, False)
    b = vx.UpdateBatch()
    b.add(0, 1)
    b.add(1, 2)
    g4.apply(b)
    bfs = vx.IncrementalBFS(g4, 0)
    assert bfs.distances == [0, 1, 2, vx.IncrementalBFS.unreachable]
    up = vx.UpdateBatch()
    up.add(2, 3)
    bfs.apply(up)
    assert bfs.distances == [0, 1, 2, 3]


def test_static_known_answer(vx):
    # tests/test_static.cpp:12-18
    g = vx.Graph(0, False)
    g.bulk_load_edges(np.array([[0, 1], [1, 2], [2, 0], [2, 3]], dtype=np.uint32))
    assert g.edge_count_directed == 8
    assert vx.IncrementalBFS(g, 0).distances == [0, 1, 1, 2]


def test_self_loop_on_new_vertex_ignored(vx):
    # tests/test_incremental.cpp:205-217
    g = vx.Graph(2, False)
    g.add_edge(9, 9)
    assert g.vertex_count == 2 and g.edge_count_directed == 0
    assert not g.has_edge(9, 9)
    g.compact()
    assert not g.has_edge(9, 9)


def test_graph_semantics_random(vx, orc):
    rng = np.random.default_rng(3)
    for directed in (False, True):
        g = vx.Graph(30, directed)
        m = orc.ModelGraph(30, directed)
        e = rng.integers(0, 30, size=(80, 2)).astype(np.uint32)
        g.bulk_load_edges(e)
        m.bulk_load_edges(e.tolist())
        check_graph(g, m)
        assert g.version == m.version
        for step in range(25):
            k = int(rng.integers(1, 40))
            s = rng.integers(0, 36 + step, size=k).astype(np.uint32)
            d = rng.integers(0, 36 + step, size=k).astype(np.uint32)
            a = (rng.random(k) < 0.6).astype(np.uint8)
            g.apply_arrays(s, d, a)
            m.apply(s, d, a)
            check_graph(g, m)
            assert g.version == m.version
            for u in range(0, m.n, 7):
                assert g.neighbors(u) == m.neighbors(u)
            if step % 8 == 7:
                g.compact()
                check_graph(g, m)


def test_row_relocation_and_big_rows(vx, orc):
    # a hub that outgrows its slack several times, merged block-wide
    g = vx.Graph(9000, False)
    m = orc.ModelGraph(9000, False)
    g.bulk_load_edges(np.array([[0, 1], [0, 2]], dtype=np.uint32))
    m.bulk_load_edges([[0, 1], [0, 2]])
    rng = np.random.default_rng(1)
    for k in (10, 100, 1000, 5000):
        v = rng.integers(1, 9000, size=k).astype(np.uint32)
        s = np.zeros(k, np.uint32)
        a = np.ones(k, np.uint8)
        g.apply_arrays(s, v, a)
        m.apply(s, v, a)
        check_graph(g, m)
    st = g.storage()
    assert st["relocated_rows"] > 0


@pytest.mark.parametrize("directed", [False, True])
def test_insert_paths_mid_relocation_and_long_rows(vx, orc, directed):
    """Rows of every insert path: lane merges (<= 32 entries), in-place
    workgroup merges of longer rows with slack (tiles shifted from the right,
    with deletions present as tombstones), relocations of rows that outgrow
    their slots, of rows above the in-place limit (20,000 entries) and of
    rows taking more than 1,024 inserts at once; adjacency equals the model
    after every batch."""
    n = 60000
    rng = np.random.default_rng(5)
    base = [(0, v) for v in range(2, 40000, 2)]          # 19,999 entries: above the in-place limit
    base += [(1, v) for v in range(10, 3000, 3)]          # ~1,000 entries: in place while slack lasts
    base += [(2, v) for v in range(100, 140)]             # 40 entries
    base += [(u, u + 1) for u in range(3, 50000, 7)]      # short rows
    g = vx.Graph(n, directed)
    m = orc.ModelGraph(n, directed)
    g.bulk_load_edges(np.array(base, dtype=np.uint32))
    m.bulk_load_edges(base)
    check_graph(g, m)
    plans = [
        [(0, 1, 1), (0, 39999, 1), (1, 11, 1), (2, 99, 1), (2, 141, 1)],
        [(1, int(v), 1) for v in rng.integers(3000, 60000, size=60)] + [(1, 13, 0), (1, 16, 0)],
        [(1, int(v), 1) for v in rng.integers(3000, 60000, size=400)],      # outgrows: relocation
        [(2, int(v), 1) for v in rng.integers(0, 60000, size=3000)],        # > 1,024 inserts in one row
        [(0, int(v), 1) for v in rng.integers(40000, 60000, size=50)] + [(0, 4, 0)],
        [(int(u), int(v), 1) for u, v in rng.integers(0, 60000, size=(5000, 2))],
    ]
    for ops in plans:
        s_, d_, a_ = ops_arrays(ops)
        g.apply_arrays(s_, d_, a_)
        m.apply(s_, d_, a_)
        check_graph(g, m)
    full, _ = orc.model_bfs(m, 0)
    assert np.array_equal(vx.IncrementalBFS(g, 0).distances_numpy(), full)


@pytest.mark.parametrize("directed", [False, True])
def test_compaction_every_row_class(vx, orc, directed):
    """Consolidation and snapshots over rows of every packing path: short
    rows (up to 64 entries, eight-lane groups), rows cut into 1,024-entry
    chunks with and without deletions (one chunk, several chunks, a row of
    20,000 entries), empty rows; adjacency equals the model after each of
    several compactions interleaved with batches, and BFS stays exact."""
    n = 50000
    rng = np.random.default_rng(11)
    lens = {0: 20000, 1: 1024, 2: 1025, 3: 3000, 4: 65, 5: 64, 6: 63, 7: 1, 8: 700}
    base = []
    for u, L in lens.items():
        base += [(u, int(v)) for v in rng.choice(np.arange(100, n), size=L, replace=False)]
    base += [(int(u), int(u) + 1) for u in range(100, 40000, 3)]
    g = vx.Graph(n, directed)
    m = orc.ModelGraph(n, directed)
    g.bulk_load_edges(np.array(base, dtype=np.uint32))
    m.bulk_load_edges(base)
    for step in range(4):
        dels = [(u, int(v), 0) for u, L in lens.items() if u % 2 == step % 2
                for v in m.neighbors(u)[: max(1, L // 3)]]
        adds = [(int(u), int(v), 1) for u, v in rng.integers(0, n, size=(3000, 2))]
        s_, d_, a_ = ops_arrays(dels + adds)
        g.apply_arrays(s_, d_, a_)
        m.apply(s_, d_, a_)
        check_graph(g, m)          # snapshot packing with tombstones present
        g.compact()
        check_graph(g, m)          # consolidated rows
        st = g.storage()
        assert st["arena_used"] >= m.edge_count_directed()
    full, _ = orc.model_bfs(m, 0)
    assert np.array_equal(vx.IncrementalBFS(g, 0).distances_numpy(), full)


# ------------------------------------------------------------ BFS


HAND = [
    ("alternate-parent", 5, True, [(0, 1), (0, 2), (1, 3), (2, 3), (3, 4)], 0, 0.75, [(1, 3, 0)], 0, False),
    ("support-fixpoint", 5, True, [(0, 1), (0, 2), (1, 3), (2, 3), (3, 4)], 0, 0.90, [(0, 1, 0), (0, 2, 0)], 4,
     False),
    ("surviving-branch", 7, True, [(0, 1), (1, 2), (2, 3), (3, 4), (0, 5), (5, 6), (6, 3)], 0, 0.75, [(1, 2, 0)],
     None, False),
    ("mixed", 6, True, [(0, 1), (1, 2), (2, 3), (3, 4), (0, 5)], 0, 0.75, [(1, 2, 0), (5, 3, 1)], None, False),
    ("undirected", 5, False, [(0, 1), (1, 2), (2, 3), (0, 4), (4, 3)], 0, 0.75, [(1, 2, 0)], None, False),
    ("final-state", 5, True, [(0, 1), (1, 2), (2, 3), (3, 4)], 0, 0.75, [(0, 3, 1), (0, 3, 0)], None, False),
    ("fallback", 20, True, [(i, i + 1) for i in range(19)], 0, 0.05, [(0, 1, 0)], None, True),
]


@pytest.mark.parametrize("case", HAND, ids=[c[0] for c in HAND])
def test_incremental_hand_cases(vx, orc, case):
    # tests/test_incremental_bfs_deletion.cpp:57-161
    name, n, directed, edges, src, fb, ops, affected, full = case
    g = vx.Graph(n, directed)
    g.bulk_load_edges(np.array(edges, dtype=np.uint32))
    bfs = vx.IncrementalBFS(g, src, fb)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(edges)
    ref = orc.ModelIncrementalBFS(mg, src, fb)
    assert bfs.distances == ref.dist.tolist()
    bfs.apply(batch_of(vx, ops))
    ref.apply(*ops_arrays(ops))
    assert bfs.distances == ref.dist.tolist(), name
    assert bfs.reachable_count == ref.reachable
    assert bfs.last_used_full_recompute == ref.diag["used_full_recompute"] == full
    assert bfs.last_deletion_candidates == ref.diag["deletion_candidates"]
    if not full:
        assert bfs.last_affected_vertices == ref.diag["affected_vertices"]
    if affected is not None:
        assert bfs.last_affected_vertices == affected


def test_out_of_range_source(vx):
    # test_incremental_bfs_deletion.cpp:163-169
    g = vx.Graph(4, True)
    g.bulk_load_edges(np.array([[0, 1], [1, 2]], dtype=np.uint32))
    bfs = vx.IncrementalBFS(g, 7, 0.75)
    assert bfs.reachable_count == 0
    assert bfs.distances == [INF] * 4


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_incremental_randomized_vs_oracle(vx, orc, directed, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(30, 200))
    e = rng.integers(0, n, size=(n * 3, 2)).astype(np.uint32)
    fb = [0.35, 0.05, 2.0][seed % 3]
    g = vx.Graph(n, directed)
    g.bulk_load_edges(e)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(e.tolist())
    bfs = vx.IncrementalBFS(g, 0, fb)
    ref = orc.ModelIncrementalBFS(mg, 0, fb)
    assert np.array_equal(bfs.distances_numpy(), ref.dist)
    for step in range(40):
        k = int(rng.integers(1, 25))
        s = rng.integers(0, n + (step % 5), size=k).astype(np.uint32)
        d = rng.integers(0, n + (step % 5), size=k).astype(np.uint32)
        # deletions aimed at existing edges half of the time
        a = (rng.random(k) < 0.5).astype(np.uint8)
        off, adj = mg.csr()
        for i in range(k):
            if not a[i] and adj.size:
                j = int(rng.integers(0, adj.size))
                s[i] = np.searchsorted(off, j, side="right") - 1
                d[i] = adj[j]
        bfs.apply_arrays(s, d, a)
        ref.apply(s, d, a)
        got = bfs.distances_numpy()
        assert np.array_equal(got, ref.dist), (step, np.nonzero(got != ref.dist)[0][:10])
        assert bfs.reachable_count == ref.reachable
        assert bfs.last_used_full_recompute == ref.diag["used_full_recompute"]
        assert bfs.last_deletion_candidates == ref.diag["deletion_candidates"]
        if not ref.diag["used_full_recompute"]:
            assert bfs.last_affected_vertices == ref.diag["affected_vertices"], step
        full, _ = orc.model_bfs(mg, 0)
        assert np.array_equal(got, full)


def test_rmat_s16_full_bfs_and_10_batches(vx, orc):
    """R-MAT scale 16, ef 16, seed 1 (undirected): full BFS bit-exact, then 10
    batches of 1,000 ops (80% inserts of random non-edges, 20% deletions of
    existing edges) repaired and compared with a full recompute each time."""
    from velographx_amd import workloads as wl

    s, d = vx.rmat_edges(16, 16, 1)
    g = vx.Graph(1 << 16, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 16, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    assert g.edge_count_directed == mg.edge_count_directed()
    lcc = wl.largest_component_from_hub(vx, g)
    root = int(wl.roots_in_component(lcc, 1, seed=1)[0])
    bfs = vx.IncrementalBFS(g, root)
    full, reached = orc.model_bfs(mg, root)
    assert np.array_equal(bfs.distances_numpy(), full)
    assert bfs.reachable_count == reached == lcc.size
    eu, ev = wl.canonical_edges(s, d, False)
    stream = wl.BatchStream(eu, ev, 1 << 16, False, seed=1)
    for i in range(10):
        bs, bd, ba = stream.next(1000, 0.8)
        bfs.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        full, reached = orc.model_bfs(mg, root)
        got = bfs.distances_numpy()
        assert np.array_equal(got, full), (i, int((got != full).sum()))
        assert bfs.reachable_count == reached
    assert g.edge_count_directed == mg.edge_count_directed()


@pytest.mark.parametrize("fan", [63, 64, 65, 127, 128, 300, 5000])
def test_top_down_chunk_boundaries(vx, orc, fan):
    """Frontiers made of one-edge rows: a 64-edge chunk can then start
    exactly at a row boundary and hold 64 row starts (regression test for
    the load-balanced search window)."""
    edges = [(0, i) for i in range(1, fan + 1)] + [(i, fan + i) for i in range(1, fan + 1)]
    edges += [(fan + i, 2 * fan + i) for i in range(1, fan + 1, 3)]
    g = vx.Graph(0, True)
    g.bulk_load_edges(np.array(edges, dtype=np.uint32))
    mg = orc.ModelGraph(0, True)
    mg.bulk_load_edges(edges)
    full, _ = orc.model_bfs(mg, 0)
    bfs = vx.IncrementalBFS(g, 0)
    for blocks in (0, 1, 2, 3):
        bfs.set_tuning(alpha=1e-9, beta=1.0, blocks=blocks, levels_per_sync=2)
        bfs.recompute()
        assert np.array_equal(bfs.distances_numpy(), full), blocks


def test_direction_switch_levels(vx, orc):
    """Force bottom-up from the first level and top-down only: both agree
    with the oracle."""
    s, d = vx.rmat_edges(12, 16, 1)
    g = vx.Graph(1 << 12, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 12, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    for root in (0, 5, 100):
        full, _ = orc.model_bfs(mg, root)
        bfs = vx.IncrementalBFS(g, root)
        for alpha, beta in ((1e9, 1e9), (1e-9, 1.0), (15.0, 18.0)):
            bfs.set_tuning(alpha=alpha, beta=beta, blocks=0, levels_per_sync=3)
            bfs.recompute()
            assert np.array_equal(bfs.distances_numpy(), full)
            st = bfs.stats()
            if alpha > 1e8:
                assert st["bottom_up_levels"] >= 1


def test_level_shape_changes_between_traversals(vx, orc):
    """The first host check comes after as many steps as this graph's
    traversals have needed; when the graph changes shape (a 3,000-level path
    gains dense shortcuts, then loses them) the traversal still has to be
    exact, including the many further chunks of the deep shape."""
    n = 3000
    path = [(i, i + 1) for i in range(n - 1)]
    g = vx.Graph(0, False)
    g.bulk_load_edges(np.array(path, dtype=np.uint32))
    mg = orc.ModelGraph(0, False)
    mg.bulk_load_edges(path)
    bfs = vx.IncrementalBFS(g, 0)
    bfs.recompute()
    rng = np.random.default_rng(7)
    hubs = [(1, v) for v in range(3, n, 2)] + [(2, v) for v in range(4, n, 2)]
    extra = [tuple(int(x) for x in e) for e in rng.integers(0, n, size=(20000, 2)) if e[0] != e[1]]
    for ops, add in ((hubs + extra, 1), (hubs, 0)):
        s = np.array([o[0] for o in ops], np.uint32)
        d = np.array([o[1] for o in ops], np.uint32)
        a = np.full(len(ops), add, np.uint8)
        g.apply_arrays(s, d, a)
        mg.apply(s, d, a)
        full, _ = orc.model_bfs(mg, 0)
        for _ in range(2):
            bfs.recompute()
            assert np.array_equal(bfs.distances_numpy(), full)


def test_traversals_after_deletions_stay_exact(vx, orc):
    """Repeated traversals of one object stay exact, also after deletions
    change the graph between them."""
    from velographx_amd import workloads as wl

    s, d = vx.rmat_edges(14, 16, 9)
    g = vx.Graph(1 << 14, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 14, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    bfs = vx.IncrementalBFS(g, 3)
    full, _ = orc.model_bfs(mg, 3)
    for _ in range(4):
        bfs.recompute()
        assert np.array_equal(bfs.distances_numpy(), full)
    assert bfs.stats()["steps"] >= bfs.stats()["levels"]
    eu, ev = wl.canonical_edges(s, d, False)
    stream = wl.BatchStream(eu, ev, 1 << 14, False, seed=11)
    for _ in range(3):
        bs, bd, ba = stream.next(500, 0.0)  # deletions only
        g.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        full, _ = orc.model_bfs(mg, 3)
        for _ in range(3):
            bfs.recompute()
            assert np.array_equal(bfs.distances_numpy(), full)


@pytest.mark.parametrize("directed", [False, True])
def test_launch_program_does_not_depend_on_history(vx, orc, directed):
    """Every launch-shape choice (direction, tail step or inline rows, claims
    by distance with a pack step) is taken on the device from the
    traversal's own counters: a root traversed after other roots, by a fresh
    object or by one re-targeted with set_source, needs the same steps and
    levels and gives the oracle's distances."""
    from velographx_amd import workloads as wl

    scale = 16
    n = 1 << scale
    s, d = vx.rmat_edges(scale, 16, 5)
    g = vx.Graph(n, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    if directed:
        roots = [int(r) for r in np.argsort(-np.bincount(s, minlength=n))[:3]]
        roots += [int(r) for r in np.random.default_rng(3).integers(0, n, 5)]
    else:
        lcc = wl.largest_component_from_hub(vx, g)
        roots = [int(r) for r in wl.roots_in_component(lcc, 8, seed=4)]
    shapes = {}
    for r in roots:
        b = vx.IncrementalBFS(g, r)
        full, _ = orc.model_bfs(mg, r)
        assert np.array_equal(b.distances_numpy(), full), r
        st = b.stats()
        shapes[r] = (st["steps"], st["levels"], st["bottom_up_levels"])
    walker = vx.IncrementalBFS(g, roots[0])
    for r in roots[::-1] + roots:
        walker.set_source(r)
        full, _ = orc.model_bfs(mg, r)
        assert np.array_equal(walker.distances_numpy(), full), r
        st = walker.stats()
        assert (st["steps"], st["levels"], st["bottom_up_levels"]) == shapes[r], r
        assert st["reachable_count"] == int(np.count_nonzero(full != INF))


@pytest.mark.parametrize("directed", [False, True])
def test_tail_and_claim_rules_at_their_extremes(vx, orc, directed):
    """The tail and claim rules at their extremes: every bottom-up level
    with rows left goes to a tail step (tail_ratio 0) or finishes them
    inline (1e12); every top-down level claims by distance and is packed
    (claim_ratio 1e12) or always queues (0); and one or two steps per host
    check.  Distances stay the oracle's."""
    from velographx_amd import workloads as wl

    scale = 15
    n = 1 << scale
    s, d = vx.rmat_edges(scale, 16, 8)
    g = vx.Graph(n, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    if directed:
        roots = [int(np.argmax(np.bincount(s, minlength=n))), 7, 1234]
    else:
        lcc = wl.largest_component_from_hub(vx, g)
        roots = [int(r) for r in wl.roots_in_component(lcc, 3, seed=9)]
    for r in roots:
        full, _ = orc.model_bfs(mg, r)
        b = vx.IncrementalBFS(g, r)
        for tail, claim, lps in ((0.0, 0.0, 0), (1e12, 1e12, 0), (0.0, 1e12, 1), (1e12, 0.0, 2),
                                 (32.0, 5.0, 1)):
            for alpha, beta in ((15.0, 18.0), (1e9, 1e9), (1e-9, 1.0)):
                b.set_tuning(alpha=alpha, beta=beta, blocks=0, levels_per_sync=lps, tail_ratio=tail,
                             claim_ratio=claim)
                b.recompute()
                assert np.array_equal(b.distances_numpy(), full), (r, tail, claim, lps, alpha, beta)


@pytest.mark.parametrize("scale,seed,directed", [(12, 1, False), (13, 2, True), (14, 4, False), (15, 5, True),
                                                 (15, 6, False)])
def test_repeated_traversals_random_graphs(vx, orc, scale, seed, directed):
    """Four traversals per root, hub roots and random ones, over R-MAT graphs
    of several sizes, both directions, all equal to the oracle."""
    n = 1 << scale
    s, d = vx.rmat_edges(scale, 16, seed)
    g = vx.Graph(n, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    rng = np.random.default_rng(seed)
    hubs = [int(r) for r in np.argsort(-np.bincount(s, minlength=n))[:4]]
    roots = hubs + [int(r) for r in rng.integers(0, n, size=8)]
    for root in roots:
        full, _ = orc.model_bfs(mg, root)
        bfs = vx.IncrementalBFS(g, root)
        for t in range(4):
            if t:
                bfs.recompute()
            assert np.array_equal(bfs.distances_numpy(), full), (root, t)
