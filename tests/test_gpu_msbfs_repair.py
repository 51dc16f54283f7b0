"""GPU parity: incremental repair of every root of a multi-source BFS.

After each update batch, MultiSourceBFS.apply must leave every root's
distances equal to the oracle's BFS of the post-batch graph (the reference's
IncrementalBFS::apply contract, incremental/bfs.hpp:29-98, for each root),
whether the batch was repaired, sent to a full pass by the policy, or
abandoned for a full pass past the affected-pair limit.  The uint8 levels,
reached counts and level count follow the repaired distances.
"""
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


def u8_of(dist):
    out = dist.astype(np.int64)
    out[dist == INF] = 255
    return out.astype(np.uint8)


def check_roots(ms, mg, orc, roots, full_check=True):
    off, adj = mg.csr()
    lv = ms.levels_u8() if full_check else None
    reached = ms.reached()
    depth = 0
    for r, root in enumerate(roots):
        full, cnt = orc.bfs(mg.n, off, adj, int(root))
        got = ms.distances(r)
        assert got.shape == full.shape
        bad = np.nonzero(got != full)[0]
        assert bad.size == 0, (r, int(root), bad[:5], got[bad[:5]], full[bad[:5]])
        if full_check:
            assert np.array_equal(lv[r], u8_of(full)), r
        assert reached[r] == cnt, (r, reached[r], cnt)
        if cnt:
            depth = max(depth, int(full[full != INF].max()) + 1)
    assert ms.level_count == depth


def make_batch(rng, mg, n_ids, k, ins_frac=0.8):
    """k ops: inserts of random pairs, deletions of existing arcs."""
    off, adj = mg.csr()
    m = len(adj)
    n_del = int(round(k * (1 - ins_frac))) if m else 0
    n_ins = k - n_del
    s = rng.integers(0, n_ids, n_ins).astype(np.uint32)
    d = rng.integers(0, n_ids, n_ins).astype(np.uint32)
    if n_del:
        pick = rng.integers(0, m, n_del)
        rows = np.searchsorted(off, pick, side="right") - 1
        s = np.concatenate([s, rows.astype(np.uint32)])
        d = np.concatenate([d, adj[pick].astype(np.uint32)])
    a = np.concatenate([np.ones(n_ins, np.uint8), np.zeros(n_del, np.uint8)])
    p = rng.permutation(len(s))
    return s[p], d[p], a[p]


def lcc_roots(orc, mg, k, rng):
    off, adj = mg.csr()
    deg = np.diff(off)
    cand = np.nonzero(deg > 0)[0]
    return [int(x) for x in rng.choice(cand, k, replace=False)]


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("k", [64, 5])
def test_repair_stream_equals_oracle(vx, orc, directed, k):
    scale = 12
    n = 1 << scale
    s, d = vx.rmat_edges(scale, 16, 7)
    g = vx.Graph(n, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    rng = np.random.default_rng(11 + k + int(directed))
    roots = lcc_roots(orc, mg, k, rng)
    ms = vx.MultiSourceBFS(g, roots)
    ms.set_repair(1, 1.0)  # always repair, never give up
    ms.run(g)
    check_roots(ms, mg, orc, roots)
    for i, size in enumerate([1, 3, 10, 40, 100, 400, 1000, 3000, 10, 1]):
        bs, bd, ba = make_batch(rng, mg, n, size, ins_frac=0.8 if i % 3 else 0.3)
        ms.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        st = ms.repair_stats()
        assert st["last_repaired"], (i, st)
        check_roots(ms, mg, orc, roots, full_check=(i % 3 == 0))
    st = ms.repair_stats()
    assert st["repaired"] == 10 and st["full_passes"] == 0


def test_deletions_of_every_shortest_parent(vx, orc):
    """Batches that delete all tree arcs into a region, so whole subtrees
    lose their support for several roots and must be re-reached."""
    n = 2000
    rng = np.random.default_rng(3)
    # a long path with random chords: deletions on the path move many
    # vertices several levels down
    s = np.arange(n - 1, dtype=np.uint32)
    d = s + 1
    cs = rng.integers(0, n, 300).astype(np.uint32)
    cd = rng.integers(0, n, 300).astype(np.uint32)
    s, d = np.concatenate([s, cs]), np.concatenate([d, cd])
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    roots = [0, 1, 500, 999, 1998, 1999, 7, 1500]
    ms = vx.MultiSourceBFS(g, roots)
    ms.set_repair(1, 1.0)
    ms.run(g)
    check_roots(ms, mg, orc, roots)
    for lo in (100, 700, 1200, 1800):
        bs = np.arange(lo, lo + 50, dtype=np.uint32)
        bd = bs + 1
        ba = np.zeros(50, np.uint8)
        ms.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        assert ms.repair_stats()["last_repaired"]
        check_roots(ms, mg, orc, roots)
    # put half of them back
    bs = np.arange(100, 150, 2, dtype=np.uint32)
    ms.apply_arrays(bs, bs + 1, np.ones(len(bs), np.uint8))
    mg.apply(bs, bs + 1, np.ones(len(bs), np.uint8))
    check_roots(ms, mg, orc, roots)


def test_growth_roots_outside_and_fallback(vx, orc):
    n = 1 << 10
    s, d = vx.rmat_edges(10, 8, 5)
    g = vx.Graph(n, True)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, True)
    mg.bulk_load_edges(np.stack([s, d], 1))
    roots = [0, 3, 17, n + 40, 0]  # one root past the last vertex, one repeated
    ms = vx.MultiSourceBFS(g, roots)
    ms.set_repair(1, 1.0)
    ms.run(g)
    check_roots(ms, mg, orc, roots)
    rng = np.random.default_rng(9)
    # growth that does not reach the outside root: repaired
    bs = np.array([5, n + 3, n + 7], np.uint32)
    bd = np.array([n + 3, n + 7, 5], np.uint32)
    ba = np.ones(3, np.uint8)
    ms.apply_arrays(bs, bd, ba)
    mg.apply(bs, bd, ba)
    assert ms.repair_stats()["last_repaired"]
    check_roots(ms, mg, orc, roots)
    # growth that brings the outside root in: a full pass
    bs = np.array([n + 40, 2], np.uint32)
    bd = np.array([2, n + 40], np.uint32)
    ms.apply_arrays(bs, bd, np.ones(2, np.uint8))
    mg.apply(bs, bd, np.ones(2, np.uint8))
    assert not ms.repair_stats()["last_repaired"]
    check_roots(ms, mg, orc, roots)
    # a tiny pair limit: the repair gives up and runs a pass
    ms.set_repair(1, 1e-6)
    bs, bd, ba = make_batch(rng, mg, mg.n, 200, ins_frac=0.2)
    ms.apply_arrays(bs, bd, ba)
    mg.apply(bs, bd, ba)
    st = ms.repair_stats()
    assert st["last_fallback"] and st["fallbacks"] == 1
    check_roots(ms, mg, orc, roots)
    # back to repairs after the pass; then always-full and adaptive modes
    ms.set_repair(1, 1.0)
    for mode in (1, 2, 0, 0):
        ms.set_repair(mode, 0.1)
        bs, bd, ba = make_batch(rng, mg, mg.n, 50)
        ms.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        assert ms.repair_stats()["last_repaired"] == (mode == 1) or mode == 0
        check_roots(ms, mg, orc, roots)


def test_graph_changed_outside_forces_a_pass(vx, orc):
    n = 1 << 10
    s, d = vx.rmat_edges(10, 8, 2)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    roots = [1, 2, 3]
    ms = vx.MultiSourceBFS(g, roots)
    ms.set_repair(1, 1.0)
    ms.run(g)
    rng = np.random.default_rng(4)
    bs, bd, ba = make_batch(rng, mg, n, 30)
    g.apply_arrays(bs, bd, ba)  # the graph moves without the traversal
    mg.apply(bs, bd, ba)
    bs, bd, ba = make_batch(rng, mg, n, 30)
    ms.apply_arrays(bs, bd, ba)
    mg.apply(bs, bd, ba)
    assert not ms.repair_stats()["last_repaired"]
    check_roots(ms, mg, orc, roots)
    bs, bd, ba = make_batch(rng, mg, n, 30)
    ms.apply_arrays(bs, bd, ba)
    mg.apply(bs, bd, ba)
    assert ms.repair_stats()["last_repaired"]
    check_roots(ms, mg, orc, roots)
