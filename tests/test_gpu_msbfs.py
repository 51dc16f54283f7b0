"""GPU parity: multi-source BFS (up to 64 roots per pass).  Every root's
distances are bit-exact against the CPU oracle's single-root BFS and against
the single-root GPU path; the uint8 level arrays equal the distances with
255 for unreachable."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = 0xFFFFFFFF


@pytest.fixture(scope="module")
def vx():
    import velographx_amd as vx

    assert vx.device_count() > 0, "no gfx950 device visible"
    return vx


def u8_of(dist):
    out = dist.astype(np.int64)
    out[dist == INF] = 255
    return out.astype(np.uint8)


def check_all_roots(ms, g, mg, orc, roots):
    ms.run(g)
    off, adj = mg.csr()
    lv = ms.levels_u8()
    assert lv.shape == (len(roots), mg.n)
    reached = ms.reached()
    depth = 0
    for r, root in enumerate(roots):
        full, cnt = orc.bfs(mg.n, off, adj, int(root))
        got = ms.distances(r)
        assert np.array_equal(got, full), (r, int(root), int((got != full).sum()))
        assert np.array_equal(lv[r], u8_of(full)), r
        assert reached[r] == cnt
        if cnt:
            depth = max(depth, int(full[full != INF].max()) + 1)
    assert ms.level_count == depth


def test_tiny_directed_and_duplicates(vx, orc):
    edges = [(0, 1), (1, 2), (2, 3), (3, 1), (4, 0), (5, 5)]
    g = vx.Graph(0, True)
    g.bulk_load_edges(np.array(edges, dtype=np.uint32))
    mg = orc.ModelGraph(0, True)
    mg.bulk_load_edges(edges)
    roots = [0, 0, 4, 3, 1, 99]  # a repeated root and one past the last vertex
    ms = vx.MultiSourceBFS(g, roots)
    check_all_roots(ms, g, mg, orc, roots)


@pytest.mark.parametrize("k", [1, 7, 63, 64])
def test_rmat_s14_roots(vx, orc, k):
    s, d = vx.rmat_edges(14, 16, 1)
    g = vx.Graph(1 << 14, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 14, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    roots = np.random.default_rng(k).integers(0, 1 << 14, size=k).astype(np.uint32)
    ms = vx.MultiSourceBFS(g, roots.tolist())
    check_all_roots(ms, g, mg, orc, roots)


def test_directed_rmat_long_rows(vx, orc):
    """Directed R-MAT: pulls run over in-rows, hub in-rows exceed one wave."""
    s, d = vx.rmat_edges(13, 24, 3)
    g = vx.Graph(1 << 13, True)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 13, True)
    mg.bulk_load_edges(np.stack([s, d], 1))
    assert max(np.diff(mg.csr(reverse=True)[0])) > 64
    roots = list(range(0, 64 * 97, 97))
    check_all_roots(vx.MultiSourceBFS(g, roots), g, mg, orc, roots)


def test_after_updates_matches_single_root(vx, orc):
    """After update batches (tombstones, relocated rows) the 64-root pass
    equals 64 single-root GPU BFS runs."""
    from velographx_amd import workloads as wl

    s, d = vx.rmat_edges(12, 16, 2)
    g = vx.Graph(1 << 12, False)
    g.bulk_load_arrays(s, d)
    eu, ev = wl.canonical_edges(s, d, False)
    stream = wl.BatchStream(eu, ev, 1 << 12, False, seed=4)
    for _ in range(5):
        g.apply_arrays(*stream.next(2000, 0.8))
    roots = np.random.default_rng(9).integers(0, 1 << 12, size=64).tolist()
    ms = vx.MultiSourceBFS(g, roots)
    ms.run(g)
    for r, root in enumerate(roots):
        single = vx.IncrementalBFS(g, root).distances_numpy()
        assert np.array_equal(ms.distances(r), single), r


def test_path_deeper_than_u8(vx):
    """A 300-vertex path: uint32 distances stay exact, the uint8 export
    refuses (depth beyond 254)."""
    n = 300
    g = vx.Graph(0, False)
    g.bulk_load_edges(np.array([(i, i + 1) for i in range(n - 1)], dtype=np.uint32))
    ms = vx.MultiSourceBFS(g, [0, n - 1])
    ms.run(g)
    assert np.array_equal(ms.distances(0), np.arange(n, dtype=np.uint32))
    assert np.array_equal(ms.distances(1), np.arange(n - 1, -1, -1).astype(np.uint32))
    with pytest.raises(Exception):
        ms.levels_u8()


def test_root_count_limits(vx):
    g = vx.Graph(4, False)
    with pytest.raises(ValueError):
        vx.MultiSourceBFS(g, [])
    with pytest.raises(ValueError):
        vx.MultiSourceBFS(g, list(range(65)))


@pytest.mark.parametrize("alpha", [1e-9, 4.0, 1e9])
@pytest.mark.parametrize("directed", [False, True])
def test_push_and_pull_levels_agree(vx, orc, alpha, directed):
    """alpha = 1e-9: every level pushes; 1e9: only the roots' level pushes,
    the rest pull; 4: mixed.  All give the oracle's distances."""
    s, d = vx.rmat_edges(13, 16, 5)
    g = vx.Graph(1 << 13, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 13, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    roots = np.random.default_rng(3).integers(0, 1 << 13, size=64).tolist()
    ms = vx.MultiSourceBFS(g, roots)
    for blocks in (0, 1, 7):
        ms.set_tuning(blocks=blocks, alpha=alpha)
        check_all_roots(ms, g, mg, orc, roots)


def test_autotune_keeps_results(vx, orc):
    s, d = vx.rmat_edges(12, 16, 6)
    g = vx.Graph(1 << 12, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 12, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    roots = list(range(0, 4096, 64))
    ms = vx.MultiSourceBFS(g, roots)
    best = ms.autotune(1)
    assert best["alpha"] > 0 and best["ms"] > 0
    assert best["short_row"] in (8, 16, 32, 64, 128) and best["chunk"] in (256, 512, 1024, 2048, 4096)
    assert best["beta"] in (0.0, 4.0, 16.0, 64.0)
    assert ms.params()["short_row"] == best["short_row"]
    check_all_roots(ms, g, mg, orc, roots)


@pytest.mark.parametrize("directed", [False, True])
def test_degree_buckets_and_beta_keep_results(vx, orc, directed):
    """Every row in one lane (short_row 1024), no row in a lane (short_row
    1), the smallest and a large chunk, the beta rule on and off, eager and
    captured: the oracle's distances each time."""
    s, d = vx.rmat_edges(13, 16, 8)
    g = vx.Graph(1 << 13, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 13, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    roots = np.random.default_rng(5).integers(0, 1 << 13, size=64).tolist()
    ms = vx.MultiSourceBFS(g, roots)
    for graph_mode in (False, True):
        ms.set_graph(graph_mode)
        for short_row, chunk, alpha, beta in ((1, 64, 4.0, 0.0), (1024, 4096, 4.0, 0.0), (8, 256, 2.0, 64.0),
                                              (128, 1 << 20, 32.0, 4.0), (32, 1024, 1e9, 16.0)):
            ms.set_params(0, alpha, beta, short_row, chunk)
            assert ms.params() == {"blocks": 0, "alpha": alpha, "beta": beta, "short_row": short_row, "chunk": chunk}
            check_all_roots(ms, g, mg, orc, roots)
    with pytest.raises(ValueError):
        ms.set_params(0, 4.0, 0.0, 0, 1024)
    with pytest.raises(ValueError):
        ms.set_params(0, 4.0, 0.0, 32, 32)
    with pytest.raises(ValueError):
        ms.set_params(0, 4.0, -1.0, 32, 1024)


def test_graph_replay_across_updates(vx, orc):
    """The captured launch program is replayed after every update batch
    (uneven sizes); each replay equals the oracle for every root."""
    from velographx_amd import workloads as wl

    s, d = vx.rmat_edges(12, 16, 4)
    g = vx.Graph(1 << 12, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << 12, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    roots = np.random.default_rng(2).integers(0, 1 << 12, size=64).tolist()
    ms = vx.MultiSourceBFS(g, roots)
    ms.set_graph(True)
    check_all_roots(ms, g, mg, orc, roots)
    eu, ev = wl.canonical_edges(s, d, False)
    stream = wl.BatchStream(eu, ev, 1 << 12, False, seed=8)
    for k in (10, 3000, 150, 20000, 40):
        bs, bd, ba = stream.next(k, 0.8)
        g.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        check_all_roots(ms, g, mg, orc, roots)
    assert ms.stats()["graph_replays"] >= 6


def test_graph_mode_deep_path(vx):
    """Deeper than one captured program (16 levels): the pass continues
    eagerly from where the replay stopped."""
    n = 100
    g = vx.Graph(0, False)
    g.bulk_load_edges(np.array([(i, i + 1) for i in range(n - 1)], dtype=np.uint32))
    ms = vx.MultiSourceBFS(g, [0, 50])
    ms.set_graph(True)
    for _ in range(2):
        ms.run(g)
        assert np.array_equal(ms.distances(0), np.arange(n, dtype=np.uint32))
        assert np.array_equal(ms.distances(1), np.abs(np.arange(n) - 50).astype(np.uint32))
        assert ms.level_count == n


@pytest.mark.parametrize("graph_mode", [True, False])
def test_program_length_follows_depth(vx, graph_mode):
    """The launch program is sized from the previous pass's depth: a shallow
    pass (a star) is followed by a 60-level pass (the star turned into a
    path), which continues over several short programs, then by the star
    again; every pass is exact."""
    n = 60
    star = np.array([(0, i) for i in range(1, n)], dtype=np.uint32)
    path = np.array([(i, i + 1) for i in range(n - 1)], dtype=np.uint32)
    g = vx.Graph(0, False)
    g.bulk_load_edges(star)
    ms = vx.MultiSourceBFS(g, [0, n - 1])
    ms.set_graph(graph_mode)

    def check(edges):
        ms.run(g)
        from collections import deque

        adj = [[] for _ in range(n)]
        for u, v in edges:
            adj[u].append(v)
            adj[v].append(u)
        depth = 0
        for r, root in enumerate((0, n - 1)):
            want = np.full(n, INF, dtype=np.uint32)
            want[root] = 0
            q = deque([root])
            while q:
                u = q.popleft()
                for v in adj[u]:
                    if want[v] == INF:
                        want[v] = want[u] + 1
                        q.append(v)
            assert np.array_equal(ms.distances(r), want), r
            depth = max(depth, int(want.max()) + 1)
        assert ms.level_count == depth

    for _ in range(2):
        check(star)
    ones = np.ones(n - 1, dtype=np.uint8)
    zeros = np.zeros(n - 1, dtype=np.uint8)
    g.apply_arrays(np.concatenate([star[:, 0], path[:, 0]]), np.concatenate([star[:, 1], path[:, 1]]),
                   np.concatenate([zeros, ones]))
    for _ in range(2):
        check(path)
    g.apply_arrays(np.concatenate([path[:, 0], star[:, 0]]), np.concatenate([path[:, 1], star[:, 1]]),
                   np.concatenate([zeros, ones]))
    check(star)
