"""GPU parity: localized residual-push PageRank repair (apply_push) against
the FP64 oracle.  After every batch the reported ranks must be within the
bound the residual gives (L1 error <= l1_tol / (1 - d), plus FP32 rounding
at precision 32), and in FP64 the residual the object reports must be the
true residual of its ranks.  Covered: pure pushes (no fallback), the
full-sweep fallback, rows that become dangling or stop being dangling (the
uniform part of the residual), vertex growth, undirected graphs, and
interleaving with the other solves (which invalidate the kept residual)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = 0.85


@pytest.fixture(scope="module")
def vx():
    import velographx_amd as vx

    assert vx.device_count() > 0, "no gfx950 device visible"
    return vx


def graphs(vx, orc, scale, seed, directed):
    s, d = vx.rmat_edges(scale, 8, seed)
    g = vx.Graph(1 << scale, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << scale, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    return g, mg


def true_residual(mg, x):
    """b + d M x - x in FP64 for the model graph."""
    n = mg.n
    off, adj = mg.csr()
    deg = np.diff(off).astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(off).astype(np.int64))
    contrib = np.where(deg > 0, x / np.where(deg > 0, deg, 1), 0.0)
    inc = np.bincount(adj, weights=contrib[rows], minlength=n)
    return (1 - D) / n + D * x[deg == 0].sum() / n + D * inc - x


def check(orc, mg, pr, tol, precision):
    ref, _ = orc.pagerank_full(mg, D, 1000, 1e-15)
    x = pr.values_numpy()
    err = np.abs(x - ref)
    floor = 0.0 if precision == 64 else 4e-7
    assert err.sum() <= tol / (1 - D) + floor, (err.sum(), tol)
    return x


def stream(wl, mg, seed):
    u = (mg.keys >> np.uint64(32)).astype(np.uint32)
    v = (mg.keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    if not mg.directed:
        keep = u < v
        u, v = u[keep], v[keep]
    return wl.BatchStream(u, v, mg.n, mg.directed, seed=seed)


@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 1e-7)])
@pytest.mark.parametrize("directed", [True, False])
def test_push_without_fallback(vx, orc, precision, tol, directed):
    from velographx_amd import workloads as wl

    g, mg = graphs(vx, orc, 12, 3, directed)
    pr = vx.IncrementalPageRank(g, D, precision)
    st = stream(wl, mg, 5)
    for b in range(8):
        bs, bd, ba = st.next(20, 0.8)
        r = pr.apply_push_arrays(bs, bd, ba, l1_tol=tol, max_push_fraction=1e9)
        mg.apply(bs, bd, ba)
        assert not r["used_fallback"] and r["rounds"] >= 1, r
        assert r["residual_l1"] <= tol
        x = check(orc, mg, pr, tol, precision)
        if precision == 64:
            assert abs(np.abs(true_residual(mg, x)).sum() - r["residual_l1"]) <= 1e-12, r
        assert not pr.last_used_full_recompute


@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 1e-7)])
def test_push_fallback_to_sweeps(vx, orc, precision, tol):
    from velographx_amd import workloads as wl

    g, mg = graphs(vx, orc, 12, 4, True)
    pr = vx.IncrementalPageRank(g, D, precision)
    st = stream(wl, mg, 6)
    for b in range(4):
        bs, bd, ba = st.next(200, 0.8)
        # b = 0: no push budget at all; then small budgets
        r = pr.apply_push_arrays(bs, bd, ba, l1_tol=tol, max_push_fraction=0.0 if b == 0 else 0.05)
        mg.apply(bs, bd, ba)
        assert r["residual_l1"] <= tol
        if b == 0:
            assert r["used_fallback"] and r["fallback_sweeps"] >= 1
        x = check(orc, mg, pr, tol, precision)
        if precision == 64:
            assert np.abs(true_residual(mg, x)).sum() <= tol * (1 + 1e-6)


def test_push_dangling_transitions(vx, orc):
    """Rows that lose their last out-arc become dangling (their share goes to
    every vertex) and dangling rows that gain one stop being so; the uniform
    part of the residual is folded into the ranks exactly."""
    n = 300
    rng = np.random.default_rng(2)
    edges = np.unique(rng.integers(0, n, size=(1200, 2)), axis=0)
    edges = edges[edges[:, 0] != edges[:, 1]].astype(np.uint32)
    g = vx.Graph(n, True)
    g.bulk_load_edges(edges)
    mg = orc.ModelGraph(n, True)
    mg.bulk_load_edges(edges)
    pr = vx.IncrementalPageRank(g, D, 64)
    out = np.bincount(edges[:, 0], minlength=n)
    for b in range(6):
        off, adj = mg.csr()
        deg = np.diff(off)
        live = np.nonzero(deg > 0)[0]
        dang = np.nonzero(deg == 0)[0]
        u = int(rng.choice(live))
        row = adj[off[u]:off[u + 1]]
        src = [u] * row.size  # remove every out-arc of u: u becomes dangling
        dst = list(row)
        add = [0] * row.size
        if dang.size:  # and give a dangling vertex an out-arc
            w = int(rng.choice(dang))
            src.append(w)
            dst.append((w + 1) % n)
            add.append(1)
        r = pr.apply_push_arrays(np.array(src, np.uint32), np.array(dst, np.uint32), np.array(add, np.uint8),
                                 l1_tol=1e-11, max_push_fraction=1e9)
        mg.apply(src, dst, add)
        assert not r["used_fallback"]
        x = check(orc, mg, pr, 1e-11, 64)
        assert abs(np.abs(true_residual(mg, x)).sum() - r["residual_l1"]) <= 1e-13
    assert out.size == n


def test_push_vertex_growth_and_interleaving(vx, orc):
    """A batch that grows the vertex range recomputes; solve_l1, recompute
    and apply replace the ranks, after which the kept residual is rebuilt
    before the next push."""
    from velographx_amd import workloads as wl

    g, mg = graphs(vx, orc, 11, 7, True)
    pr = vx.IncrementalPageRank(g, D, 64)
    st = stream(wl, mg, 8)
    n = mg.n
    r = pr.apply_push_arrays(np.array([3, n + 10], np.uint32), np.array([n + 10, 5], np.uint32),
                             np.array([1, 1], np.uint8), l1_tol=1e-10)
    mg.apply([3, n + 10], [n + 10, 5], [1, 1])
    assert r["used_fallback"]
    check(orc, mg, pr, 1e-10, 64)
    for step in range(6):
        bs, bd, ba = st.next(15, 0.8)
        if step == 1:
            pr.solve_l1(200, 1e-12, True)
        if step == 3:
            pr.recompute()
        if step == 4:
            b = vx.UpdateBatch()
            b.add(1, 2)
            pr.apply(b)
            mg.apply([1], [2], [1])
        r = pr.apply_push_arrays(bs, bd, ba, l1_tol=1e-10, max_push_fraction=1e9)
        mg.apply(bs, bd, ba)
        x = check(orc, mg, pr, 1e-10, 64)
        assert abs(np.abs(true_residual(mg, x)).sum() - r["residual_l1"]) <= 1e-12, step


def test_push_update_batch_object(vx, orc):
    g = vx.Graph(6, True)
    g.bulk_load_edges(np.array([[0, 1], [1, 2], [2, 0], [3, 2]], dtype=np.uint32))
    mg = orc.ModelGraph(6, True)
    mg.bulk_load_edges([(0, 1), (1, 2), (2, 0), (3, 2)])
    pr = vx.IncrementalPageRank(g, D, 64)
    b = vx.UpdateBatch()
    b.add(4, 0)
    b.remove(3, 2)
    b.add(3, 2)  # the last op on an arc wins
    r = pr.apply_push(b, l1_tol=1e-13, max_push_fraction=1e9)
    mg.apply([4, 3, 3], [0, 2, 2], [1, 0, 1])
    assert not r["used_fallback"]
    check(orc, mg, pr, 1e-13, 64)


def test_push_selector_skips_and_probes(vx, orc):
    """A batch whose first push round is over budget marks its size: the next
    batches of at least half that size go to the warm sweeps directly
    (skipped_push), every 8th tries the pushes again, and a push phase that
    finishes clears the mark.  Every batch stays within the residual's bound."""
    from velographx_amd import workloads as wl

    g, mg = graphs(vx, orc, 12, 9, True)
    pr = vx.IncrementalPageRank(g, D, 64)
    st = stream(wl, mg, 10)
    tol = 1e-10
    seen = []
    for b in range(11):
        bs, bd, ba = st.next(40, 0.8)
        r = pr.apply_push_arrays(bs, bd, ba, l1_tol=tol, max_push_fraction=0.0 if b == 0 else 1e9)
        mg.apply(bs, bd, ba)
        seen.append((r["used_fallback"], r["skipped_push"]))
        x = check(orc, mg, pr, tol, 64)
        assert abs(np.abs(true_residual(mg, x)).sum() - r["residual_l1"]) <= 1e-12, b
    assert seen[0] == (True, False)
    assert all(seen[b] == (True, True) for b in range(1, 8)), seen
    assert seen[8] == (False, False) and seen[9] == (False, False), seen
    # a smaller batch (below half the marked size) is never skipped
    bs, bd, ba = st.next(40, 0.8)
    pr.apply_push_arrays(bs, bd, ba, l1_tol=tol, max_push_fraction=0.0)
    mg.apply(bs, bd, ba)
    bs, bd, ba = st.next(10, 0.8)
    r = pr.apply_push_arrays(bs, bd, ba, l1_tol=tol, max_push_fraction=1e9)
    mg.apply(bs, bd, ba)
    assert not r["skipped_push"] and not r["used_fallback"]
    check(orc, mg, pr, tol, 64)
