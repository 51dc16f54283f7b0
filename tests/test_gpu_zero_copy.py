"""GPU parity: update batches taken from device tensors and DLPack views of
the results.  A batch applied from device memory gives the same graph,
distances and repair diagnostics as the same batch from host arrays, both
equal to the oracle; the DLPack views equal the copied-out results."""
import numpy as np
import pytest

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


def to_dev(bs, bd, ba):
    import torch

    return (torch.from_numpy(bs.view(np.int32)).cuda(), torch.from_numpy(bd.view(np.int32)).cuda(),
            torch.from_numpy(ba.astype(np.uint8)).cuda())


def batches(n, seed, sizes, grow=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in sizes:
        hi = n + grow
        bs = rng.integers(0, hi, k).astype(np.uint32)
        bd = rng.integers(0, hi, k).astype(np.uint32)
        if k >= 4:
            bd[: k // 4] = bs[: k // 4]  # self-loops
        ba = (rng.random(k) < 0.7).astype(np.uint8)
        out.append((bs, bd, ba))
    return out


def csr_all(g):
    out = g.export_csr()
    inn = g.export_csr(True) if g.directed else out
    return [np.asarray(x) for x in (*out, *inn)]


@pytest.mark.parametrize("directed", [False, True])
def test_graph_apply_tensors_equals_host(vx, orc, directed):
    s, d = vx.rmat_edges(11, 16, 4)
    n = 1 << 11
    gh, gd = vx.Graph(n, directed), vx.Graph(n, directed)
    gh.bulk_load_arrays(s, d)
    gd.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    for bs, bd, ba in batches(n, 1, (1, 17, 500, 6000), grow=40):
        gh.apply_arrays(bs, bd, ba)
        gd.apply_tensors(*to_dev(bs, bd, ba))
        mg.apply(bs, bd, ba)
        assert gd.version == gh.version and gd.vertex_count == gh.vertex_count == mg.n
        assert all(np.array_equal(x, y) for x, y in zip(csr_all(gd), csr_all(gh)))
        off, adj = gd.export_csr()
        moff, madj = mg.csr()
        assert np.array_equal(off, moff) and np.array_equal(adj, madj)
    # an all-loop batch grows nothing and changes no row
    import torch

    before = csr_all(gd)
    loop = torch.full((5,), n + 1000, dtype=torch.int32, device="cuda")
    gd.apply_tensors(loop, loop.clone(), torch.ones(5, dtype=torch.bool, device="cuda"))
    assert gd.vertex_count == mg.n and all(np.array_equal(x, y) for x, y in zip(csr_all(gd), before))


def test_bfs_apply_tensors_equals_host(vx, orc):
    s, d = vx.rmat_edges(12, 16, 6)
    n = 1 << 12
    gh, gd = vx.Graph(n, False), vx.Graph(n, False)
    gh.bulk_load_arrays(s, d)
    gd.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    bh, bdv = vx.IncrementalBFS(gh, 3), vx.IncrementalBFS(gd, 3)
    import torch

    side = torch.cuda.Stream()
    for bs, bd, ba in batches(n, 2, (3, 90, 2500, 400, 12000), grow=25):
        bh.apply_arrays(bs, bd, ba)
        with torch.cuda.stream(side):  # the batch is written on another stream
            t = to_dev(bs, bd, ba)
            bdv.apply_tensors(*t)
        mg.apply(bs, bd, ba)
        off, adj = mg.csr()
        want, _ = orc.bfs(mg.n, off, adj, 3)
        assert np.array_equal(bdv.distances_numpy(), want)
        assert np.array_equal(bh.distances_numpy(), want)
        for attr in ("reachable_count", "last_deletion_candidates", "last_affected_vertices",
                     "last_used_full_recompute"):
            assert getattr(bdv, attr) == getattr(bh, attr), attr
        view = bdv.distances_tensor()
        assert view.dtype == torch.uint32 and view.device.type == "cuda"
        assert np.array_equal(view.cpu().numpy(), want)


def test_dlpack_views(vx, orc):
    import torch

    s, d = vx.rmat_edges(10, 16, 8)
    g = vx.Graph(1 << 10, True)
    g.bulk_load_arrays(s, d)
    bfs = vx.IncrementalBFS(g, 0)
    v = bfs.distances_tensor()
    assert v.shape == (g.vertex_count,)
    assert np.array_equal(v.cpu().numpy(), bfs.distances_numpy())
    del bfs  # the view keeps its owner alive
    host = v.cpu().numpy()
    assert host[0] == 0 and int((host != 0xFFFFFFFF).sum()) > 1
    pr = vx.IncrementalPageRank(g)
    t = pr.values_tensor()
    assert t.dtype == torch.float64 and np.array_equal(t.cpu().numpy(), np.asarray(pr.values))
    pr32 = vx.IncrementalPageRank(g, 0.85, 32)
    t32 = pr32.values_tensor()
    assert t32.dtype == torch.float32
    assert np.abs(t32.cpu().numpy().astype(np.float64) - np.asarray(pr32.values)).max() == 0.0
    cap = pr.values_dlpack()  # an unconsumed capsule frees its view
    del cap


def test_device_batch_errors(vx):
    import torch

    g = vx.Graph(8, False)
    a = torch.zeros(3, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError):
        g.apply_tensors(a, a[:2], torch.zeros(3, dtype=torch.uint8, device="cuda"))
    with pytest.raises(TypeError):
        g.apply_tensors(a.cpu(), a, torch.zeros(3, dtype=torch.uint8, device="cuda"))
    with pytest.raises(TypeError):
        g.apply_tensors(a.float(), a, torch.zeros(3, dtype=torch.uint8, device="cuda"))
    e = torch.zeros(0, dtype=torch.int32, device="cuda")
    v = g.version
    g.apply_tensors(e, e, torch.zeros(0, dtype=torch.uint8, device="cuda"))
    assert g.version == v


def _delay(torch, stream):
    """Queues a few tens of milliseconds of device work on `stream`: a bounded
    spin where torch has one, else a chain of matrix products."""
    with torch.cuda.stream(stream):
        if hasattr(torch.cuda, "_sleep"):
            torch.cuda._sleep(100_000_000)
        else:
            a = torch.randn(2048, 2048, device="cuda")
            for _ in range(40):
                a = a @ a
                a = a / a.norm()


def test_view_reader_on_a_side_stream_is_waited_for(vx, orc):
    """A tensor view read on a side stream behind slow work, then an apply
    that changes the distances: the apply waits on the device for the
    reader, so the reader copies the distances the view had."""
    import torch

    s, d = vx.rmat_edges(12, 8, 3)
    n = 1 << 12
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    bfs = vx.IncrementalBFS(g, 0)
    before = bfs.distances_numpy()
    # a batch that moves many distances: every vertex of the component gets an arc to the root
    reached = np.flatnonzero(before != 0xFFFFFFFF).astype(np.uint32)
    bs = np.zeros(reached.size, np.uint32)
    bd = reached
    ba = np.ones(reached.size, np.uint8)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        view = bfs.distances_tensor()  # read on `side`
    _delay(torch, side)
    with torch.cuda.stream(side):
        copy = view.clone()
    bfs.apply_arrays(bs, bd, ba)  # host order: after the reader was queued
    side.synchronize()
    assert np.array_equal(copy.cpu().numpy(), before)
    mg.apply(bs, bd, ba)
    off, adj = mg.csr()
    want, _ = orc.bfs(mg.n, off, adj, 0)
    assert not np.array_equal(want, before)
    assert np.array_equal(bfs.distances_numpy(), want)
    assert np.array_equal(view.cpu().numpy(), want)  # same memory, current values
    del view, copy


def test_views_survive_array_growth(vx, orc):
    """A batch that grows the vertex count past the arrays' capacity while
    views are held: the views keep the old allocations (old contents), the
    objects move to new ones, and releasing the views frees the old ones."""
    import torch

    s, d = vx.rmat_edges(10, 8, 5)
    n = 1 << 10
    g = vx.Graph(n, True)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, True)
    mg.bulk_load_edges(np.stack([s, d], 1))
    bfs = vx.IncrementalBFS(g, 1)
    pr = vx.IncrementalPageRank(g)
    dv, rv = bfs.distances_tensor(), pr.values_tensor()
    d0, r0 = bfs.distances_numpy(), np.asarray(pr.values)
    big = 40 * n  # well past the vertex capacity
    bs = np.array([1, big - 1, 5], np.uint32)
    bd = np.array([big - 1, 2, 6], np.uint32)
    ba = np.ones(3, np.uint8)
    bfs.apply_arrays(bs, bd, ba)
    pr.apply(vx.UpdateBatch.from_arrays(bs, bd, ba))  # the vertex count changed: a full solve
    mg.apply(bs, bd, ba)
    torch.cuda.synchronize()
    assert np.array_equal(dv.cpu().numpy(), d0)
    assert np.array_equal(rv.cpu().numpy(), r0)
    off, adj = mg.csr()
    want, _ = orc.bfs(mg.n, off, adj, 1)
    assert np.array_equal(bfs.distances_numpy(), want)
    ref, _ = orc.pagerank_full(mg)
    assert np.abs(np.asarray(pr.values) - ref).max() < 1e-10
    del dv, rv  # the old allocations go now
    bfs.recompute()
    assert np.array_equal(bfs.distances_numpy(), want)
    v2 = bfs.distances_tensor()
    assert np.array_equal(v2.cpu().numpy(), want)


def test_batch_from_a_slow_producer_stream(vx, orc):
    """A device batch written by device work on a side stream queued behind
    slow work: applying it with that stream current reads the finished ids
    (the engine waits for the producer), not the zeros the buffers held."""
    import torch

    s, d = vx.rmat_edges(11, 16, 9)
    n = 1 << 11
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    bfs = vx.IncrementalBFS(g, 0)
    rng = np.random.default_rng(4)
    bs = rng.integers(1, n, 3000).astype(np.uint32)
    bd = rng.integers(1, n, 3000).astype(np.uint32)
    ba = np.ones(3000, np.uint8)
    ba[:1000] = 0
    host = [torch.from_numpy(x.view(np.int32) if x.dtype == np.uint32 else x).cuda() for x in (bs, bd, ba)]
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    src = torch.zeros(3000, dtype=torch.int32, device="cuda")
    dst = torch.zeros(3000, dtype=torch.int32, device="cuda")
    add = torch.zeros(3000, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    _delay(torch, side)
    with torch.cuda.stream(side):
        src.copy_(host[0])
        dst.copy_(host[1])
        add.copy_(host[2])
        bfs.apply_tensors(src, dst, add)
    mg.apply(bs, bd, ba)
    off, adj = mg.csr()
    want, _ = orc.bfs(mg.n, off, adj, 0)
    assert np.array_equal(bfs.distances_numpy(), want)
