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
