"""GPU parity: graphs built from edge lists that are already in device memory.

from_numpy_edges / from_arrow_columns / from_arrow_table / from_scipy_csr
(python/bindings.cpp:21-36) given torch GPU tensors (or DLPack producers on
the device) build the graph on the device.  Each result must equal the
oracle's model of the reference semantics -- Graph(0 or n, directed) plus one
UpdateBatch of additions (last op wins, self-loops dropped, vertex count
max id + 1, version 1 for a non-empty batch) -- and the host path's graph;
invalid input raises the host path's errors before any graph exists.
"""
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


def csr_all(g):
    out = g.export_csr()
    inn = g.export_csr(True) if g.directed else out
    return [np.asarray(x) for x in (*out, *inn)]


def model_from_edges(orc, s, d, directed, n=0):
    mg = orc.ModelGraph(n, directed)
    if len(s):
        mg.apply(np.asarray(s, np.uint32), np.asarray(d, np.uint32), np.ones(len(s), np.uint8))
    return mg


def check_against(g, mg, version):
    assert g.vertex_count == mg.n
    assert g.version == version
    off, adj = g.export_csr()
    moff, madj = mg.csr()
    assert np.array_equal(off, moff) and np.array_equal(adj, madj)
    if mg.directed:
        off, adj = g.export_csr(True)
        moff, madj = mg.csr(reverse=True)
        assert np.array_equal(off, moff) and np.array_equal(adj, madj)


def edge_list(vx, scale, seed):
    s, d = vx.rmat_edges(scale, 8, seed)
    rng = np.random.default_rng(seed)
    k = len(s)
    # duplicates (both orientations) and self-loops on top of the R-MAT list
    dup = rng.integers(0, k, k // 10)
    s = np.concatenate([s, s[dup], d[dup[: len(dup) // 2]]])
    d = np.concatenate([d, d[dup], s[dup[: len(dup) // 2]]])
    loops = rng.integers(0, len(s), 50)
    d[loops] = s[loops]
    return s.astype(np.uint32), d.astype(np.uint32)


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("dtype", ["int32", "uint32", "int64"])
def test_from_numpy_edges_device_equals_oracle_and_host(vx, orc, directed, dtype):
    import torch

    s, d = edge_list(vx, 11, 3)
    e = np.stack([s, d], 1)
    t = torch.from_numpy(e.astype(dtype)).cuda()
    g = vx.from_numpy_edges(t, directed)
    mg = model_from_edges(orc, s, d, directed)
    check_against(g, mg, 1)
    gh = vx.from_numpy_edges(e, directed)
    assert all(np.array_equal(x, y) for x, y in zip(csr_all(g), csr_all(gh)))
    assert g.vertex_count == gh.vertex_count and g.version == gh.version


def test_strided_and_column_inputs(vx, orc):
    import torch

    s, d = edge_list(vx, 10, 5)
    mg = model_from_edges(orc, s, d, False)
    # (N, 2) stored column-major: element (i, j) at j * N + i
    cm = torch.from_numpy(np.stack([s, d]).astype(np.int64)).cuda().t()
    assert cm.stride() == (1, len(s))
    check_against(vx.from_numpy_edges(cm, False), mg, 1)
    # every other row of a larger list
    e = np.stack([s, d], 1)
    big = np.zeros((2 * len(s), 2), np.int32)
    big[::2] = e
    tb = torch.from_numpy(big).cuda()[::2]
    check_against(vx.from_numpy_edges(tb, False), mg, 1)
    # two device columns, and a table of them
    cs = torch.from_numpy(s.view(np.int32)).cuda()
    cd = torch.from_numpy(d.astype(np.int64)).cuda()  # mixed widths are widened
    check_against(vx.from_arrow_columns(cs, cd, False), mg, 1)

    class Table:
        def __init__(self, cols):
            self.cols = cols

        def column(self, name):
            return self.cols[name]

    check_against(vx.from_arrow_table(Table({"a": cs, "b": cd}), "a", "b", False), mg, 1)


def test_empty_and_loop_lists(vx, orc):
    import torch

    g = vx.from_numpy_edges(torch.zeros((0, 2), dtype=torch.int32, device="cuda"))
    assert g.vertex_count == 0 and g.version == 0 and g.edge_count_directed == 0
    loops = torch.tensor([[7, 7], [3, 3]], dtype=torch.int32, device="cuda")
    g = vx.from_numpy_edges(loops)
    gh = vx.from_numpy_edges(loops.cpu().numpy())
    assert (g.vertex_count, g.version, g.edge_count_directed) == (gh.vertex_count, gh.version, 0) == (0, 1, 0)
    # int64 ids in range are narrowed on the device
    one = torch.tensor([[0, 5]], dtype=torch.int64, device="cuda")
    g = vx.from_numpy_edges(one, True)
    assert g.vertex_count == 6 and g.neighbors(0) == [5] and g.version == 1


@pytest.mark.parametrize(
    "bad",
    [
        np.array([[0, 1], [2, -1]], np.int64),
        np.array([[0, 1], [-5, 3]], np.int32),
        np.array([[0, (1 << 31) - 1]], np.int64),
        np.array([[0, 1 << 40]], np.int64),
    ],
)
def test_out_of_range_ids_raise(vx, bad):
    import torch

    with pytest.raises(ValueError, match="vertex ids"):
        vx.from_numpy_edges(torch.from_numpy(bad).cuda())


def test_shape_and_dtype_errors(vx):
    import torch

    with pytest.raises(ValueError, match=r"shape \(N, 2\)"):
        vx.from_numpy_edges(torch.zeros((4, 3), dtype=torch.int32, device="cuda"))
    with pytest.raises(TypeError):
        vx.from_numpy_edges(torch.zeros((4, 2), dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError, match="equal length"):
        vx.from_arrow_columns(torch.zeros(3, dtype=torch.int32, device="cuda"),
                              torch.zeros(4, dtype=torch.int32, device="cuda"))


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("wide", [False, True])
def test_from_scipy_csr_device_equals_host(vx, orc, directed, wide):
    import scipy.sparse as sp
    import torch

    s, d = edge_list(vx, 10, 9)
    n = (1 << 10) + 37  # trailing vertices without arcs keep the count
    m = sp.csr_matrix((np.ones(len(s)), (s, d)), shape=(n, n))
    m.sum_duplicates()
    m.sort_indices()
    host = vx.from_scipy_csr(m, directed)

    class DevCsr:
        pass

    dc = DevCsr()
    it = np.int64 if wide else np.int32
    dc.indptr = torch.from_numpy(m.indptr.astype(it)).cuda()
    dc.indices = torch.from_numpy(m.indices.astype(it)).cuda()
    dc.shape = m.shape
    g = vx.from_scipy_csr(dc, directed)
    assert g.vertex_count == host.vertex_count == n and g.version == host.version == 1
    assert all(np.array_equal(x, y) for x, y in zip(csr_all(g), csr_all(host)))
    # the model: the batch of the entries the host path keeps
    coo = m.tocoo()
    keep = np.ones(len(coo.row), bool) if directed else coo.col >= coo.row
    mg = model_from_edges(orc, coo.row[keep], coo.col[keep], directed, n)
    check_against(g, mg, 1)
    # a torch sparse CSR tensor on the device
    ts = torch.sparse_csr_tensor(dc.indptr.to(torch.int64), dc.indices.to(torch.int64),
                                 torch.ones(dc.indices.numel(), device="cuda"), size=m.shape)
    check_against(vx.from_scipy_csr(ts, directed), mg, 1)


def test_csr_errors_and_empty(vx):
    import torch

    class DevCsr:
        def __init__(self, indptr, indices, n):
            self.indptr = torch.tensor(indptr, dtype=torch.int64, device="cuda")
            self.indices = torch.tensor(indices, dtype=torch.int64, device="cuda")
            self.shape = (n, n)

    with pytest.raises(ValueError, match="invalid CSR offsets"):
        vx.from_scipy_csr(DevCsr([0, 2, 1, 3], [1, 2, 0], 3))
    with pytest.raises(ValueError, match="invalid CSR offsets"):
        vx.from_scipy_csr(DevCsr([0, 2, 2, 9], [1, 2, 0], 3))
    with pytest.raises(ValueError, match="out-of-range vertex id"):
        vx.from_scipy_csr(DevCsr([0, 1, 2, 3], [1, 3, 0], 3))
    with pytest.raises(ValueError, match="indptr length"):
        vx.from_scipy_csr(DevCsr([0, 1, 2], [1, 2], 3))
    # entries past the last offset are not read; an undirected CSR whose only
    # entries lie below the diagonal keeps nothing: version 0, no arcs
    g = vx.from_scipy_csr(DevCsr([0, 0, 1, 2], [0, 1, 99], 3))
    assert g.vertex_count == 3 and g.version == 0 and g.edge_count_directed == 0
    g = vx.from_scipy_csr(DevCsr([0, 0, 1, 2], [0, 1, 99], 3), True)
    assert g.version == 1 and g.neighbors(1) == [0] and g.neighbors(2) == [1]
    g = vx.from_scipy_csr(DevCsr([0, 0, 0], [], 2))
    assert g.vertex_count == 2 and g.version == 0


@pytest.mark.parametrize("form", ["arrays", "pairs"])
def test_host_bulk_load_id_rule(vx, form):
    """Host edge lists are uploaded and checked on the device; ids at or past
    2^31 - 1, which that check does not size, take the host rule of
    dynamic_graph.hpp:502-529: self-loops are dropped whatever their ids,
    the graph grows to max id + 1 over the other edges, and an id that would
    take the vertex count past 2^31 is refused with the graph unchanged."""
    big = (1 << 32) - 1

    def load(g, e):
        e = np.asarray(e, np.uint64).astype(np.uint32)
        if form == "arrays":
            g.bulk_load_arrays(np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1]))
        else:
            g.bulk_load_edges(e)

    g = vx.Graph(5, False)
    load(g, [[0, 1], [1, 2]])
    assert (g.vertex_count, g.version, g.edge_count_directed) == (5, 1, 4)
    # loops only, at ids the device check refuses: nothing kept, no growth
    load(g, [[big, big], [1 << 31, 1 << 31], [3, 3]])
    assert (g.vertex_count, g.version, g.edge_count_directed) == (5, 2, 0)
    # a loop at 2^31 - 1 beside ordinary edges: the edges decide the extent
    load(g, [[0, 6], [(1 << 31) - 1, (1 << 31) - 1], [6, 2]])
    assert (g.vertex_count, g.version, g.edge_count_directed) == (7, 3, 4)
    assert g.neighbors(6) == [0, 2] and g.neighbors(0) == [6]
    # a non-loop id of 2^31 is refused; the graph is left as it was
    with pytest.raises(ValueError, match="vertex ids"):
        load(g, [[0, 1], [2, 1 << 31]])
    assert (g.vertex_count, g.version, g.edge_count_directed) == (7, 3, 4)
    assert g.neighbors(6) == [0, 2]
