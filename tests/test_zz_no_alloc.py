"""The allocation-free BFS entry point (vgx_bfs_distances_noalloc), called
through the C ABI: its distances equal those of the allocating entry point
(vgx_bfs_create, through the bindings) bit for bit, and the oracle's, on
calls in a row where the second graph is smaller (the scratch array keeps
the previous call's contents), both directions; the size limit and a short
output buffer are refused with an error."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = 0xFFFFFFFF


@pytest.fixture(scope="module")
def env():
    import torch

    import velographx_amd as vx

    assert vx.device_count() > 0, "no gfx950 device visible"
    lib = ctypes.CDLL(os.path.join(os.path.dirname(vx.__file__), "libvgx.so"))
    f = lib.vgx_bfs_distances_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64,
                  ctypes.POINTER(ctypes.c_uint64)]
    lib.vgx_last_error.restype = ctypes.c_char_p
    return torch, vx, lib


def noalloc(env, g, source, cap=None):
    torch, _, lib = env
    n = g.vertex_count
    out = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # the fill runs on torch's stream, the BFS on the graph's
    reach = ctypes.c_uint64(123)
    rc = lib.vgx_bfs_distances_noalloc(g._handle, source, out.data_ptr(), n if cap is None else cap,
                                       ctypes.byref(reach))
    return rc, out[:n].cpu().numpy().view(np.uint32), reach.value, lib.vgx_last_error().decode()


def graph(vx, orc, scale, seed, directed):
    s, d = vx.rmat_edges(scale, 16, seed)
    g = vx.Graph(1 << scale, directed)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(1 << scale, directed)
    mg.bulk_load_edges(np.stack([s, d], 1))
    return g, mg


@pytest.mark.parametrize("directed", [False, True])
def test_noalloc_equals_allocating_path_second_call_smaller(env, orc, directed):
    _, vx, _ = env
    big, mbig = graph(vx, orc, 16, 3, directed)
    small, msmall = graph(vx, orc, 11, 4, directed)
    hub = int(np.argmax(big.degrees()))
    for g, mg, roots in ((big, mbig, [hub, 5, 1000]), (small, msmall, [0, 77]), (big, mbig, [hub])):
        for r in roots:
            rc, dist, reach, err = noalloc(env, g, r)
            assert rc == 0, err
            want = vx.IncrementalBFS(g, r).distances_numpy()
            assert np.array_equal(dist, want), (g.vertex_count, r)
            full, reached = orc.model_bfs(mg, r)
            assert np.array_equal(dist, full)
            assert reach == reached


def test_noalloc_limits_and_out_of_range_source(env, orc):
    _, vx, _ = env
    n = 1 << 20
    g = vx.Graph(n, False)
    s, d = vx.rmat_edges(20, 4, 2)
    g.bulk_load_arrays(s, d)
    assert g.vertex_count == n
    hub = int(np.argmax(g.degrees()))
    rc, dist, reach, err = noalloc(env, g, hub)
    assert rc == 0, err
    assert np.array_equal(dist, vx.IncrementalBFS(g, hub).distances_numpy())
    rc, dist, reach, err = noalloc(env, g, n + 3)
    assert rc == 0 and reach == 0 and np.all(dist == INF)
    rc, _, _, err = noalloc(env, g, hub, cap=n - 1)
    assert rc == 1 and "fewer entries" in err
    over = vx.Graph(n + 1, False)
    over.add_edge(0, n)
    rc, _, _, err = noalloc(env, over, 0)
    assert rc == 1 and "at most 1048576 vertices" in err
