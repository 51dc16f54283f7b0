"""GPU parity: weighted personalized PageRank forward and adjoint kernels
behind torch.autograd.Function, against the FP64 oracle (oracle/ppr.py):
an exact dense-solve reference gradient, FP64 central finite differences on
200 sampled arcs, and the same truncated iteration in FP64 at a larger
scale.  Tolerances are FP32 accumulation bounds, stated per check."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = 0.85


@pytest.fixture(scope="module")
def env():
    import torch

    import velographx_amd as vx
    from velographx_amd import autograd as ag

    assert vx.device_count() > 0, "no gfx950 device visible"
    return torch, vx, ag


def weighted_rmat(vx, ag, scale, seed):
    s, d = vx.rmat_edges(scale, 16, seed)
    g = vx.Graph(1 << scale, True)
    g.bulk_load_arrays(s, d)
    src, dst = ag.arc_list(g)
    w = (1.0 - np.random.default_rng(seed).random(src.size)).astype(np.float32)  # uniform on (0, 1]
    return g, src.astype(np.int64), dst.astype(np.int64), w


def problem(n, K, seed):
    rng = np.random.default_rng(seed)
    S = np.zeros((n, K), np.float32)
    S[rng.integers(0, n, K), np.arange(K)] = 1.0
    C = rng.standard_normal((n, K)).astype(np.float32)
    return S, C


def run_gpu(torch, ag, g, w, S, C, iterations):
    wt = torch.tensor(w, device="cuda", requires_grad=True)
    st = torch.tensor(S, device="cuda", requires_grad=True)
    x = ag.personalized_pagerank(g, wt, st, D, iterations)
    (x * torch.tensor(C, device="cuda")).sum().backward()
    torch.cuda.synchronize()
    return x.detach().cpu().numpy(), wt.grad.cpu().numpy(), st.grad.cpu().numpy()


def rel_inf(got, want):
    return float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-30))


def test_dense_solve_reference_gradient_s12(env):
    """R-MAT scale 12, 32 one-hot teleport columns, random linear loss:
    X, dLoss/dw and dLoss/dS against the exact FP64 dense solve.  120 sweeps
    leave a truncation of d^121 < 3e-9; the FP32 sums are bounded at 1e-5 of
    each quantity's largest entry."""
    torch, vx, ag = env
    g, src, dst, w = weighted_rmat(vx, ag, 12, 1)
    from oracle import ppr

    n = g.vertex_count
    S, C = problem(n, 32, 2)
    x, gw, gs = run_gpu(torch, ag, g, w, S, C, 120)
    X, rw, rs = ppr.gradients(n, src, dst, w.astype(np.float64), S, C, D, dense=True)
    assert np.max(np.abs(x - X)) <= 1e-6
    assert rel_inf(gs, rs) <= 1e-5, rel_inf(gs, rs)
    assert rel_inf(gw, rw) <= 1e-5, rel_inf(gw, rw)


def test_finite_differences_200_arcs_s10(env):
    """FP64 central differences of the exact solve on 200 sampled arcs of a
    weighted R-MAT scale 10 graph against the GPU adjoint's dLoss/dw."""
    torch, vx, ag = env
    g, src, dst, w = weighted_rmat(vx, ag, 10, 3)
    from oracle import ppr

    n = g.vertex_count
    S, C = problem(n, 32, 4)
    _, gw, _ = run_gpu(torch, ag, g, w, S, C, 120)
    arcs = np.random.default_rng(5).choice(src.size, size=200, replace=False)
    fd = ppr.finite_difference_w(n, src, dst, w.astype(np.float64), S, C, D, arcs, eps=1e-6)
    scale = np.max(np.abs(fd))
    assert np.max(np.abs(gw[arcs] - fd)) <= 1e-5 * scale, np.max(np.abs(gw[arcs] - fd)) / scale


def test_same_iteration_fp64_s16(env):
    """Scale 16 with the oracle running the same 60 sweeps in FP64: the
    difference is FP32 rounding only."""
    torch, vx, ag = env
    g, src, dst, w = weighted_rmat(vx, ag, 16, 7)
    from oracle import ppr

    n = g.vertex_count
    S, C = problem(n, 32, 8)
    x, gw, gs = run_gpu(torch, ag, g, w, S, C, 60)
    X, rw, rs = ppr.gradients(n, src, dst, w.astype(np.float64), S, C, D, iterations=60)
    assert np.max(np.abs(x - X)) <= 1e-6
    assert rel_inf(gs, rs) <= 1e-5, rel_inf(gs, rs)
    assert rel_inf(gw, rw) <= 1e-5, rel_inf(gw, rw)


def test_more_than_32_columns_and_undirected(env):
    """K = 40 runs as column blocks of 32 and 8; an undirected graph has one
    weight per arc direction."""
    torch, vx, ag = env
    from oracle import ppr

    s, d = vx.rmat_edges(9, 8, 11)
    g = vx.Graph(1 << 9, False)
    g.bulk_load_arrays(s, d)
    src, dst = (a.astype(np.int64) for a in ag.arc_list(g))
    w = (1.0 - np.random.default_rng(1).random(src.size)).astype(np.float32)
    n = g.vertex_count
    S, C = problem(n, 40, 12)
    x, gw, gs = run_gpu(torch, ag, g, w, S, C, 150)
    X, rw, rs = ppr.gradients(n, src, dst, w.astype(np.float64), S, C, D, dense=True)
    assert np.max(np.abs(x - X)) <= 1e-6
    assert rel_inf(gs, rs) <= 1e-5
    assert rel_inf(gw, rw) <= 1e-5


def test_dangling_and_isolated_vertices(env):
    """Vertices without out-arcs send no mass; their teleport stays put."""
    torch, vx, ag = env
    from oracle import ppr

    edges = np.array([[0, 1], [0, 2], [1, 2], [2, 0], [3, 2]], dtype=np.uint32)
    g = vx.Graph(6, True)
    g.bulk_load_edges(edges)
    src, dst = (a.astype(np.int64) for a in ag.arc_list(g))
    w = np.array([0.5, 1.0, 0.25, 1.0, 0.75], np.float32)
    S = np.eye(6, 3, -2, dtype=np.float32)  # seeds at 2, 3, 4
    C = np.random.default_rng(0).standard_normal((6, 3)).astype(np.float32)
    x, gw, gs = run_gpu(torch, ag, g, w, S, C, 200)
    X, rw, rs = ppr.gradients(6, src, dst, w.astype(np.float64), S, C, D, dense=True)
    assert np.allclose(x, X, atol=1e-6)
    assert np.allclose(gw, rw, atol=1e-5)
    assert np.allclose(gs, rs, atol=1e-5)
    assert x[4, 2] == pytest.approx(1 - D) and x[5].sum() == 0.0


def test_errors(env):
    torch, vx, ag = env
    g = vx.Graph(0, True)
    g.bulk_load_edges(np.array([[0, 1], [1, 2]], dtype=np.uint32))
    w = torch.ones(2, device="cuda", requires_grad=True)
    S = torch.eye(3, 1, device="cuda")
    with pytest.raises(ValueError):
        ag.personalized_pagerank(g, w.detach().cpu(), S)
    with pytest.raises(ValueError):
        ag.personalized_pagerank(g, torch.ones(3, device="cuda"), S)
    x = ag.personalized_pagerank(g, w, S)
    g.add_edge(2, 0)
    with pytest.raises(RuntimeError):
        x.sum().backward()
