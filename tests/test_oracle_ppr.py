"""Pins the FP64 weighted personalized PageRank oracle (oracle/ppr.py) by
what the mathematics fixes; the reference has no weighted or personalized
PageRank to compare with."""
import numpy as np
import pytest

from oracle import ppr


def random_weighted(n, m, seed, dangling=True):
    rng = np.random.default_rng(seed)
    keys = np.unique(rng.integers(0, n, size=(m, 2)) @ np.array([n, 1]))
    src, dst = keys // n, keys % n
    keep = src != dst
    src, dst = src[keep], dst[keep]
    if not dangling:  # give every vertex an out-arc
        extra = np.setdiff1d(np.arange(n), src)
        src = np.concatenate([src, extra])
        dst = np.concatenate([dst, (extra + 1) % n])
    w = 1.0 - rng.random(src.size)  # uniform on (0, 1]
    return src, dst, w


def one_hot(n, cols, seed):
    S = np.zeros((n, len(cols)))
    S[cols, np.arange(len(cols))] = 1.0
    return S


def test_zero_damping_is_identity():
    src, dst, w = random_weighted(30, 120, 1)
    S = np.random.default_rng(2).random((30, 3))
    assert np.allclose(ppr.forward(30, src, dst, w, S, 0.0), S, atol=0, rtol=0)
    assert np.allclose(ppr.forward(30, src, dst, w, S, 0.0, iterations=5), S, atol=0, rtol=0)


def test_cycle_closed_form_and_orientation():
    """On the directed cycle 0->1->...->m-1->0 a one-hot seed at 0 gives
    X_k = (1-d) d^k / (1 - d^m); weights scale out, since every row has one
    arc.  The reversed orientation (a transposed P) fails the same check."""
    m, d = 7, 0.85
    src = np.arange(m)
    dst = (src + 1) % m
    w = np.linspace(0.2, 1.0, m)
    S = one_hot(m, [0], 0)
    want = (1 - d) * d ** np.arange(m) / (1 - d**m)
    for dense in (True, False):
        got = ppr.forward(m, src, dst, w, S, d, dense=dense)[:, 0]
        assert np.allclose(got, want, rtol=1e-13, atol=0)
    wrong = ppr.forward(m, dst, src, w, S, d)[:, 0]
    assert not np.allclose(wrong, want, rtol=1e-6)


def test_weights_split_mass_by_ratio():
    """Vertex 0 with arcs to 1 (weight 1) and 2 (weight 3), d = 0.5: the
    single pass gives X_1 = 0.5*0.5*0.25, X_2 = 0.5*0.5*0.75."""
    src, dst, w = np.array([0, 0]), np.array([1, 2]), np.array([1.0, 3.0])
    X = ppr.forward(3, src, dst, w, one_hot(3, [0], 0), 0.5)[:, 0]
    assert np.allclose(X, [0.5, 0.0625, 0.1875], rtol=1e-15, atol=0)


def test_solve_equals_jacobi_limit_and_dense():
    src, dst, w = random_weighted(60, 400, 3)
    S = one_hot(60, [0, 5, 17, 59], 0)
    exact = ppr.forward(60, src, dst, w, S, 0.85)
    dense = ppr.forward(60, src, dst, w, S, 0.85, dense=True)
    it = ppr.forward(60, src, dst, w, S, 0.85, iterations=300)
    assert np.allclose(exact, dense, rtol=1e-12, atol=1e-15)
    assert np.allclose(exact, it, rtol=1e-12, atol=1e-15)
    # substochastic P: total mass never exceeds the teleport mass
    assert np.all(exact.sum(0) <= 1.0 + 1e-12)


def test_no_dangling_conserves_mass():
    src, dst, w = random_weighted(50, 300, 4, dangling=False)
    S = one_hot(50, [1, 2, 3], 0)
    X = ppr.forward(50, src, dst, w, S, 0.85)
    assert np.allclose(X.sum(0), 1.0, rtol=1e-12)


@pytest.mark.parametrize("seed", [5, 6])
def test_adjoint_gradients_equal_finite_differences(seed):
    """dLoss/dw and dLoss/dS of Loss = sum(C * X) against FP64 central
    differences of the exact solve, with dangling rows present."""
    n = 40
    src, dst, w = random_weighted(n, 200, seed)
    rng = np.random.default_rng(seed + 10)
    S = one_hot(n, rng.integers(0, n, 3), 0)
    C = rng.standard_normal((n, 3))
    X, gw, gs = ppr.gradients(n, src, dst, w, S, C, 0.85, dense=True)
    fd = ppr.finite_difference_w(n, src, dst, w, S, C, 0.85, np.arange(src.size), eps=1e-6)
    assert np.allclose(gw, fd, rtol=1e-6, atol=1e-9)
    # teleport gradient: Loss is linear in S, so one difference per entry is exact
    for (v, k) in [(0, 0), (7, 1), (n - 1, 2)]:
        Sp = S.copy()
        Sp[v, k] += 1.0
        diff = ppr.loss(n, src, dst, w, Sp, C, 0.85) - ppr.loss(n, src, dst, w, S, C, 0.85)
        assert np.isclose(gs[v, k], diff, rtol=1e-10, atol=1e-13)


def test_gradient_planted_mistake_is_caught():
    """Dropping the normalisation term of dw (a plausible slip) breaks the
    agreement with finite differences."""
    n = 30
    src, dst, w = random_weighted(n, 150, 8)
    S = one_hot(n, [3, 4], 0)
    C = np.random.default_rng(1).standard_normal((n, 2))
    X = ppr.forward(n, src, dst, w, S, 0.85)
    L = ppr.adjoint(n, src, dst, w, C, 0.85)
    _, W = ppr.transition(n, src, dst, w)
    slipped = 0.85 * np.einsum("ek,ek->e", X[src], L[dst]) / W[src]
    fd = ppr.finite_difference_w(n, src, dst, w, S, C, 0.85, np.arange(src.size))
    assert not np.allclose(slipped, fd, rtol=1e-3, atol=1e-6)


D = 0.85


def _small_problem(seed, n=200, m=1500, K=4):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    key = np.unique(src[keep] * n + dst[keep])
    src, dst = key // n, key % n  # out-CSR order: by source, then destination
    w = 1.0 - rng.random(src.size)
    S = np.zeros((n, K))
    S[rng.integers(0, n, K), np.arange(K)] = 1.0
    C = rng.standard_normal((n, K))
    return n, src, dst, w, S, C, rng


@pytest.mark.parametrize("seed", [1, 2])
def test_c_sweeps_equal_numpy_sweeps(seed):
    """orc_ppr_sweeps over the in-rows and the out-rows gives the forward and
    adjoint Jacobi sweeps of the numpy oracle."""
    n, src, dst, w, S, C, _ = _small_problem(seed)
    for it in (0, 1, 7, 60):
        assert np.allclose(ppr.forward_c(n, src, dst, w, S, D, it), ppr.forward(n, src, dst, w, S, D, it),
                           rtol=0, atol=1e-15)
        W = np.bincount(src, weights=w, minlength=n)
        p = w / W[src]
        L = ppr.sweeps_c(n, ppr.csr_rows(n, src, dst, p), C, D, it)
        assert np.allclose(L, ppr.adjoint(n, src, dst, w, C, D, it), rtol=0, atol=1e-13)


@pytest.mark.parametrize("seed", [3, 4])
def test_rank_one_finite_differences_equal_dense_ones(seed):
    """The rank-one evaluation of the perturbed losses gives the central
    differences of the exact dense solve (to the latter's cancellation
    error), and both approach the exact adjoint gradient; the same
    differences on the transposed graph do not."""
    n, src, dst, w, S, C, rng = _small_problem(seed)
    arcs = rng.choice(src.size, 40, replace=False)
    dense = ppr.finite_difference_w(n, src, dst, w, S, C, D, arcs, eps=1e-6)
    r1, X, L = ppr.finite_difference_w_rank1(n, src, dst, w, S, C, D, arcs, eps=1e-5, iterations=300)
    scale = np.max(np.abs(dense))
    assert np.max(np.abs(r1 - dense)) <= 1e-6 * scale
    _, gw, gs = ppr.gradients(n, src, dst, w, S, C, D, dense=True)
    assert np.max(np.abs(gw[arcs] - r1)) <= 1e-8 * np.max(np.abs(gw[arcs]))
    assert np.max(np.abs(X - ppr.forward(n, src, dst, w, S, D, None, dense=True))) <= 1e-12
    assert np.max(np.abs((1 - D) * L - gs)) <= 1e-12
    # planted mistake: the transposed graph's differences are different numbers
    key = np.argsort(dst * n + src)
    t_src, t_dst, t_w = dst[key], src[key], w[key]
    t_arcs = np.argsort(key)[arcs]
    wrong, _, _ = ppr.finite_difference_w_rank1(n, t_src, t_dst, t_w, S, C, D, t_arcs, eps=1e-5, iterations=300)
    assert np.max(np.abs(wrong - dense)) > 1e-3 * scale
