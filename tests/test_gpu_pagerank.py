"""GPU parity for PageRank: FP64 against the CPU oracle (which reproduces the
reference's full_solve / localized apply), FP32 against the FP64 oracle with
the 1e-6 L1 / max-abs bound."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vx():
    import velographx_amd as vx

    assert vx.device_count() > 0, "no gfx950 device visible"
    return vx


def batch_of(vx, ops):
    b = vx.UpdateBatch()
    for u, v, a in ops:
        (b.add if a else b.remove)(u, v)
    return b


def twins(vx, orc, n, directed, edges):
    g = vx.Graph(n, directed)
    m = orc.ModelGraph(n, directed)
    if edges:
        g.bulk_load_edges(np.array(edges, dtype=np.uint32))
        m.bulk_load_edges(edges)
    return g, m


@pytest.mark.parametrize("directed", [True, False])
def test_full_solve_fp64_matches_oracle(vx, orc, directed):
    rng = np.random.default_rng(7)
    n = 300
    edges = [tuple(x) for x in rng.integers(0, n, size=(900, 2)).tolist()]
    g, m = twins(vx, orc, n, directed, edges)
    pr = vx.IncrementalPageRank(g)
    ref, info = orc.pagerank_full(m)
    got = np.asarray(pr.values)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() < 1e-13
    assert pr.last_full_recompute_converged == info["converged"]
    assert pr.last_repair_iterations == info["iterations"]
    assert abs(pr.last_residual_linf - info["residual_linf"]) < 1e-14
    pr.recompute(1000, 1e-14)
    ref2, _ = orc.pagerank_full(m, 0.85, 1000, 1e-14)
    assert np.abs(np.asarray(pr.values) - ref2).max() < 1e-13


def test_damping_validation(vx):
    g = vx.Graph(3, True)
    with pytest.raises(ValueError):
        vx.IncrementalPageRank(g, 1.5)


def test_localized_reference_cases(vx, orc):
    # tests/test_pagerank_localized.cpp:31-81
    cyc = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4)]
    g, m = twins(vx, orc, 8, True, [])
    g.apply(batch_of(vx, [(u, v, 1) for u, v in cyc]))
    m.apply([u for u, _ in cyc], [v for _, v in cyc], [1] * 8)
    pr = vx.IncrementalPageRank(g)
    ref = orc.ModelPageRank(m)
    assert pr.last_full_recompute_converged and pr.last_residual_linf <= 1e-12
    pr.apply(batch_of(vx, [(0, 2, 1)]), 128, 1e-13, 0.95)
    ref.apply([0], [2], [1], 128, 1e-13, 0.95)
    assert pr.last_repair_iterations == ref.diag.last_repair_iterations > 1
    assert pr.last_repaired_vertices == ref.diag.last_repaired_vertices
    assert 3 <= pr.last_repaired_vertices < g.vertex_count
    assert np.abs(np.asarray(pr.values) - ref.values).max() < 1e-12
    v = pr.validate_against_full(1000, 1e-13, 1e-7, 1e-8)
    assert v.reference_converged and v.within_tolerance and not v.fallback_applied
    assert v.l1_error <= 1e-7 and v.linf_error <= 1e-8
    # active-fraction fallback
    g2, m2 = twins(vx, orc, 8, True, cyc)
    pr2 = vx.IncrementalPageRank(g2)
    pr2.apply(batch_of(vx, [(0, 2, 1)]), 64, 1e-12, 0.20)
    assert pr2.last_repaired_vertices == g2.vertex_count
    assert pr2.last_full_recompute_converged
    # dangling transition (:86-101)
    g3, m3 = twins(vx, orc, 4, True, [(0, 1), (1, 2)])
    pr3 = vx.IncrementalPageRank(g3)
    assert abs(sum(pr3.values) - 1.0) < 1e-10
    pr3.apply(batch_of(vx, [(2, 0, 1)]), 128, 1e-13, 0.95)
    assert pr3.last_repaired_vertices == g3.vertex_count
    assert pr3.last_full_recompute_converged and pr3.last_residual_linf <= 1e-12
    # validated fallback (:122-131)
    g4, _ = twins(vx, orc, 8, True, cyc)
    pr4 = vx.IncrementalPageRank(g4)
    met = pr4.apply_validated(batch_of(vx, [(0, 2, 1)]), 1, 1e-3, 0.99, 1000, 1e-13, 1e-12, 1e-13)
    assert met.reference_converged and (met.fallback_applied or met.within_tolerance)
    assert pr4.validate_against_full(1000, 1e-13, 1e-12, 1e-13).within_tolerance


def test_growth_and_ci_smoke(vx):
    # tests/test_incremental.cpp:176-197 and ci.yml:243-253
    g = vx.Graph(2, False)
    g.add_edge(0, 1)
    pr = vx.IncrementalPageRank(g)
    pr.apply(batch_of(vx, [(2, 3, 1)]))
    assert g.vertex_count == 4 and pr.last_repaired_vertices == 4
    assert pr.validate_against_full().within_tolerance
    g7 = vx.Graph(5, True)
    g7.apply(batch_of(vx, [(0, 1, 1), (1, 2, 1), (2, 0, 1), (2, 3, 1), (3, 4, 1)]))
    p7 = vx.IncrementalPageRank(g7)
    p7.apply(batch_of(vx, [(4, 0, 1)]))
    assert len(p7.values) == 5 and p7.last_repaired_vertices > 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_randomized_apply_vs_oracle(vx, orc, seed):
    rng = np.random.default_rng(seed)
    n = 40
    # cycles keep every vertex non-dangling so localized repair really runs
    edges = [(i, (i + 1) % n) for i in range(n)] + [tuple(x) for x in rng.integers(0, n, size=(30, 2)).tolist()]
    g, m = twins(vx, orc, n, True, edges)
    pr = vx.IncrementalPageRank(g)
    ref = orc.ModelPageRank(m)
    for step in range(15):
        ops = [(int(a), int(b), int(rng.random() < 0.75)) for a, b in rng.integers(0, n, size=(3, 2))]
        ops = [(a, b, ad) for a, b, ad in ops if ad or (a + 1) % n != b]  # keep the cycle
        pr.apply(batch_of(vx, ops))
        ref.apply([o[0] for o in ops], [o[1] for o in ops], [o[2] for o in ops])
        assert np.abs(np.asarray(pr.values) - ref.values).max() < 1e-11, step
        assert pr.last_repaired_vertices == ref.diag.last_repaired_vertices, step
        assert pr.last_repair_iterations == ref.diag.last_repair_iterations, step


def test_fp32_rmat_within_1e6_of_fp64(vx, orc):
    """Directed R-MAT scale 16: FP32 ranks and sums, solved to an L1 residual
    below 1e-7; L1 and max-abs error against the FP64 oracle below 1e-6."""
    s, d = vx.rmat_edges(16, 16, 1)
    g = vx.Graph(1 << 16, True)
    g.bulk_load_arrays(s, d)
    m = orc.ModelGraph(1 << 16, True)
    m.bulk_load_edges(np.stack([s, d], 1))
    ref, info = orc.pagerank_full(m, 0.85, 1000, 1e-15)
    pr = vx.IncrementalPageRank(g, 0.85, 32)
    pr.solve_l1(500, 1e-7)
    got = pr.values_numpy()
    err = np.abs(got - ref)
    assert err.sum() <= 1e-6, err.sum()
    assert err.max() <= 1e-6, err.max()
    assert pr.last_residual_l1 < 1e-7
    # FP64 on the device reproduces the oracle closely
    p64 = vx.IncrementalPageRank(g, 0.85, 64)
    p64.solve_l1(1000, 1e-13)
    assert np.abs(p64.values_numpy() - ref).max() < 1e-12


def test_fp32_incremental_batches(vx, orc):
    """Directed R-MAT scale 14 with 10 batches (80/20): every batch's FP32
    result stays within 1e-6 (L1 and max-abs) of the FP64 oracle."""
    from velographx_amd import workloads as wl

    s, d = vx.rmat_edges(14, 16, 1)
    n = 1 << 14
    g = vx.Graph(n, True)
    g.bulk_load_arrays(s, d)
    m = orc.ModelGraph(n, True)
    m.bulk_load_edges(np.stack([s, d], 1))
    pr = vx.IncrementalPageRank(g, 0.85, 32)
    eu, ev = wl.canonical_edges(s, d, True)
    st = wl.BatchStream(eu, ev, n, True, seed=3)
    for _ in range(10):
        bs, bd, ba = st.next(500, 0.8)
        pr.apply(vx.UpdateBatch.from_arrays(bs, bd, ba))
        m.apply(bs, bd, ba)
        pr.solve_l1(500, 1e-7, True)
        ref, _ = orc.pagerank_full(m, 0.85, 1000, 1e-15)
        err = np.abs(pr.values_numpy() - ref)
        assert err.sum() <= 1e-6 and err.max() <= 1e-6
