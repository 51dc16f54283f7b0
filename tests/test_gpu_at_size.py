"""GPU parity at the benchmark configurations' full sizes, checked against the
CPU oracle on the host (nothing here compares the GPU with itself):

* undirected R-MAT scale 22: the 16 seeded roots, bit-exact against the
  oracle's queue BFS; the stream of 100 update batches of 10,000 ops (80%
  inserts of non-edges, 20% deletions) answered through the adaptive
  selector with consolidation every 20 batches, checked every 10 batches
  against the oracle's BFS of the oracle's own graph after the same ops;
* Erdos-Renyi, 4,194,304 vertices, mean degree 16: 16 roots, then 20
  update batches of 10,000 ops through the adaptive selector, checked every
  10 batches;
* undirected R-MAT scale 20: all 64 roots of a multi-source pass (captured
  launch program) after each of 4 uneven update batches, and all 64 roots
  of the multi-source incremental repair after each of 6 batches of 10 to
  60,000 ops (4 forced repairs, 2 under the adaptive policy);
* directed R-MAT scale 22: the FP32 PageRank solve to an L1 residual below
  1e-7, its warm-started re-solve after 3 batches of 10,000 ops and the
  residual-push repair after 3 more (10,000, 10,000 and 100 ops), within
  1e-6 (L1 and max-abs) of the oracle's FP64 solve;
* directed weighted R-MAT scale 20, 32 one-hot teleport columns: the
  forward and the teleport gradient against the oracle's FP64 sweeps, and
  dLoss/dw on 200 sampled arcs against FP64 central differences (each
  perturbed loss evaluated exactly by the rank-one identity of
  oracle/ppr.finite_difference_w_rank1).

The oracle graph is oracle.CsrModel (pinned against ModelGraph in
tests/test_oracle.py); BFS is orc_bfs, PageRank orc_pagerank_full.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

INF = 0xFFFFFFFF
THREADS = 8


def log(msg):
    print(f"[at-size] {msg}", flush=True)


@pytest.fixture(scope="module")
def vx():
    import velographx_amd as vx

    assert vx.device_count() > 0, "no gfx950 device visible"
    return vx


@pytest.fixture(scope="module")
def wl():
    from velographx_amd import workloads

    return workloads


@pytest.fixture(scope="module")
def rmat22(vx):
    s, d = vx.rmat_edges(22, 16, 1)
    return s, d


def oracle_bfs_many(orc, model, roots):
    off, adj = model.csr()
    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(lambda r: orc.bfs(model.n, off, adj, int(r))[0], roots))


def component_of_hub(orc, model):
    deg = np.diff(model.off)
    dist, _ = orc.bfs(model.n, model.off, model.adj, int(np.argmax(deg)))
    return np.nonzero(dist != INF)[0].astype(np.uint32)


@pytest.fixture(scope="module")
def s22(vx, wl, orc, rmat22):
    s, d = rmat22
    n = 1 << 22
    model = orc.CsrModel(n, False)
    model.bulk_load_edges(s, d)
    lcc = component_of_hub(orc, model)
    roots = wl.roots_in_component(lcc, 16, seed=1)
    log(f"s22 undirected: {model.edge_count_directed()} arcs, component {lcc.size}")
    return model, lcc, roots


def test_config2_s22_bfs_16_roots(vx, orc, rmat22, s22):
    s, d = rmat22
    model, lcc, roots = s22
    g = vx.Graph(1 << 22, False)
    g.bulk_load_arrays(s, d)
    assert g.edge_count_directed == model.edge_count_directed()
    expect = oracle_bfs_many(orc, model, roots)
    walker = vx.IncrementalBFS(g, int(lcc[0]))
    for r, want in zip(roots, expect):
        walker.set_source(int(r))
        got = walker.distances_numpy()
        assert np.array_equal(got, want), int(r)
        assert walker.stats()["reachable_count"] == lcc.size
    # the same roots by fresh objects (construction = first traversal)
    for r, want in zip(roots[:4], expect[:4]):
        assert np.array_equal(vx.IncrementalBFS(g, int(r)).distances_numpy(), want), int(r)
    log("s22 BFS: 16 roots equal the oracle")


def test_config2_s22_selector_stream_100_batches(vx, wl, orc, rmat22, s22):
    s, d = rmat22
    model, lcc, roots = s22
    n = 1 << 22
    root = int(roots[0])
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    eu, ev = model.undirected_edges()
    stream = wl.BatchStream(eu, ev, n, False, seed=1)
    batches = [stream.next(10000, 0.8) for _ in range(100)]
    oracle = orc.CsrModel(n, False)
    oracle.off, oracle.adj = model.off.copy(), model.adj.copy()
    inc = vx.IncrementalBFS(g, root, 2.0)
    sel = vx.AdaptiveSelector(inc)
    pending = []
    full = 0
    for i, (bs, bd, ba) in enumerate(batches):
        tr = sel.apply_arrays(bs, bd, ba)
        full += int(tr["chose_full"] or tr["internal_fallback"])
        if (i + 1) % 20 == 0:
            g.compact()
        pending.append((bs, bd, ba))
        if (i + 1) % 10 == 0:
            # the ten batches in order as one op list: the last op on an arc wins
            oracle.apply(*(np.concatenate(x) for x in zip(*pending)))
            pending = []
            assert g.edge_count_directed == oracle.edge_count_directed(), i
            want = oracle_bfs_many(orc, oracle, [root])[0]
            assert np.array_equal(inc.distances_numpy(), want), i
            assert inc.reachable_count == int(np.count_nonzero(want != INF))
    log(f"s22 selector stream: 100 batches, {full} answered by full recompute, all checks equal the oracle")


def test_config2_erdos_renyi_4m_bfs(vx, wl, orc):
    n = 1 << 22
    u, v = wl.erdos_renyi(n, 8 * n, seed=1)
    model = orc.CsrModel(n, False)
    model.bulk_load_edges(u, v)
    lcc = component_of_hub(orc, model)
    roots = wl.roots_in_component(lcc, 16, seed=1)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(u, v)
    assert g.edge_count_directed == model.edge_count_directed()
    expect = oracle_bfs_many(orc, model, roots)
    walker = vx.IncrementalBFS(g, int(roots[0]))
    for r, want in zip(roots, expect):
        walker.set_source(int(r))
        assert np.array_equal(walker.distances_numpy(), want), int(r)
    log(f"ER: 16 roots equal the oracle (component {lcc.size})")
    # the repairs leg's update stream on the same graph, through the selector
    eu, ev = model.undirected_edges()
    stream = wl.BatchStream(eu, ev, n, False, seed=1)
    root = int(roots[0])
    inc = vx.IncrementalBFS(g, root, 2.0)
    sel = vx.AdaptiveSelector(inc)
    pending = []
    for i in range(20):
        bs, bd, ba = stream.next(10000, 0.8)
        sel.apply_arrays(bs, bd, ba)
        pending.append((bs, bd, ba))
        if (i + 1) % 10 == 0:
            model.apply(*(np.concatenate(x) for x in zip(*pending)))
            pending = []
            assert g.edge_count_directed == model.edge_count_directed(), i
            want = oracle_bfs_many(orc, model, [root])[0]
            assert np.array_equal(inc.distances_numpy(), want), i
    log("ER: 20 update batches of 10,000 ops through the selector equal the oracle")


def test_config3_s20_64_roots_after_uneven_batches(vx, wl, orc):
    n = 1 << 20
    s, d = vx.rmat_edges(20, 16, 1)
    model = orc.CsrModel(n, False)
    model.bulk_load_edges(s, d)
    lcc = component_of_hub(orc, model)
    roots = [int(r) for r in wl.roots_in_component(lcc, 64, seed=1)]
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    ms = vx.MultiSourceBFS(g, roots)
    ms.set_graph(True)
    ms.run(g)
    expect = oracle_bfs_many(orc, model, roots)
    for i, want in enumerate(expect):
        assert np.array_equal(ms.distances(i), want), (i, roots[i])
    eu, ev = model.undirected_edges()
    stream = wl.BatchStream(eu, ev, n, False, seed=3)
    sizes = [int(k) for k in wl.log_uniform_sizes(3, 1.0, 5.0, seed=1)] + [100000]
    for b, k in enumerate(sizes):
        bs, bd, ba = stream.next(k, 0.8)
        g.apply_arrays(bs, bd, ba)
        model.apply(bs, bd, ba)
        ms.run(g)
        expect = oracle_bfs_many(orc, model, roots)
        for i, want in enumerate(expect):
            assert np.array_equal(ms.distances(i), want), (b, k, i, roots[i])
    log(f"s20 multi-source: 64 roots equal the oracle after batches of {sizes} ops")


def test_config3_s20_64_root_repairs(vx, wl, orc):
    """The 64-root incremental repair at full size: every batch repaired
    (never handed to a pass), then the adaptive policy; all 64 roots against
    the oracle after every batch."""
    n = 1 << 20
    s, d = vx.rmat_edges(20, 16, 1)
    model = orc.CsrModel(n, False)
    model.bulk_load_edges(s, d)
    lcc = component_of_hub(orc, model)
    roots = [int(r) for r in wl.roots_in_component(lcc, 64, seed=1)]
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    ms = vx.MultiSourceBFS(g, roots)
    ms.set_graph(True)
    ms.set_repair(1, 1.0)
    ms.run(g)
    eu, ev = model.undirected_edges()
    stream = wl.BatchStream(eu, ev, n, False, seed=5)
    plan = [(10, 1), (300, 1), (3000, 1), (30000, 1), (25, 0), (60000, 0)]
    for b, (k, mode) in enumerate(plan):
        ms.set_repair(mode, 1.0 if mode == 1 else 0.1)
        bs, bd, ba = stream.next(k, 0.8)
        ms.apply_arrays(bs, bd, ba)
        model.apply(bs, bd, ba)
        st = ms.repair_stats()
        if mode == 1:
            assert st["last_repaired"], (b, k, st)
        expect = oracle_bfs_many(orc, model, roots)
        for i, want in enumerate(expect):
            assert np.array_equal(ms.distances(i), want), (b, k, i, roots[i])
        lv = ms.levels_u8()
        for i, want in enumerate(expect[:8]):
            w8 = want.astype(np.int64)
            w8[want == INF] = 255
            assert np.array_equal(lv[i], w8.astype(np.uint8)), (b, i)
        log(f"s20 multi-source repair: batch of {k} ops ({'repaired' if st['last_repaired'] else 'full pass'}, "
            f"{st['last_affected_pairs']} affected pairs, {st['last_lowered']} lowered), 64 roots equal the oracle")


def test_config4_s22_directed_fp32_pagerank(vx, wl, orc, rmat22):
    s, d = rmat22
    n = 1 << 22
    model = orc.CsrModel(n, True)
    model.bulk_load_edges(s, d)
    g = vx.Graph(n, True)
    g.bulk_load_arrays(s, d)
    pr = vx.IncrementalPageRank(g, 0.85, 32)

    def check(tag):
        ref, info = orc.pagerank_full(model, 0.85, 150, 1e-15)
        got = pr.values_numpy().astype(np.float64)
        assert got.size == ref.size
        err = np.abs(got - ref)
        log(f"PageRank {tag}: L1 {err.sum():.3e}, max-abs {err.max():.3e} (oracle {info['iterations']} sweeps)")
        assert err.sum() <= 1e-6 and err.max() <= 1e-6, tag
        assert abs(got.sum() - 1.0) <= 1e-5

    pr.solve_l1(200, 1e-7, False)
    assert pr.last_residual_l1 < 1e-7
    check("solve")
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(model.off).astype(np.int64))
    stream = wl.BatchStream(rows, model.adj, n, True, seed=2)
    for b in range(3):
        bs, bd, ba = stream.next(10000, 0.8)
        g.apply_arrays(bs, bd, ba)
        model.apply(bs, bd, ba)
        pr.solve_l1(200, 1e-7, True)
        check(f"batch {b}")
    # the residual-push repair on the same stream (10,000 ops, then 100)
    for b, k in enumerate((10000, 10000, 100)):
        bs, bd, ba = stream.next(k, 0.8)
        r = pr.apply_push_arrays(bs, bd, ba, l1_tol=1e-7)
        model.apply(bs, bd, ba)
        assert r["residual_l1"] <= 1e-7, r
        check(f"push batch {b} ({k} ops, fallback {r['used_fallback']}, {r['rounds']} rounds)")


def test_config5_s20_ppr_fd_200_arcs(vx):
    import torch

    from oracle import ppr
    from velographx_amd import autograd as ag

    n = 1 << 20
    K = 32
    D = 0.85
    T = 150  # sweeps: truncation d^151 < 3e-11 of each quantity
    s, d = vx.rmat_edges(20, 16, 1)
    g = vx.Graph(n, True)
    g.bulk_load_arrays(s, d)
    src, dst = (a.astype(np.int64) for a in ag.arc_list(g))
    rng = np.random.default_rng(1)
    w = (1.0 - rng.random(src.size)).astype(np.float32)  # uniform on (0, 1]
    S = np.zeros((n, K), np.float32)
    S[rng.integers(0, n, K), np.arange(K)] = 1.0
    C = rng.standard_normal((n, K)).astype(np.float32)
    wt = torch.tensor(w, device="cuda", requires_grad=True)
    st = torch.tensor(S, device="cuda", requires_grad=True)
    x = ag.personalized_pagerank(g, wt, st, D, T)
    (x * torch.tensor(C, device="cuda")).sum().backward()
    torch.cuda.synchronize()
    x, gw, gs = x.detach().cpu().numpy(), wt.grad.cpu().numpy(), st.grad.cpu().numpy()
    # 200 arcs: 25 from each of 8 sampled sources with at least 25 out-arcs
    outdeg = np.bincount(src, minlength=n)
    cand = np.nonzero(outdeg >= 25)[0]
    sources = np.sort(rng.choice(cand, 8, replace=False))
    starts = np.searchsorted(src, sources)
    arcs = np.concatenate([st0 + np.sort(rng.choice(outdeg[u], 25, replace=False))
                           for u, st0 in zip(sources, starts)])
    assert arcs.size == 200 and np.all(src[arcs] == np.repeat(sources, 25))
    fd, X, L = ppr.finite_difference_w_rank1(n, src, dst, w.astype(np.float64), S, C, D, arcs, eps=1e-5,
                                             iterations=T)
    ex = float(np.max(np.abs(x - X)))
    es = float(np.max(np.abs(gs - (1 - D) * L)) / np.max(np.abs((1 - D) * L)))
    ew = float(np.max(np.abs(gw[arcs] - fd)) / np.max(np.abs(fd)))
    log(f"PPR s20: X max-abs {ex:.2e}, dS rel {es:.2e}, dw vs FD rel {ew:.2e} on {arcs.size} arcs")
    assert ex <= 1e-6
    assert es <= 1e-5
    assert ew <= 1e-5
