"""k-core: the oracle against the reference's own vector, closed forms and
the compiled reference (CPU); the device peeling against the oracle (GPU,
marked)."""
import numpy as np
import pytest

from oracle import kcore as okc
from tests import refdriver


@pytest.fixture(scope="module")
def orc():
    from oracle import oracle as orc

    orc.build()
    return orc


def model(orc, n, edges):
    g = orc.ModelGraph(n, False)
    if len(edges):
        g.bulk_load_edges(np.asarray(edges))
    return g


def cores(g):
    off, adj = g.csr()
    return okc.core_numbers(g.n, off, adj)


def test_oracle_reference_vector(orc):
    # tests/test_incremental.cpp:51-56
    g = model(orc, 4, [(0, 1), (0, 2), (0, 3), (1, 2)])
    assert cores(g).tolist() == [2, 2, 2, 1]


def test_oracle_closed_forms(orc):
    n = 9
    complete = [(i, j) for i in range(n) for j in range(i + 1, n)]
    assert cores(model(orc, n, complete)).tolist() == [n - 1] * n
    cycle = [(i, (i + 1) % 12) for i in range(12)]
    assert cores(model(orc, 12, cycle)).tolist() == [2] * 12
    tree = [(i, (i - 1) // 2) for i in range(1, 20)]
    assert cores(model(orc, 22, tree)).tolist() == [1] * 20 + [0, 0]
    # a 5-clique with a pendant path: the clique keeps 4, the path 1
    clique = [(i, j) for i in range(5) for j in range(i + 1, 5)]
    path = [(4, 5), (5, 6), (6, 7)]
    assert cores(model(orc, 8, clique + path)).tolist() == [4] * 5 + [1, 1, 1]


def test_oracle_planted_mistake_fails(orc):
    """Reading the degree at removal instead of the running maximum gives
    wrong cores on the reference's own vector."""
    g = model(orc, 4, [(0, 1), (0, 2), (0, 3), (1, 2)])
    off, adj = g.csr()
    off = np.asarray(off)
    degree = np.diff(off)
    assert cores(g).tolist() != degree.tolist()


def random_batches(n, seed, count, grow=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        k = int(rng.integers(1, 80))
        bs = rng.integers(0, n + grow, k).astype(np.uint32)
        bd = rng.integers(0, n + grow, k).astype(np.uint32)
        ba = (rng.random(k) < 0.6).astype(np.uint8)
        out.append((bs, bd, ba))
    return out


@pytest.mark.skipif(not refdriver.available(), reason="oracle/_ref (compiled reference) not built")
def test_oracle_matches_compiled_reference(orc):
    n = 300
    rng = np.random.default_rng(3)
    e = rng.integers(0, n, size=(900, 2))
    stream = random_batches(n, 4, 12, grow=10)
    sc = refdriver.Scenario()
    sc.graph(n, False)
    sc.bulk([(int(u), int(v)) for u, v in e])
    sc.cmd("kcore_new")
    for bs, bd, ba in stream:
        sc.batch(zip(bs.tolist(), bd.tolist(), ba.tolist()))
        sc.cmd("kcore_apply")
    recs = sc.run()
    got = [r[1] for r in recs if r[0] == "kcore"]
    diag = [r[1] for r in recs if r[0] == "kcorediag"]
    g = model(orc, n, e)
    assert np.array_equal(got[0], cores(g)) and diag[0] == n
    for (bs, bd, ba), want, dg in zip(stream, got[1:], diag[1:]):
        off0, adj0 = g.csr()
        before = (g.n, np.asarray(off0), np.asarray(adj0))
        g.apply(bs, bd, ba)
        off1, adj1 = g.csr()
        assert np.array_equal(cores(g), want)
        assert okc.affected_count(before, (g.n, np.asarray(off1), np.asarray(adj1)), bs, bd) == dg


@pytest.mark.gpu
@pytest.mark.parametrize("scale", [10, 13])
def test_engine_matches_oracle(orc, scale):
    import velographx_amd as vx

    assert vx.device_count() > 0
    n = 1 << scale
    s, d = vx.rmat_edges(scale, 16, 30 + scale)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    mg = orc.ModelGraph(n, False)
    mg.bulk_load_edges(np.stack([s, d], 1))
    kc = vx.IncrementalKCore(g)
    assert np.array_equal(kc.core_numpy(), cores(mg))
    assert kc.last_repaired_vertices == n
    for bs, bd, ba in random_batches(n, 40 + scale, 4, grow=20) + [
            (np.arange(0, 200, dtype=np.uint32), np.arange(1, 201, dtype=np.uint32), np.zeros(200, np.uint8))]:
        off0, adj0 = mg.csr()
        before = (mg.n, np.asarray(off0), np.asarray(adj0))
        kc.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        assert np.array_equal(kc.core_numpy(), cores(mg))
        off1, adj1 = mg.csr()
        after = (mg.n, np.asarray(off1), np.asarray(adj1))
        assert kc.last_repaired_vertices == okc.affected_count(before, after, bs, bd)
    kc.apply_arrays(np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint8))
    assert kc.last_repaired_vertices == 0 and np.array_equal(kc.core_numpy(), cores(mg))
    kc.recompute()
    assert np.array_equal(kc.core_numpy(), cores(mg))


@pytest.mark.gpu
def test_engine_closed_forms_and_errors(orc):
    import velographx_amd as vx

    n = 40
    complete = np.array([(i, j) for i in range(n) for j in range(i + 1, n)], dtype=np.uint32)
    g = vx.Graph(0, False)
    g.bulk_load_edges(complete)
    assert vx.IncrementalKCore(g).core == [n - 1] * n
    b = vx.UpdateBatch()
    b.add(3, 4)
    b.remove(0, 1)
    kc = vx.IncrementalKCore(g)
    kc.apply(b)
    mg = model(orc, n, complete)
    mg.apply(np.array([3, 0], np.uint32), np.array([4, 1], np.uint32), np.array([1, 0], np.uint8))
    assert np.array_equal(kc.core_numpy(), cores(mg))
    with pytest.raises(ValueError):
        vx.IncrementalKCore(vx.Graph(4, True))
    assert vx.IncrementalKCore(vx.Graph(0, False)).core == []
    iso = vx.Graph(5, False)
    assert vx.IncrementalKCore(iso).core == [0] * 5


@pytest.mark.gpu
def test_region_repair_on_separate_components(orc):
    """Batches that touch one of several components re-peel that component
    only (kcore.hpp:31-56), growth included, and match full recomputes."""
    import velographx_amd as vx

    assert vx.device_count() > 0
    rng = np.random.default_rng(9)
    parts = []
    for c in range(8):  # eight components of 500 vertices, ids interleaved
        ids = np.arange(c, 4000, 8)
        e = rng.integers(0, ids.size, size=(3000, 2))
        parts.append(np.stack([ids[e[:, 0]], ids[e[:, 1]]], 1))
    e = np.concatenate(parts).astype(np.uint32)
    g = vx.Graph(4000, False)
    g.bulk_load_edges(e)
    mg = orc.ModelGraph(4000, False)
    mg.bulk_load_edges(e)
    kc = vx.IncrementalKCore(g)
    for step in range(12):
        c = step % 8
        ids = np.arange(c, 4000, 8)
        k = 40
        bs = ids[rng.integers(0, ids.size, k)].astype(np.uint32)
        bd = ids[rng.integers(0, ids.size, k)].astype(np.uint32)
        if step == 5:
            bd[0] = 4100 + step  # a new vertex joins component c
        ba = (rng.random(k) < 0.5).astype(np.uint8)
        off0, adj0 = mg.csr()
        before = (mg.n, np.asarray(off0), np.asarray(adj0))
        kc.apply_arrays(bs, bd, ba)
        mg.apply(bs, bd, ba)
        off1, adj1 = mg.csr()
        after = (mg.n, np.asarray(off1), np.asarray(adj1))
        assert np.array_equal(kc.core_numpy(), cores(mg))
        rep = okc.affected_count(before, after, bs, bd)
        assert kc.last_repaired_vertices == rep and rep < mg.n


@pytest.mark.skipif(not refdriver.available(), reason="oracle/_ref (compiled reference) not built")
def test_bench_digests_equal_the_reference_programs(orc):
    """The digests bench.py compares at s20 (every core number; the
    partition labelled by smallest members) are those oracle/ref_driver.cpp
    prints: on a small graph its digests of the reference's IncrementalKCore
    and IncrementalComponents equal bench.vertex_digest of the oracle's core
    numbers and smallest-member labels, and a single changed core number
    changes the digest."""
    import os
    import subprocess
    import tempfile

    import bench
    from oracle import components as occ

    rng = np.random.default_rng(21)
    n = 3000
    e = rng.integers(0, n, size=(12000, 2)).astype(np.uint32)
    g = model(orc, n, e)
    off, adj = g.csr()
    core = cores(g)
    labels = np.asarray(occ.min_labels(n, off, adj), dtype=np.uint32)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "e.bin")
        e.tofile(p)
        k = subprocess.run([refdriver.REF_BIN, "bench-kcore", p, str(n)], capture_output=True, text=True,
                           check=True).stdout.split()
        c = subprocess.run([refdriver.REF_BIN, "bench-cc", p, str(n)], capture_output=True, text=True,
                           check=True).stdout.split()
    assert int(k[3]) == bench.vertex_digest(core)
    assert int(c[3]) == bench.vertex_digest(labels)
    changed = core.copy()
    changed[17] += 1
    assert bench.vertex_digest(changed) != int(k[3])
