"""Consolidation policy and controller (CPU): the oracle against the
reference's own test vectors and the compiled reference header, then the
engine's C++ controller (through the bindings) against the oracle."""
import random

import numpy as np
import pytest

from oracle import consolidation as oc
from tests import refdriver


# tests/test_consolidation_controller.cpp:17-52, as (signal, epoch, expected
# decision) steps; mark_consolidated after the steps at epochs 13 and 14.
QUIET = (1.05, 1.10, False, False, False)
LATENCY_NOISE = (1.08, 1.30, False, True, True)
LATENCY_WITH_GROWTH = (1.12, 1.30, False, True, True)
STORAGE = (1.26, 1.05, True, False, True)
REFERENCE_STEPS = (
    [(LATENCY_NOISE, e, False) for e in range(1, 9)]
    + [(LATENCY_WITH_GROWTH, e, e == 13) for e in range(9, 14)]
    + [(STORAGE, 14, True)]
    + [(LATENCY_WITH_GROWTH, e, False) for e in range(15, 24)]
    + [(LATENCY_WITH_GROWTH, 24, False), (LATENCY_WITH_GROWTH, 25, False), (QUIET, 26, False)]
    + [(LATENCY_WITH_GROWTH, e, e == 31) for e in range(27, 32)]
)
MARKS = (13, 14)


def run_reference_steps(observe, mark, streak):
    for signal, epoch, want in REFERENCE_STEPS:
        assert observe(signal, epoch) == want, epoch
        if epoch == 8:
            assert streak() == 0  # :25-27
        if epoch in MARKS:
            mark(epoch)
            assert streak() == 0


def test_oracle_reference_vectors():
    # test_consolidation_controller.cpp:8-15
    assert oc.scale_aware_policy(99_999_999) == (1.25, 1.25)
    assert oc.scale_aware_policy(100_000_000) == (1.50, 1.25)
    assert oc.scale_aware_policy(250_000_000, 1.40) == (1.50, 1.40)
    c = oc.Controller()
    run_reference_steps(c.observe, c.mark_consolidated, lambda: c.streak)
    # test_dynamic.cpp:132-139
    assert not oc.evaluate(120, 100, 120.0, 100.0)[4]
    s = oc.evaluate(125, 100, 100.0, 100.0)
    assert s[2] and s[4]
    s = oc.evaluate(110, 100, 125.0, 100.0)
    assert s[3] and s[4]
    # zero canonical size / non-positive canonical latency read as ratio 1
    assert oc.evaluate(500, 0, 3.0, 0.0)[:2] == (1.0, 1.0)


def test_oracle_planted_mistakes_fail():
    """Plausible slips in the controller (a strict bound, no cooldown reset,
    the growth gate dropped) each break a reference vector."""

    class Strict(oc.Controller):
        def observe(self, signal, epoch):
            sr, lr, se, le, sc = signal
            return super().observe((sr, lr, se, le and lr > 1.30, sc), epoch)

    class NoGrowthGate(oc.Controller):
        def observe(self, signal, epoch):
            sr, lr, se, le, sc = signal
            return super().observe((max(sr, 2.0) if not se else sr, lr, se, le, sc), epoch)

    class NoCooldown(oc.Controller):
        def mark_consolidated(self, epoch):
            self.streak = 0

    for cls in (Strict, NoGrowthGate, NoCooldown):
        c = cls()
        with pytest.raises(AssertionError):
            run_reference_steps(c.observe, c.mark_consolidated, lambda: c.streak)


def random_signals(seed, count):
    rng = random.Random(seed)
    out = []
    for e in range(1, count + 1):
        sr = rng.choice([1.0, 1.05, 1.09, 1.10, 1.12, 1.2, 1.25, 1.3, 1.5])
        lr = rng.choice([0.9, 1.0, 1.2, 1.25, 1.3, 2.0])
        se = sr >= 1.25 and rng.random() < 0.5
        le = lr >= 1.25
        out.append(((sr, lr, se, le, se or le), e))
    return out


@pytest.mark.skipif(not refdriver.available(), reason="oracle/_ref (compiled reference) not built")
def test_oracle_matches_compiled_reference():
    sc = refdriver.Scenario()
    arcs = [0, 1, 99_999_999, 100_000_000, 3 * 10**8]
    for a in arcs:
        sc.cmd("cons_policy", a, 1.3)
    evals = [(cur, canon, cl, kl, ms, ml) for cur in (0, 99, 100, 125, 150) for canon in (0, 100)
             for cl, kl in ((100.0, 100.0), (125.0, 100.0), (1.0, 0.0), (3.0, -1.0)) for ms, ml in ((1.25, 1.25), (1.5, 1.1))]
    for e in evals:
        sc.cmd("cons_eval", *e)
    configs = [(10, 5, 1.10), (3, 2, 1.0), (0, 1, 1.2)]
    streams = []
    for ci, cfg in enumerate(configs):
        sc.cmd("cons_ctl", *cfg)
        stream = random_signals(ci, 400)
        streams.append((cfg, stream))
        rng = random.Random(100 + ci)
        for (sig, epoch) in stream:
            sr, lr, se, le, scn = sig
            sc.cmd("cons_observe", repr(sr), repr(lr), int(se), int(le), int(scn), epoch)
            if rng.random() < 0.1:
                sc.cmd("cons_mark", epoch)
    recs = sc.run()
    it = iter(recs)
    for a in arcs:
        tag, val = next(it)
        assert tag == "cpolicy" and val == oc.scale_aware_policy(a, 1.3)
    for e in evals:
        tag, val = next(it)
        assert tag == "csignal" and val == oc.evaluate(*e[:4], policy=e[4:]), e
    for ci, (cfg, stream) in enumerate(streams):
        c = oc.Controller(*cfg)
        rng = random.Random(100 + ci)
        for sig, epoch in stream:
            tag, (fire, streak, last) = next(it)
            assert tag == "cobserve"
            assert c.observe(sig, epoch) == fire and c.streak == streak and c.last == last, (cfg, epoch)
            if rng.random() < 0.1:
                c.mark_consolidated(epoch)


@pytest.fixture(scope="module")
def vx():
    import velographx_amd as vx

    return vx


def to_vx_signal(vx, s):
    return vx.ConsolidationSignal(*s)


def test_engine_policy_and_evaluate(vx):
    for a in (0, 99_999_999, 100_000_000, 10**9):
        for lat in (1.25, 1.4):
            p = vx.scale_aware_consolidation_policy(a, lat)
            assert (p.max_storage_growth_ratio, p.max_neighbor_latency_ratio) == oc.scale_aware_policy(a, lat)
    p0 = vx.ConsolidationPolicy()
    assert (p0.max_storage_growth_ratio, p0.max_neighbor_latency_ratio) == (1.25, 1.25)
    for cur in (0, 99, 100, 110, 120, 125, 150):
        for canon in (0, 100):
            for cl, kl in ((120.0, 100.0), (125.0, 100.0), (3.0, 0.0), (3.0, -2.0)):
                for pol in ((1.25, 1.25), (1.5, 1.1)):
                    s = vx.evaluate_consolidation(cur, canon, cl, kl, vx.ConsolidationPolicy(*pol))
                    got = (s.storage_growth_ratio, s.neighbor_latency_ratio, s.storage_limit_exceeded,
                           s.latency_limit_exceeded, s.should_consolidate)
                    assert got == oc.evaluate(cur, canon, cl, kl, pol)


def test_engine_controller_reference_vectors(vx):
    c = vx.ConsolidationController()
    run_reference_steps(lambda s, e: c.observe(to_vx_signal(vx, s), e), c.mark_consolidated,
                        c.latency_breach_streak)
    assert c.last_consolidation_epoch() == 14


@pytest.mark.parametrize("cfg", [(10, 5, 1.10), (3, 2, 1.0), (0, 1, 1.2)])
def test_engine_controller_matches_oracle(vx, cfg):
    c = vx.ConsolidationController(vx.ConsolidationControllerConfig(*cfg))
    o = oc.Controller(*cfg)
    rng = random.Random(7)
    for sig, epoch in random_signals(sum(cfg[:2]), 500):
        assert c.observe(to_vx_signal(vx, sig), epoch) == o.observe(sig, epoch), epoch
        assert c.latency_breach_streak() == o.streak
        if rng.random() < 0.1:
            c.mark_consolidated(epoch)
            o.mark_consolidated(epoch)
        assert c.last_consolidation_epoch() == o.last


def test_probe_digest_pins():
    # splitmix64 from state 0: its first three published outputs
    z = oc.mix(np.array([0, oc.GOLDEN, (2 * oc.GOLDEN) & oc.MASK64], dtype=np.uint64))
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    # the probe rows with exact integers
    for n in (1, 7, 1000, 1 << 20, 4194304):
        want = [((i * oc.PROBE_MULT) & oc.MASK64) % n for i in range(5000)]
        assert oc.probe_rows(n, 5000).tolist() == want
    # digest by exact Python integers on a small random CSR (rows unsorted)
    rng = np.random.default_rng(3)
    n = 50
    rows = [list(rng.integers(0, n, size=int(rng.integers(0, 9)))) for _ in range(n)]
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    adj = np.array([v for r in rows for v in r], dtype=np.uint32)

    def mix_int(x):
        x = (x + oc.GOLDEN) & oc.MASK64
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & oc.MASK64
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & oc.MASK64
        return x ^ (x >> 31)

    want = 0
    for i in range(300):
        u = ((i * oc.PROBE_MULT) & oc.MASK64) % n
        want = (want + mix_int(len(rows[u])) + sum(mix_int(int(v)) for v in rows[u])) & oc.MASK64
    assert oc.probe_digest(off, adj, n, 300) == want
    # order inside rows does not matter; contents do
    adj2 = np.concatenate([np.array(r[::-1], dtype=np.uint32) for r in rows])
    assert oc.probe_digest(off, adj2, n, 300) == want
    adj3 = adj.copy()
    adj3[0] ^= 1
    assert oc.probe_digest(off, adj3, n, 10**4) != oc.probe_digest(off, adj, n, 10**4)
