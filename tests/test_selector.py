"""The adaptive repair-vs-recompute policy: the oracle's restatement of the
reference policy (oracle/selector.py) is pinned by hand cases that walk each
branch of benchmarks/publication_policy_bfs.cpp:162-233, then the library's
native policy (vgx_policy_*, host code, no device work) must make the same
decisions with the same predictions on random replayed traces."""
import os

import numpy as np
import pytest

from oracle import selector as osel


def test_large_branch_hand_cases():
    p = osel.Policy(1_000_000, initial_full_us=1000.0, initial_reachable_fraction=0.6)
    # preflight: >= 5% of the arcs in one batch
    assert p.decide(0.05)["reason"] == "large_preflight_full"
    # shallow shortest-parent deletions in the first batch
    assert p.decide(0.001, shallow=0.01, first_batch=True)["reason"] == "large_shallow_cold_start"
    # no incremental sample yet: probe the incremental arm
    d = p.decide(0.001, first_batch=True)
    assert (d["chose_full"], d["reason"]) == (False, "large_initial_incremental_probe")
    assert d["predicted_incremental_us"] == 0.0 and d["predicted_full_us"] == 1000.0
    # the probe cost 5000 us with 10% of the vertices affected
    p.observe(d, False, 5000.0, 100_000)
    d = p.decide(0.001)
    # 5000 * sqrt(1) * (1 + 3*0.1) = 6500; 6500*(1-0.35) = 4225 > 1000*1.2
    assert d["predicted_incremental_us"] == pytest.approx(6500.0)
    assert (d["chose_full"], d["reason"]) == (True, "large_uncertainty_confident_full")
    # a quarter of the update fraction halves the scale factor (sqrt), clamp at 0.5
    d = p.decide(0.001 / 16)
    assert d["predicted_incremental_us"] == pytest.approx(5000 * 0.5 * 1.3)
    # after the full arm: affected fraction resets, full EMA moves by 1/4
    d = p.decide(0.001)
    p.observe(d, False, 3000.0, 0)
    assert p.ema_full == pytest.approx(0.75 * 1000 + 0.25 * 3000)
    assert p.prev_affected_fraction == 0.0


def test_large_branch_overlap_keeps_incremental():
    p = osel.Policy(500_000, 1000.0, 0.5)
    d = p.decide(0.002)
    p.observe(d, False, 900.0, 10)
    d = p.decide(0.002)
    # 900 * 1 * (1 + 3*2e-5) * 0.65 < 1000 * 1.2
    assert (d["chose_full"], d["reason"]) == (False, "large_uncertainty_overlap_incremental")
    # a repair that fell back counts as a full-arm sample
    p.observe(d, True, 2000.0, 0)
    assert p.ema_full == pytest.approx(0.75 * 1000 + 0.25 * 2000) and p.have_full_err


def test_small_branch_hand_cases():
    # very sparse reach: always full
    p = osel.Policy(10_000, 100.0, 0.01)
    assert p.decide(0.001)["reason"] == "scale_very_sparse_reach"
    # sparse reach guard needs update fraction >= 0.25%
    p = osel.Policy(10_000, 100.0, 0.1)
    assert p.decide(0.003)["reason"] == "scale_sparse_reach_guard"
    assert p.decide(0.002)["reason"] == "scale_insufficient_evidence_incremental"
    # fresh model with incremental 1.25x slower than full -> full
    p = osel.Policy(10_000, 100.0, 0.9)
    d = p.decide(0.001)
    assert d["reason"] == "scale_insufficient_evidence_incremental"
    p.observe(d, False, 120.0, 100)  # affected fraction 0.01
    d = p.decide(0.001)
    # 120 * (1 + 2*0.01) = 122.4 <= 125: still incremental, now with a fresh model
    assert (d["chose_full"], d["reason"]) == (False, "scale_confidence_incremental")
    p.observe(d, False, 200.0, 100)
    d = p.decide(0.001)
    # EMA 0.75*120 + 0.25*200 = 140; 140*1.02 = 142.8 > 125
    assert d["predicted_incremental_us"] == pytest.approx(142.8)
    assert (d["chose_full"], d["reason"]) == (True, "scale_fresh_cost_model")
    # staleness: five batches without a full sample -> not fresh
    for _ in range(5):
        p.observe({"chose_full": False, "predicted_incremental_us": 0.0, "predicted_full_us": 0.0,
                   "update_fraction": 0.001}, False, 200.0, 100)
    assert p.decide(0.001)["reason"] == "scale_insufficient_evidence_incremental"


def test_shallow_fraction_rule():
    dist = np.array([0, 1, 2, 3, 4, 0xFFFFFFFF], np.uint32)
    src = [0, 1, 3, 2, 4, 0]
    dst = [1, 2, 4, 1, 5, 2]
    add = [0, 0, 0, 0, 0, 1]
    # directed: (0,1) d=1, (1,2) d=2 count; (3,4) depth 4 no; (2,1) reversed no; (4,5) unreachable; add skipped
    assert osel.shallow_parent_deletion_fraction(src, dst, add, dist, True) == pytest.approx(2 / 6)
    # undirected also counts (2,1): d(1)+1 = d(2) = 2
    assert osel.shallow_parent_deletion_fraction(src, dst, add, dist, False) == pytest.approx(3 / 6)


@pytest.mark.parametrize("vertices", [50_000, 4_194_304])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_native_policy_matches_oracle(vertices, seed):
    vx = pytest.importorskip("velographx_amd")
    rng = np.random.default_rng(seed)
    reach = float(rng.choice([0.01, 0.1, 0.6]))
    full0 = float(rng.uniform(100, 5000))
    ref = osel.Policy(vertices, full0, reach)
    nat = vx.AdaptivePolicy(vertices, full0, reach)
    for i in range(200):
        uf = float(10 ** rng.uniform(-6, -1))
        shallow = float(rng.choice([0.0, 0.0, 0.01])) if i == 0 else 0.0
        r = ref.decide(uf, shallow, i == 0)
        t = nat.decide(uf, shallow, i == 0)
        assert t["reason"] == r["reason"], i
        assert t["chose_full"] == r["chose_full"], i
        assert t["predicted_incremental_us"] == pytest.approx(r["predicted_incremental_us"], rel=1e-12, abs=0)
        assert t["predicted_full_us"] == pytest.approx(r["predicted_full_us"], rel=1e-12, abs=0)
        fallback = bool(not r["chose_full"] and rng.random() < 0.1)
        cost = float(rng.uniform(50, 8000))
        affected = int(rng.integers(0, vertices // 4))
        ref.observe(r, fallback, cost, affected)
        t.update({"execution_us": cost, "internal_fallback": fallback, "affected_vertices": affected})
        nat.observe(t)


POLICY_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                          "policy_bfs")


@pytest.mark.skipif(not os.path.exists(POLICY_BIN),
                    reason="oracle/_ref/policy_bfs (the reference's publication_policy_bfs.cpp) not built")
@pytest.mark.parametrize("scale,edgefactor,batch", [(11, 8, 300), (12, 16, 400), (18, 2, 8000), (18, 2, 3000)])
def test_policy_replay_of_the_compiled_reference(scale, edgefactor, batch, tmp_path):
    """The reference's own harness (benchmarks/publication_policy_bfs.cpp,
    compiled as it is) runs its adaptive policy on a sliding-window stream
    over an R-MAT edge list; each batch's recorded features and measured
    execution time, replayed through oracle/selector.py, give the same
    reason, choice and predictions (the harness prints six significant
    digits, so predictions agree to that).  Scale 18 exercises the
    large-graph rules, scale 11-12 the small-graph ones."""
    import json
    import subprocess

    vx = pytest.importorskip("velographx_amd")
    s, d = vx.rmat_edges(scale, edgefactor, 3)
    path = tmp_path / "edges.txt"
    path.write_text("\n".join(f"{u} {v}" for u, v in zip(s.tolist(), d.tolist())))
    out = subprocess.run([POLICY_BIN, str(path), "0", "0.9", str(batch), "0.01"], capture_output=True, text=True,
                         timeout=300, check=True).stdout
    run = json.loads(out)
    ad = next(p for p in run["policies"] if p["name"] == "adaptive")
    traces, exec_us, fallback = ad["trace"], ad["execution_us"], ad["internal_full_fallback"]
    n = run["vertices"]
    pol = osel.Policy(n, traces[0]["predicted_full_us"], traces[0]["reachable_fraction"])
    reasons = set()
    for i, t in enumerate(traces):
        r = pol.decide(t["update_fraction"], t["shallow_parent_deletion_fraction"], i == 0)
        assert r["reason"] == t["reason"], (i, r, t)
        assert r["chose_full"] == t["explicit_full"], i
        for key in ("predicted_incremental_us", "predicted_full_us"):
            assert r[key] == pytest.approx(t[key], rel=2e-5, abs=1e-9), (i, key)
        reasons.add(t["reason"])
        # the next batch's trace records the affected fraction this one left
        nxt = traces[i + 1]["previous_affected_fraction"] if i + 1 < len(traces) else 0.0
        pol.observe(r, fallback[i], exec_us[i], nxt * n)
    assert len(traces) >= 5 and len(reasons) >= 2
