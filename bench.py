#!/usr/bin/env python3
"""Benchmark of the traversal hot path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
undirected R-MAT scale 22, edge factor 16, A/B/C/D = .57/.19/.19/.05, seed 1
(4,194,304 vertices, ~64.2M edges after loop/duplicate removal).

* step  = one direction-optimizing BFS from one of 16 seeded roots in the
          largest component (answer-ready: distances in HBM, stream synced);
* value = BFS GTEPS over the K timed steps (Graph500: m_c / t per root,
          harmonic mean; every root lies in the same component, so this is
          sum(m_c) / sum(t));
* then 100 update batches of 10,000 ops (80% inserts of random non-edges,
  20% deletions of live edges) are applied and repaired to exact distances,
  with a full consolidation every 20 batches -> repairs/s, update-ops/s;
* roofline: algorithmic bytes of one BFS (4*A_c + 12*V_c) over the summed
  duration of its level-kernel launches, timed with HIP events on the
  graph's stream;
* cpu_baseline: the reference engine (oracle/_ref/vgx_ref, compiled from
  the reference sources) or, where it is absent, the oracle's queue BFS,
  timed on this host on a bounded sample of the same roots.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BFS GTEPS + update-batch repairs/s at R-MAT s22; batched vs plain, 1 MI355X"
HBM_PEAK_GBPS = 8000.0
INF = 0xFFFFFFFF


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=16)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--scale", type=int, default=22)
    p.add_argument("--edgefactor", type=int, default=16)
    p.add_argument("--roots", type=int, default=16)
    p.add_argument("--batches", type=int, default=100)
    p.add_argument("--batch-size", type=int, default=10000)
    p.add_argument("--consolidate-every", type=int, default=20)
    p.add_argument("--cpu-budget", type=float, default=20.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-repairs", action="store_true")
    return p.parse_args()


def cpu_baseline(s, d, n, roots, m_c, budget):
    """Time the reference engine's BFS (one thread) on a bounded sample."""
    ref = os.path.join(ROOT, "oracle", "_ref", "vgx_ref")
    cores = 1
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "edges.bin")
            np.stack([s, d], 1).astype(np.uint32).tofile(path)
            cmd = [ref, "bench-bfs", path, str(n), "0", str(budget)] + [str(int(r)) for r in roots]
            out = subprocess.run(cmd, capture_output=True, text=True, timeout=600, check=True).stdout
        times = []
        for line in out.splitlines():
            w = line.split()
            if w and w[0] == "root":
                times.append(float(w[2]))
        kind = "reference"
        sample = f"IncrementalBFS(graph, root).recompute() of the reference engine on the same graph, {len(times)} roots"
    else:
        from oracle import oracle as orc

        orc.build()
        g = orc.ModelGraph(n, False)
        g.bulk_load_edges(np.stack([s, d], 1))
        off, adj = g.csr()
        times = []
        t_all = time.perf_counter()
        for r in roots:
            t0 = time.perf_counter()
            orc.bfs(n, off, adj, int(r))
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_all > budget:
                break
        kind = "port"
        sample = f"oracle queue BFS (C restatement) on the same graph, {len(times)} roots"
    if not times:
        return None
    gteps = m_c * len(times) / sum(times) / 1e9
    return {"value": gteps, "unit": "GTEPS", "cores": cores, "kind": kind, "sample": sample,
            "ms_per_bfs": 1e3 * sum(times) / len(times)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import velographx_amd as vx
    from velographx_amd import workloads as wl

    assert vx.device_count() > 0, "no gfx950 device visible"
    n = 1 << args.scale
    t0 = time.time()
    s, d = vx.rmat_edges(args.scale, args.edgefactor, 1)
    log(f"R-MAT s{args.scale} ef{args.edgefactor}: {s.size} tuples in {time.time() - t0:.1f}s")
    g = vx.Graph(n, False)
    t0 = time.time()
    g.bulk_load_arrays(s, d)
    log(f"bulk load: {g.edge_count_directed} arcs in {time.time() - t0:.2f}s")
    lcc = wl.largest_component_from_hub(vx, g)
    roots = wl.roots_in_component(lcc, args.roots, seed=1)
    deg = g.degrees()
    arcs_c = int(deg[lcc].astype(np.int64).sum())
    m_c = arcs_c / 2.0
    v_c = int(lcc.size)
    bytes_bfs = 4.0 * arcs_c + 12.0 * v_c
    log(f"LCC: {v_c} vertices, {arcs_c} arcs; bytes/BFS {bytes_bfs / 1e6:.1f} MB")

    bfs = vx.IncrementalBFS(g, int(roots[0]))
    for i in range(args.warmup):
        bfs_w = vx.IncrementalBFS(g, int(roots[i % len(roots)]))
        del bfs_w
    # timed BFS steps, each answer-ready
    objs = [vx.IncrementalBFS(g, int(r)) for r in roots]
    for o in objs:
        o.set_profiling(True)
    torch.cuda.synchronize()
    times = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        o = objs[k % len(objs)]
        t0 = time.perf_counter()
        o.recompute()
        times.append(time.perf_counter() - t0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    launches = sum(o.stats()["level_launches"] for o in objs)
    kernel_ms = sum(o.stats()["level_kernel_ms"] for o in objs)
    st0 = objs[0].stats()
    reached = {o.stats()["reachable_count"] for o in objs}
    assert reached == {v_c}, reached
    gteps = m_c * args.steps / sum(times) / 1e9
    ms_per_step = 1e3 * elapsed / args.steps
    kernel_ms_per_bfs = kernel_ms / args.steps
    achieved = bytes_bfs / (kernel_ms_per_bfs * 1e-3) / 1e9
    log(f"BFS: {ms_per_step:.3f} ms/step, {gteps:.1f} GTEPS, kernel {kernel_ms_per_bfs:.3f} ms/BFS "
        f"over {launches / args.steps:.1f} launches, levels {st0['levels']} (bottom-up {st0['bottom_up_levels']})")
    # parity spot check against a top-down-only traversal of the same root
    chk = vx.IncrementalBFS(g, int(roots[0]))
    ref_d = chk.distances_numpy()
    chk.set_tuning(alpha=1e-9, beta=1.0, blocks=0, levels_per_sync=8)
    chk.recompute()
    exact_bfs = bool(np.array_equal(ref_d, chk.distances_numpy()))
    del chk

    result = {
        "metric": METRIC,
        "value": gteps,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"R-MAT s{args.scale} ef{args.edgefactor} undirected seed 1: direction-optimizing BFS "
                        f"from {args.roots} LCC roots, then {args.batches} x {args.batch_size}-op update batches",
            "vertices": n, "arcs": int(g.edge_count_directed), "lcc_vertices": v_c, "lcc_arcs": arcs_c,
            "roots": args.roots, "batches": args.batches, "batch_size": args.batch_size,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "k_bfs_level",
            "bytes_per_bfs": bytes_bfs,
            "avg_launch_us": 1e3 * kernel_ms / max(launches, 1),
            "launches_per_bfs": launches / args.steps,
            "wall_frac": bytes_bfs / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS,
        },
        "bfs": {"levels": st0["levels"], "bottom_up_levels": st0["bottom_up_levels"],
                "exact_vs_top_down": exact_bfs},
    }

    if not args.no_repairs and args.batches > 0:
        t0 = time.time()
        eu, ev = wl.canonical_edges(s, d, False)
        stream = wl.BatchStream(eu, ev, n, False, seed=1)
        batches = [stream.next(args.batch_size, 0.8) for _ in range(args.batches)]
        log(f"generated {len(batches)} batches in {time.time() - t0:.1f}s")
        inc = vx.IncrementalBFS(g, int(roots[0]))
        batch_t = []
        full = 0
        exact = True
        for i, (bs, bd, ba) in enumerate(batches):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            inc.apply_arrays(bs, bd, ba)
            if args.consolidate_every and (i + 1) % args.consolidate_every == 0:
                g.compact()
            batch_t.append(time.perf_counter() - t0)
            full += int(inc.last_used_full_recompute)
            if i % 10 == 9 or i == len(batches) - 1:
                ver = vx.IncrementalBFS(g, int(roots[0]))
                if not np.array_equal(ver.distances_numpy(), inc.distances_numpy()):
                    exact = False
                del ver
        tot = sum(batch_t)
        result["repairs"] = {
            "repairs_per_s": len(batch_t) / tot,
            "update_ops_per_s": len(batch_t) * args.batch_size / tot,
            "mean_batch_ms": 1e3 * tot / len(batch_t),
            "p50_batch_ms": 1e3 * float(np.median(batch_t)),
            "full_recompute_batches": full,
            "exact": exact,
            "consolidate_every": args.consolidate_every,
        }
        log(f"repairs: {result['repairs']}")

    if not args.no_cpu:
        t0 = time.time()
        cb = cpu_baseline(s, d, n, roots, m_c, args.cpu_budget)
        result["cpu_baseline"] = cb
        log(f"cpu baseline in {time.time() - t0:.1f}s: {cb}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
