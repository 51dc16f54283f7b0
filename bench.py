#!/usr/bin/env python3
"""Benchmark of the traversal hot path on one MI355X.

Headline workload: undirected R-MAT scale 22, edge factor 16,
A/B/C/D = .57/.19/.19/.05, seed 1 (4,194,304 vertices, ~64.2M edges after
loop/duplicate removal).

* step  = one direction-optimizing BFS from one of 16 seeded roots in the
          largest component (answer-ready: distances in HBM, stream synced),
          the first traversal from that root: one traversal object is
          re-targeted to the step's root; the warm-up traverses other roots
          of the component.  No state is kept per root (every launch-shape
          choice is made on the device from the traversal's own counters),
          so the figure does not change with --warmup;
* value = plain single-root BFS GTEPS over the K timed steps (Graph500:
          m_c / t; every root lies in the same component, so the harmonic
          mean is sum(m_c) / sum(t));
* roofline: algorithmic bytes of one BFS (4*A_c + 12*V_c) over its GPU
  time, measured in the timed region with HIP events on the graph's stream
  around every traversal (init and step kernels); traffic = HBM bytes per
  BFS from rocprofv3 FETCH_SIZE and WRITE_SIZE passes over a child run, each
  scaled by its reading for a 1 GiB copy;
* cpu_baseline: the reference engine (oracle/_ref/vgx_ref, compiled from the
  reference sources) timed on this host on the same roots: the value is its
  root-parallel BFS over all host threads (at most 16, one GPU's share of a
  box), as its thread-scaling benchmark runs it, with the one-thread figure
  and the CPU model beside it; where the reference build is absent, the
  oracle's queue BFS on one thread.

Secondary legs, each in its own object of the same JSON line:
* batched   -- 64 roots per multi-source pass on the same graph (aggregate
               GTEPS, batched / plain ratio, bit-exact check of the first 16
               roots against the plain path);
* consolidation -- device snapshot of the same graph (timed, GB/s), the
               on-device neighbour probe, and 24 epochs of 1,000,000-op
               batches under the consolidation controller with snapshot
               cutover; the reference engine's snapshot and probe beside it;
* static    -- the static CsrGraph of the same edges: build, bfs_distances,
               20 pagerank sweeps; the reference's beside it;
* kcore     -- all core numbers and connected components of undirected
               R-MAT s20 on the device; the reference's IncrementalKCore and
               IncrementalComponents constructions beside them;
* repairs   -- 100 update batches of 10,000 ops (80% inserts of random
               non-edges, 20% deletions of live edges) answered with exact
               distances, the adaptive selector choosing localized repair or
               full recompute per batch, consolidation every 20 batches;
               repairs/s, next to the always-repair and always-recompute arms
               and the reference engine's IncrementalBFS::apply timed on a
               bounded sample of the same batches;
* contrast  -- Erdos-Renyi, 4,194,304 vertices, mean degree 16: the same
               16-root BFS and the same 100-batch repair stream (adaptive
               selector, consolidation every 20; always-recompute beside);
* batching  -- R-MAT scale 20: 64 roots per hipGraph-captured multi-source
               pass (autotuned), then 256 update batches of log-uniform size
               10 .. 100,000, answered by two arms on two copies: apply and
               replay the pass, and the 64-root repair (captured repair
               programs, a pass where its cost model prefers one); the arms
               compared on all 64 roots and sampled roots checked bit-exact
               against the single-root path;
* pagerank  -- directed R-MAT scale 22: FP32 pull PageRank to an L1
               residual below 1e-7, then 50 batches of 10,000 ops answered
               by the residual-push repair and, on a second copy, by a
               warm-started FP32 re-solve; L1 / max-abs error of both
               against the FP64 solve on a third copy, per batch (that
               solve checked against the CPU oracle's on the first and last
               batch); then 10 batches each of 100 and 1,000 ops;
* ppr       -- directed weighted R-MAT scale 20, 32 one-hot teleport
               columns: weighted personalized PageRank forward + adjoint
               through torch.autograd, 100 sweeps each.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
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
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "vgx_ref")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=16)
    p.add_argument("--warmup", type=int, default=4,
                   help="untimed traversals from other roots of the same component before the timed ones")
    p.add_argument("--scale", type=int, default=22)
    p.add_argument("--edgefactor", type=int, default=16)
    p.add_argument("--roots", type=int, default=16)
    p.add_argument("--batched-roots", type=int, default=64)
    p.add_argument("--batches", type=int, default=100)
    p.add_argument("--batch-size", type=int, default=10000)
    p.add_argument("--consolidate-every", type=int, default=20)
    p.add_argument("--cpu-budget", type=float, default=20.0)
    p.add_argument("--pr-batches", type=int, default=50)
    p.add_argument("--batch-scale", type=int, default=20)
    p.add_argument("--uneven-batches", type=int, default=256)
    p.add_argument("--no-batching", action="store_true")
    p.add_argument("--ppr-scale", type=int, default=20)
    p.add_argument("--ppr-iterations", type=int, default=100)
    p.add_argument("--ppr-reps", type=int, default=5)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-repairs", action="store_true")
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--no-consolidation", action="store_true")
    p.add_argument("--no-static", action="store_true")
    p.add_argument("--no-kcore", action="store_true")
    p.add_argument("--cons-epochs", type=int, default=24)
    p.add_argument("--cons-batch", type=int, default=1000000)
    p.add_argument("--no-contrast", action="store_true")
    p.add_argument("--no-pagerank", action="store_true")
    p.add_argument("--no-ppr", action="store_true")
    p.add_argument("--only-bfs", action="store_true", help="plain BFS leg only (for profiling)")
    p.add_argument("--legs", default="",
                   help="comma list of secondary legs to run beside the plain BFS (batched, consolidation, static, "
                        "repairs, contrast, batching, kcore, pagerank, ppr); default all")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                   help="only run 5 device-to-device copies of BYTES (HBM counter calibration)")
    p.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-traffic passes")
    p.add_argument("--pmc-workload", choices=PMC_WORKLOADS, default="rmat", help=argparse.SUPPRESS)
    p.add_argument("--pmc-child", action="store_true",
                   help="(internal) profiled child: a 1 GiB copy, then the plain BFS roots")
    return p.parse_args()


# ----------------------------------------------------------------- CPU legs


def _edges_file(td, s, d):
    path = os.path.join(td, "edges.bin")
    np.stack([s, d], 1).astype(np.uint32).tofile(path)
    return path


def cpu_model():
    """The host CPU's model name (the "model name" line lscpu shows)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_threads():
    """Host threads for the all-core baseline: the CPU share this process may
    run on, at most 16 (one GPU's share of a box)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline_all_cores(s, d, roots, m_c, threads):
    """The reference's root-parallel BFS (benchmarks/thread_scaling_bfs.cpp:
    78-108): every thread takes the next root off a shared counter and runs
    bfs_distances on one CsrGraph; the 16 roots once each."""
    with tempfile.TemporaryDirectory() as td:
        cmd = [REF_BIN, "bench-bfs-mt", _edges_file(td, s, d), str(threads), "1"] + [str(int(r)) for r in roots]
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=900, check=True).stdout
    w = [line.split() for line in out.splitlines() if line.startswith("mt ")][0]
    queries, sec = int(w[2]), float(w[3])
    return {"value": m_c * queries / sec / 1e9, "unit": "GTEPS", "cores": threads, "kind": "reference",
            "sample": f"bfs_distances(CsrGraph, root) of the reference engine, the {queries} roots shared by "
                      f"{threads} threads (its thread-scaling benchmark), wall time",
            "ms_per_bfs_wall": 1e3 * sec / queries}


def cpu_baseline(s, d, n, roots, m_c, budget):
    """Time the reference engine's BFS on a bounded sample: all host threads
    root-parallel (the reported value) and one thread."""
    times = []
    if os.path.exists(REF_BIN):
        with tempfile.TemporaryDirectory() as td:
            cmd = [REF_BIN, "bench-bfs", _edges_file(td, s, d), str(n), "0", str(budget)]
            cmd += [str(int(r)) for r in roots]
            out = subprocess.run(cmd, capture_output=True, text=True, timeout=900, check=True).stdout
        for line in out.splitlines():
            w = line.split()
            if w and w[0] == "root":
                times.append(float(w[2]))
        kind = "reference"
        sample = (f"IncrementalBFS(graph, root) construction (a full queue BFS) of the reference engine on the "
                  f"same graph, first {len(times)} of the same roots, one thread")
    else:
        from oracle import oracle as orc

        orc.build()
        g = orc.ModelGraph(n, False)
        g.bulk_load_edges(np.stack([s, d], 1))
        off, adj = g.csr()
        t_all = time.perf_counter()
        for r in roots:
            t0 = time.perf_counter()
            orc.bfs(n, off, adj, int(r))
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_all > budget:
                break
        kind = "port"
        sample = f"oracle queue BFS (C restatement) on the same graph, first {len(times)} of the same roots"
    if not times:
        return None
    one = {"value": m_c * len(times) / sum(times) / 1e9, "unit": "GTEPS", "cores": 1, "kind": kind,
           "sample": sample, "ms_per_bfs": 1e3 * sum(times) / len(times)}
    if kind != "reference":
        one["cpu_model"] = cpu_model()
        return one
    out = cpu_baseline_all_cores(s, d, roots, m_c, cpu_threads())
    out["cpu_model"] = cpu_model()
    out["single_thread"] = one
    return out


def cpu_repairs(s, d, n, root, batches, budget):
    """The reference engine's IncrementalBFS::apply on the first batches of
    the same stream (one thread, bounded by ``budget`` seconds)."""
    if not os.path.exists(REF_BIN):
        return None
    with tempfile.TemporaryDirectory() as td:
        bpath = os.path.join(td, "batches.bin")
        with open(bpath, "wb") as f:
            for bs, bd, ba in batches:
                np.array([bs.size], np.uint32).tofile(f)
                np.stack([bs, bd, ba.astype(np.uint32)], 1).astype(np.uint32).tofile(f)
        cmd = [REF_BIN, "bench-repair", _edges_file(td, s, d), str(n), "0", str(int(root)), bpath, str(budget)]
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=1200, check=True).stdout
    t = [float(line.split()[1]) for line in out.splitlines() if line.startswith("batch ")]
    full = sum(int(line.split()[3]) for line in out.splitlines() if line.startswith("batch "))
    if not t:
        return None
    return {"repairs_per_s": len(t) / sum(t), "mean_batch_ms": 1e3 * sum(t) / len(t), "batches": len(t),
            "full_recompute_batches": full, "cores": 1, "kind": "reference",
            "sample": f"IncrementalBFS::apply of the reference engine on the first {len(t)} batches"}


def cpu_pagerank(s, d, n, iterations):
    if not os.path.exists(REF_BIN):
        return None
    with tempfile.TemporaryDirectory() as td:
        cmd = [REF_BIN, "bench-pr", _edges_file(td, s, d), str(n), "1", str(iterations)]
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=900, check=True).stdout
    w = [line.split() for line in out.splitlines() if line.startswith("pr ")][0]
    secs, its = float(w[1]), int(w[2])
    return {"ms_per_iteration": 1e3 * secs / max(its, 1), "iterations": its, "cores": 1, "kind": "reference",
            "sample": f"IncrementalPageRank::recompute of the reference engine (FP64), {its} sweeps"}


# ------------------------------------------------------ HBM traffic (PMC)

COPY_BYTES = 1 << 30


def _counter_rows(directory):
    import csv

    for root, _, files in os.walk(directory):
        for f in files:
            if f.endswith("counter_collection.csv"):
                with open(os.path.join(root, f)) as fh:
                    return sorted(csv.DictReader(fh), key=lambda r: int(r["Dispatch_Id"]))
    return []


def _per_bfs(rows):
    """Counter sums (KiB) of the kernels of each traversal (the init and the
    step kernels), split at k_bfs_init (one per traversal)."""
    groups = []
    for r in rows:
        name = r["Kernel_Name"]
        if "k_bfs_init" in name:
            groups.append(float(r["Counter_Value"]))
        elif groups and "k_bfs_step" in name:
            groups[-1] += float(r["Counter_Value"])
    return groups


def _per_sweep(rows, with_finish):
    """Counter values (KiB) of each FP32 PageRank sweep: k_pr_iter<float>,
    plus the k_pr_finish<float> launch that follows it when with_finish."""
    groups = []
    for r in rows:
        name = r["Kernel_Name"]
        if "k_pr_iter<float>" in name:
            groups.append(float(r["Counter_Value"]))
        elif with_finish and groups and "k_pr_finish<float>" in name:
            groups[-1] += float(r["Counter_Value"])
    return groups


# Profiled workloads: what the child runs and how its dispatches group into
# units (one traversal, one sweep).
PMC_WORKLOADS = ("rmat", "er", "pagerank")


def measure_traffic(args, steps, workload="rmat"):
    """HBM bytes per unit of a workload from two rocprofv3 --pmc passes
    (FETCH_SIZE, WRITE_SIZE) over a child run of this script: per traversal
    of the headline BFS ("rmat") or of the Erdos-Renyi BFS ("er"), per FP32
    PageRank sweep ("pagerank", k_pr_iter<float> alone and with its
    k_pr_finish<float>).  Each counter is scaled by its own reading for a
    1 GiB device copy made first in the same child, i.e. bytes = counter *
    (2^30 / counter_of_copy); the median over the child's last `steps` units
    is reported."""
    import shutil
    import signal

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        with tempfile.TemporaryDirectory(dir="/tmp") as td:
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", td, "-o", "p", "--", sys.executable,
                   os.path.abspath(__file__), "--pmc-child", "--pmc-workload", workload, "--scale", str(args.scale),
                   "--edgefactor", str(args.edgefactor), "--roots", str(args.roots), "--steps", str(steps)]
            env = dict(os.environ, TMPDIR="/tmp")
            proc = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                                    start_new_session=True)
            try:
                _, err = proc.communicate(timeout=300)
            except subprocess.TimeoutExpired:
                os.killpg(proc.pid, signal.SIGKILL)
                proc.communicate()
                log(f"traffic pass {workload} {counter} timed out")
                return None
            rows = _counter_rows(td)
        copies = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
        units = {"unit": _per_sweep(rows, True), "k_pr_iter": _per_sweep(rows, False)} if workload == "pagerank" \
            else {"unit": _per_bfs(rows)}
        if not copies or not units["unit"]:
            log(f"traffic pass {workload} {counter}: no data (rc={proc.returncode})")
            return None
        # the calibration copy is by far the largest copy dispatch of the child
        scale = (COPY_BYTES / 1024.0) / max(copies)
        key = "bytes_per_sweep" if workload == "pagerank" else "bytes_per_bfs"
        rec = {"copy_scale": scale}
        for name, groups in units.items():
            med = float(np.median(groups[-steps:]))
            label = key if name == "unit" else "bytes_per_k_pr_iter"
            rec[label] = med * 1024.0 * scale
            if name == "unit":
                rec["kib_per_unit"] = med
        out[counter] = rec
    return out


def pmc_child(args):
    """Profiled child: a 1 GiB device copy (counter calibration), then the
    workload as its timed region runs it -- the headline BFS (warm-up roots,
    then one traversal per root by a re-targeted object), the same on the
    Erdos-Renyi graph, or FP32 PageRank solves on directed R-MAT."""
    import torch

    import velographx_amd as vx
    from velographx_amd import workloads as wl

    a = torch.ones(COPY_BYTES, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    n = 1 << args.scale
    if args.pmc_workload == "pagerank":
        s, d = vx.rmat_edges(args.scale, args.edgefactor, 1)
        g = vx.Graph(n, True)
        g.bulk_load_arrays(s, d)
        pr = vx.IncrementalPageRank(g, 0.85, 32)
        pr.solve_l1(200, 1e-7, False)
        pr.solve_l1(200, 1e-7, False)
        torch.cuda.synchronize()
        return
    if args.pmc_workload == "er":
        n = 1 << 22
        s, d = wl.erdos_renyi(n, 8 * n, seed=1)
    else:
        s, d = vx.rmat_edges(args.scale, args.edgefactor, 1)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    lcc, _ = component(vx, wl, g)
    roots = wl.roots_in_component(lcc, args.roots, seed=1)
    warm = warmup_roots(wl, lcc, roots, 4)
    walker = vx.IncrementalBFS(g, int(warm[0]))
    for r in warm:
        walker.set_source(int(r))
    for k in range(args.steps):
        walker.set_source(int(roots[k % len(roots)]))
    torch.cuda.synchronize()


# ----------------------------------------------------------------- GPU legs


def timed_bfs(vx, g, roots, steps, warmup, warm_roots):
    """K answer-ready BFS traversals, step k from roots[k % len(roots)], by
    one traversal object re-targeted to the step's root (set_source), after
    `warmup` untimed traversals from other roots (warm_roots).  Nothing is
    kept per root: the launch program is steered on the device by the
    traversal's own counters, so the first traversal from a root is the one
    timed, and a root traversed again runs the same program.  Returns a dict
    with the wall seconds per step, the elapsed wall time, the GPU ms per
    traversal (HIP events on the graph's stream around each one), the step
    launches per traversal and their mean duration (a separate pass with
    events around every launch), the GPU ms of the same roots traversed a
    second time in a row, and the traversal object."""
    import torch

    walker = vx.IncrementalBFS(g, int(warm_roots[0]))
    for k in range(warmup):
        walker.set_source(int(warm_roots[k % len(warm_roots)]))
    walker.set_profiling(2)  # GPU time of every whole traversal; resets the counters
    torch.cuda.synchronize()
    times = []
    t_start = time.perf_counter()
    for k in range(steps):
        r = int(roots[k % len(roots)])
        t0 = time.perf_counter()
        walker.set_source(r)
        times.append(time.perf_counter() - t0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    kernel_ms = walker.stats()["traversal_ms"] / steps
    # the same roots again, each traversed twice in a row (the second timed):
    # equal to the first traversal within noise, since nothing is kept
    walker.set_profiling(0)
    again = []
    for r in roots:
        walker.set_source(int(r))
        walker.set_profiling(2)
        walker.recompute()
        again.append(walker.stats()["traversal_ms"])
        walker.set_profiling(0)
    # step launches per traversal and their mean duration
    walker.set_profiling(1)
    for r in roots:
        walker.set_source(int(r))
    torch.cuda.synchronize()
    st = walker.stats()
    launches = st["level_launches"] / len(roots)
    launch_ms = st["level_kernel_ms"] / max(st["level_launches"], 1)
    walker.set_profiling(0)
    return {"times": times, "elapsed": elapsed, "kernel_ms": kernel_ms, "launches": launches,
            "launch_ms": launch_ms, "repeat_ms": float(np.mean(again)), "walker": walker}


def warmup_roots(wl, lcc, roots, count):
    """Roots of the same component for the untimed warm-up, none of them a
    timed root."""
    cand = wl.roots_in_component(lcc, len(roots) + count + 16, seed=7)
    timed = set(int(r) for r in roots)
    out = [int(r) for r in cand if int(r) not in timed]
    return out[:max(count, 1)]


def component(vx, wl, g):
    lcc = wl.largest_component_from_hub(vx, g)
    deg = g.degrees()
    arcs_c = int(deg[lcc].astype(np.int64).sum())
    return lcc, arcs_c


def leg_batched(vx, wl, g, lcc, m_c, k, plain_dist, plain_gteps):
    import torch

    roots = wl.roots_in_component(lcc, k, seed=1)
    ms = vx.MultiSourceBFS(g, roots.tolist())
    tuned = ms.autotune(2)  # outside the timed region
    log(f"batched autotune: {tuned}")
    ms.run(g)  # warm-up
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        ms.run(g)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    exact = all(np.array_equal(ms.distances(i), x) for i, x in enumerate(plain_dist[:k]))
    reached = ms.reached()
    agg = m_c * k / t / 1e9
    out = {"roots": k, "ms_per_pass": 1e3 * t, "aggregate_gteps": agg, "batched_vs_plain": agg / plain_gteps,
           "tuning": {k: tuned[k] for k in ("blocks", "alpha", "beta", "short_row", "chunk")},
           "levels": int(ms.level_count), "exact_vs_plain": bool(exact),
           "all_roots_reach_component": bool(np.all(reached == lcc.size))}
    log(f"batched: {out}")
    return out


def leg_batching(vx, wl, args):
    """R-MAT scale 20: 64 roots per captured multi-source pass, then 256
    update batches of log-uniform size (10 .. 100,000 ops, 80/20).  Two arms
    over the same batches, each on its own copy of the graph: the
    recompute arm applies the batch and replays the captured pass; the
    repair arm (MultiSourceBFS.apply) repairs every root in captured repair
    programs or, when its cost model prefers it, replays the pass.  Both
    arms' 64 roots are compared after every 32nd batch, and 8 roots against
    the single-root path."""
    import torch

    scale = args.batch_scale
    n = 1 << scale
    s, d = vx.rmat_edges(scale, args.edgefactor, 1)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    g2 = vx.Graph(n, False)
    g2.bulk_load_arrays(s, d)
    lcc, arcs_c = component(vx, wl, g)
    roots = wl.roots_in_component(lcc, 64, seed=1).tolist()
    ms = vx.MultiSourceBFS(g, roots)
    tuned = ms.autotune(2)
    ms.set_graph(True)
    ms.run(g)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        ms.run(g)
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / reps
    ms.set_graph(False)
    ms.run(g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ms.run(g)
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / reps
    ms.set_graph(True)
    ms2 = vx.MultiSourceBFS(g2, roots)
    ms2.set_params(tuned["blocks"], tuned["alpha"], tuned["beta"], tuned["short_row"], tuned["chunk"])
    ms2.set_graph(True)
    ms2.set_repair(0, 0.1)
    ms2.run(g2)
    sizes = wl.log_uniform_sizes(args.uneven_batches, 1.0, 5.0, seed=1)
    eu, ev = wl.canonical_edges(s, d, False)
    stream = wl.BatchStream(eu, ev, n, False, seed=3)
    batches = [stream.next(int(k), 0.8) for k in sizes]
    t_apply, t_pass, t_rep = [], [], []
    repaired = []
    exact = True
    arms_equal = True
    checked = 0
    singles = [vx.IncrementalBFS(g, int(r)) for r in roots[:8]]
    for i, (bs, bd, ba) in enumerate(batches):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.apply_arrays(bs, bd, ba)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms.run(g)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        ms2.apply_arrays(bs, bd, ba)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        t_apply.append(t1 - t0)
        t_pass.append(t2 - t1)
        t_rep.append(t3 - t2)
        repaired.append(bool(ms2.repair_stats()["last_repaired"]))
        if i % 32 == 31 or i == len(batches) - 1:  # verification outside the timers
            for j, o in enumerate(singles):
                o.recompute()
                exact &= bool(np.array_equal(ms.distances(j), o.distances_numpy()))
            arms_equal &= bool(np.array_equal(ms.levels_u8(), ms2.levels_u8()))
            checked += 1
    ops = int(sum(int(k) for k in sizes))
    t_rec = np.array(t_apply) + np.array(t_pass)
    t_rep = np.array(t_rep)
    rep_mask = np.array(repaired)
    by_size = {}
    for lo, hi in ((10, 100), (100, 1000), (1000, 10000), (10000, 100001)):
        sel = (sizes >= lo) & (sizes < hi)
        if sel.any():
            by_size[f"{lo}-{hi - 1 if hi > 100000 else hi}"] = {
                "batches": int(sel.sum()), "repaired": int((rep_mask & sel).sum()),
                "recompute_ms": 1e3 * float(np.mean(t_rec[sel])), "repair_arm_ms": 1e3 * float(np.mean(t_rep[sel]))}
    st = ms2.repair_stats()
    out = {"workload": f"undirected R-MAT s{scale} ef{args.edgefactor} seed 1, 64 LCC roots per captured pass; "
                       f"{len(batches)} batches of log-uniform size 10..100000 (80/20)",
           "arcs": int(g.edge_count_directed), "pass_ms_graph": 1e3 * t_graph, "pass_ms_eager": 1e3 * t_eager,
           "aggregate_gteps_graph": arcs_c / 2.0 * 64 / t_graph / 1e9,
           "tuning": {k: tuned[k] for k in ("blocks", "alpha", "beta", "short_row", "chunk")},
           "batches": len(batches), "update_ops": ops,
           "recompute_arm": {"mean_apply_ms": 1e3 * float(np.mean(t_apply)),
                             "mean_pass_ms": 1e3 * float(np.mean(t_pass)),
                             "batches_per_s": len(batches) / float(t_rec.sum())},
           "repair_arm": {"batches_per_s": len(batches) / float(t_rep.sum()),
                          "mean_batch_ms": 1e3 * float(t_rep.mean()),
                          "repaired_batches": int(rep_mask.sum()), "full_pass_batches": int(st["full_passes"]),
                          "fallbacks": int(st["fallbacks"]), "policy": "per-size-class running means (mode 0)"},
           "by_batch_size": by_size,
           "batches_per_s": len(batches) / float(t_rep.sum()),
           "repair_vs_recompute": float(t_rec.sum() / t_rep.sum()),
           "graph_replays": int(ms.stats()["graph_replays"]),
           "exact_vs_single_root": exact, "arms_equal_all_64_roots": arms_equal, "verified_batches": checked,
           "verified_roots_per_check": len(singles)}
    log(f"batching: {out}")
    return out


def cpu_snapshot(s, d, n, repeats=2):
    """The reference engine's consolidate_to_csr_snapshot and its 4096-row
    neighbour probe on the same graph (one thread)."""
    if not os.path.exists(REF_BIN):
        return None
    with tempfile.TemporaryDirectory() as td:
        cmd = [REF_BIN, "bench-snapshot", _edges_file(td, s, d), str(n), "0", str(repeats)]
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=900, check=True).stdout
    snaps = [float(line.split()[1]) for line in out.splitlines() if line.startswith("snapshot ")]
    probe = [float(line.split()[1]) for line in out.splitlines() if line.startswith("probe ")]
    if not snaps:
        return None
    return {"snapshot_ms": 1e3 * min(snaps), "probe_ns": probe[0] if probe else None, "cores": 1,
            "kind": "reference",
            "sample": f"consolidate_to_csr_snapshot of the reference engine on the bulk-loaded graph "
                      f"(best of {len(snaps)}), then its 4096-row neighbour probe"}


def leg_consolidation(vx, wl, g, s, d, n, args):
    """Consolidation at the BFS graph's scale: the device snapshot of the
    bulk-loaded graph (timed; algorithmic bytes = row table, degrees and
    arena read, packed arena and row table, capacities and degrees
    written), the on-device neighbour probe (4096 rows, the reference
    benchmark's default), then a controller-driven steady state on a
    snapshot copy: per epoch one update batch (80% inserts of random pairs,
    20% deletions of loaded edges), the storage and
    probe-latency signal under the reference benchmark's default policy
    (1.25 / 1.25), the controller's decision and, on a trigger, a snapshot
    cut over in place of the graph."""
    import torch

    st = g.storage()
    arcs = int(g.edge_count_directed)
    nv = int(g.vertex_count)
    stored = int(st["arena_used"])
    snap_bytes = 12.0 * nv + 4.0 * stored + 4.0 * arcs + 16.0 * nv
    vx.consolidate_to_csr_snapshot(g)  # warm-up: allocator pools
    t_snap = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        snap = vx.consolidate_to_csr_snapshot(g)
        t_snap.append(time.perf_counter() - t0)
        del snap
    t_best = min(t_snap)
    probe_ns, _ = g.probe_neighbors(4096, 5)
    keep = vx.consolidate_to_csr_snapshot(g)
    h = keep.graph
    exact = h.probe_neighbors(nv, 1)[1] == g.probe_neighbors(nv, 1)[1]
    rng = np.random.default_rng(31)
    n_ins = int(0.8 * args.cons_batch)
    n_del = args.cons_batch - n_ins
    policy = vx.ConsolidationPolicy(1.25, 1.25)
    ctl = vx.ConsolidationController()
    canonical_bytes = h.storage_bytes()
    canonical_ns, _ = h.probe_neighbors(4096, 5)
    t_apply, t_cut = [], []
    triggers = {"storage": 0, "latency": 0}
    hw_storage, hw_latency = 1.0, 1.0
    for epoch in range(1, args.cons_epochs + 1):
        # inserts of uniform random pairs, deletions of bulk-loaded edges
        pick = rng.integers(0, s.size, n_del)
        bs = np.concatenate([rng.integers(0, nv, n_ins).astype(np.uint32), s[pick]])
        bd = np.concatenate([rng.integers(0, nv, n_ins).astype(np.uint32), d[pick]])
        ba = np.concatenate([np.ones(n_ins, np.uint8), np.zeros(n_del, np.uint8)])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.apply_arrays(bs, bd, ba)
        torch.cuda.synchronize()
        t_apply.append(time.perf_counter() - t0)
        ns, _ = h.probe_neighbors(4096, 3)
        sig = vx.evaluate_consolidation(h.storage_bytes(), canonical_bytes, ns, canonical_ns, policy)
        hw_storage = max(hw_storage, sig.storage_growth_ratio)
        hw_latency = max(hw_latency, sig.neighbor_latency_ratio)
        if ctl.observe(sig, epoch):
            triggers["storage" if sig.storage_limit_exceeded else "latency"] += 1
            t0 = time.perf_counter()
            nxt = vx.consolidate_to_csr_snapshot(h)
            t_cut.append(time.perf_counter() - t0)
            exact &= nxt.graph.probe_neighbors(nv, 1)[1] == h.probe_neighbors(nv, 1)[1]
            keep, h = nxt, nxt.graph
            canonical_bytes = h.storage_bytes()
            canonical_ns, _ = h.probe_neighbors(4096, 5)
            ctl.mark_consolidated(epoch)
    out = {"workload": f"undirected R-MAT s{args.scale} ef{args.edgefactor} seed 1: device snapshot of the "
                       f"bulk-loaded graph; {args.cons_epochs} epochs of {args.cons_batch}-op batches (80% random "
                       f"inserts, 20% deletions of loaded edges) "
                       f"under the consolidation controller (policy 1.25/1.25)",
           "arcs": arcs, "snapshot_ms": 1e3 * t_best, "snapshot_bytes": snap_bytes,
           "snapshot_gbps": snap_bytes / t_best / 1e9, "snapshot_roofline_frac": snap_bytes / t_best / 1e9 / HBM_PEAK_GBPS,
           "probe_ns_4096": probe_ns, "epochs": args.cons_epochs,
           "mean_apply_ms": 1e3 * float(np.mean(t_apply)), "consolidations": len(t_cut),
           "storage_triggers": triggers["storage"], "latency_triggers": triggers["latency"],
           "mean_cutover_ms": 1e3 * float(np.mean(t_cut)) if t_cut else None,
           "high_water_storage_ratio": hw_storage, "high_water_latency_ratio": hw_latency,
           "digest_equal_after_every_cutover": bool(exact)}
    del keep, h
    log(f"consolidation: {out}")
    return out


def leg_static(vx, s, d, root):
    """The static CsrGraph of the same edges: construction from the edge
    array (host copy, upload, device build), bfs_distances from one root
    (answer returned to the host, as the reference's vector), and exactly 20
    pagerank sweeps (tolerance 0)."""
    import torch

    # the (N, 2) edge array is the input, as the reference's edge vector is
    # built before its timer starts
    edges = np.stack([s, d], 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = vx.CsrGraph(edges, False)
    t_build = time.perf_counter() - t0
    del edges
    vx.bfs_distances(g, int(root))
    t_bfs = []
    for _ in range(3):
        t0 = time.perf_counter()
        dist = vx.bfs_distances(g, int(root))
        t_bfs.append(time.perf_counter() - t0)
    reached = int((dist != 0xFFFFFFFF).sum())
    vx.pagerank(g, 0.85, 2, 0.0)
    t0 = time.perf_counter()
    pr = vx.pagerank(g, 0.85, 20, 0.0)
    t_pr = time.perf_counter() - t0
    out = {"workload": "static CsrGraph of the same R-MAT edges: build, bfs_distances from the first root, "
                       "pagerank(0.85, 20 sweeps, tolerance 0)",
           "vertices": int(g.vertex_count()), "entries": int(g.edge_entry_count()),
           "build_ms": 1e3 * t_build, "bfs_distances_ms": 1e3 * min(t_bfs), "reached": reached,
           "pagerank_ms_per_sweep": 1e3 * t_pr / 20, "rank_sum": float(pr.sum())}
    del g
    log(f"static: {out}")
    return out


def leg_kcore(vx, args, with_cpu):
    """Core numbers and connected components of undirected R-MAT at the
    batching scale (s20): the device peeling and the device labelling of the
    bulk-loaded graph (best of 3 each), and the reference's IncrementalKCore
    and IncrementalComponents constructions on the same edges."""
    import torch

    scale = args.batch_scale
    n = 1 << scale
    s, d = vx.rmat_edges(scale, args.edgefactor, 1)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(s, d)
    kc = vx.IncrementalKCore(g)  # warm-up
    t = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kc.recompute()
        t.append(time.perf_counter() - t0)
    core = kc.core_numpy()
    st = kc.stats()
    del kc
    cc = vx.IncrementalComponents(g)  # warm-up
    tc = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cc.recompute()
        tc.append(time.perf_counter() - t0)
    labels = cc.labels_numpy()
    ncomp = int((labels == np.arange(labels.size, dtype=np.uint32)).sum())
    cc_rounds = int(cc.stats()["rounds"])
    del cc
    out = {"workload": f"undirected R-MAT s{scale} ef{args.edgefactor} seed 1: all core numbers",
           "arcs": int(g.edge_count_directed), "ms": 1e3 * min(t), "max_core": int(core.max()),
           "levels": int(st["levels"]), "rounds": int(st["rounds"]),
           "arcs_per_s": g.edge_count_directed / min(t),
           "components": {"ms": 1e3 * min(tc), "count": ncomp, "rounds": cc_rounds}}
    if with_cpu and os.path.exists(REF_BIN):
        with tempfile.TemporaryDirectory() as td:
            cmd = [REF_BIN, "bench-kcore", _edges_file(td, s, d), str(n)]
            res = subprocess.run(cmd, capture_output=True, text=True, timeout=900, check=True).stdout
        w = [line.split() for line in res.splitlines() if line.startswith("kcore ")][0]
        out["cpu_reference"] = {"ms": 1e3 * float(w[1]), "max_core": int(w[2]), "cores": 1, "kind": "reference",
                                "sample": "IncrementalKCore(graph) construction of the reference engine on the "
                                          "same edges"}
        out["max_core_equal"] = int(w[2]) == out["max_core"]
        # every core number, by the digest of the whole vector
        out["core_numbers_equal"] = int(w[3]) == vertex_digest(core)
        with tempfile.TemporaryDirectory() as td:
            cmd = [REF_BIN, "bench-cc", _edges_file(td, s, d), str(n)]
            res = subprocess.run(cmd, capture_output=True, text=True, timeout=900, check=True).stdout
        w = [line.split() for line in res.splitlines() if line.startswith("cc ")][0]
        out["components"]["cpu_reference"] = {"ms": 1e3 * float(w[1]), "count": int(w[2]), "cores": 1,
                                              "kind": "reference",
                                              "sample": "IncrementalComponents(graph) construction of the "
                                                        "reference engine on the same edges"}
        # the partition, by the digest of each vertex's smallest component member
        out["components"]["partition_equal"] = int(w[2]) == ncomp and int(w[3]) == vertex_digest(labels)
    del g
    log(f"kcore: {out}")
    return out


def vertex_digest(values):
    """The reference driver's order-free digest of a per-vertex array: the
    sum mod 2^64 of splitmix64's output function of (v << 32 | value[v])."""
    v = np.arange(values.size, dtype=np.uint64)
    x = (v << np.uint64(32)) | values.astype(np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
        return int(np.sum(x, dtype=np.uint64))


def cpu_static(s, d, root):
    """The reference's CsrGraph construction, bfs_distances and 3 pagerank
    sweeps on the same edges (one thread)."""
    if not os.path.exists(REF_BIN):
        return None
    with tempfile.TemporaryDirectory() as td:
        cmd = [REF_BIN, "bench-static", _edges_file(td, s, d), "0", str(int(root)), "3"]
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=900, check=True).stdout
    w = [line.split() for line in out.splitlines() if line.startswith("static ")][0]
    return {"build_ms": 1e3 * float(w[1]), "bfs_distances_ms": 1e3 * float(w[2]),
            "pagerank_ms_per_sweep": 1e3 * float(w[3]) / 3, "cores": 1, "kind": "reference",
            "sample": "CsrGraph(edges) construction, bfs_distances from the same root and pagerank(0.85, 3 sweeps, "
                      "tolerance 0) of the reference engine"}


def _run_repairs(vx, g, root, batches, args, policy):
    """One pass over the update stream with the given arm policy:
    "adaptive" (the selector picks repair or recompute per batch, the
    reference harness's set-up with the in-object fallback disabled),
    "incremental" (always the localized repair, in-object fallback 0.35),
    "incremental_device" (the same with every batch uploaded to HBM before
    the timer and applied from there) or "full" (graph apply + full
    recompute).  Consolidation every
    `consolidate_every` batches inside the timer; verification against a
    fresh BFS every 10 batches outside it."""
    import collections

    import torch

    inc = vx.IncrementalBFS(g, int(root), 2.0 if policy == "adaptive" else 0.35)
    sel = vx.AdaptiveSelector(inc) if policy == "adaptive" else None
    dev = None
    if policy == "incremental_device":  # batches resident in HBM before the timer
        dev = [(torch.from_numpy(bs.view(np.int32)).cuda(), torch.from_numpy(bd.view(np.int32)).cuda(),
                torch.from_numpy(ba.astype(np.uint8)).cuda()) for bs, bd, ba in batches]
        torch.cuda.synchronize()
    batch_t, reasons = [], collections.Counter()
    full = 0
    exact = True
    for i, (bs, bd, ba) in enumerate(batches):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if policy == "adaptive":
            tr = sel.apply_arrays(bs, bd, ba)
        elif policy == "incremental":
            inc.apply_arrays(bs, bd, ba)
        elif policy == "incremental_device":
            inc.apply_tensors(*dev[i])
        else:
            g.apply_arrays(bs, bd, ba)
            inc.recompute()
        if args.consolidate_every and (i + 1) % args.consolidate_every == 0:
            g.compact()
        batch_t.append(time.perf_counter() - t0)
        if policy == "adaptive":
            reasons[tr["reason"]] += 1
            full += int(tr["chose_full"] or tr["internal_fallback"])
        elif policy in ("incremental", "incremental_device"):
            full += int(inc.last_used_full_recompute)
        else:
            full += 1
        if i % 10 == 9 or i == len(batches) - 1:
            ver = vx.IncrementalBFS(g, int(root))
            exact &= bool(np.array_equal(ver.distances_numpy(), inc.distances_numpy()))
            del ver
    tot = sum(batch_t)
    out = {"repairs_per_s": len(batch_t) / tot, "update_ops_per_s": len(batch_t) * args.batch_size / tot,
           "mean_batch_ms": 1e3 * tot / len(batch_t), "p50_batch_ms": 1e3 * float(np.median(batch_t)),
           "full_recompute_batches": full, "exact": exact}
    if reasons:
        out["decisions"] = dict(reasons)
    return out


def leg_repairs(vx, wl, g, s, d, n, root, args):
    t0 = time.time()
    eu, ev = wl.canonical_edges(s, d, False)
    stream = wl.BatchStream(eu, ev, n, False, seed=1)
    batches = [stream.next(args.batch_size, 0.8) for _ in range(args.batches)]
    log(f"generated {len(batches)} batches in {time.time() - t0:.1f}s")
    out = _run_repairs(vx, g, root, batches, args, "adaptive")
    out["policy"] = "adaptive selector (publication-preflight-v1), consolidation every %d" % args.consolidate_every
    out["consolidate_every"] = args.consolidate_every
    # the two fixed arms on fresh copies of the same graph, for comparison
    # ("incremental_device": the same repairs with each batch already in HBM)
    for policy in ("incremental", "incremental_device", "full"):
        g2 = vx.Graph(n, False)
        g2.bulk_load_arrays(s, d)
        r = _run_repairs(vx, g2, root, batches, args, policy)
        out[f"always_{policy}"] = {k: r[k] for k in ("repairs_per_s", "mean_batch_ms", "full_recompute_batches",
                                                     "exact")}
        del g2
    log(f"repairs: {out}")
    return out, batches


def leg_contrast(vx, wl, steps, args):
    """Erdos-Renyi contrast at the plain workload's size: the BFS from 16
    roots of its largest component, then the same update stream as the
    repairs leg (100 batches of 10,000 ops, 80/20, consolidation every 20)
    answered through the adaptive selector, with the always-recompute arm
    on a copy for comparison."""
    n = 1 << 22
    t0 = time.time()
    u, v = wl.erdos_renyi(n, 8 * n, seed=1)
    g = vx.Graph(n, False)
    g.bulk_load_arrays(u, v)
    lcc, arcs_c = component(vx, wl, g)
    roots = wl.roots_in_component(lcc, 16, seed=1)
    log(f"ER n={n}: {g.edge_count_directed} arcs, LCC {lcc.size}, built in {time.time() - t0:.1f}s")
    t_bfs = timed_bfs(vx, g, roots, steps, 4, warmup_roots(wl, lcc, roots, 4))
    times, elapsed, kernel_ms = t_bfs["times"], t_bfs["elapsed"], t_bfs["kernel_ms"]
    bytes_bfs = 4.0 * arcs_c + 12.0 * lcc.size
    st = t_bfs["walker"].stats()
    out = {"workload": "Erdos-Renyi n=4194304, 33554432 uniform pairs (mean degree 16), seed 1",
           "arcs": int(g.edge_count_directed), "lcc_vertices": int(lcc.size),
           "gteps": arcs_c / 2.0 * steps / sum(times) / 1e9, "ms_per_bfs": 1e3 * elapsed / steps,
           "gpu_ms_per_bfs": kernel_ms, "bytes_per_bfs": bytes_bfs,
           "roofline_frac": bytes_bfs / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
           "levels": st["levels"], "bottom_up_levels": st["bottom_up_levels"], "steps": st["steps"]}
    del t_bfs
    if args.batches:
        eu, ev = wl.canonical_edges(u, v, False)
        stream = wl.BatchStream(eu, ev, n, False, seed=1)
        batches = [stream.next(args.batch_size, 0.8) for _ in range(args.batches)]
        rep = _run_repairs(vx, g, int(roots[0]), batches, args, "adaptive")
        rep["policy"] = "adaptive selector (publication-preflight-v1), consolidation every %d" % args.consolidate_every
        g2 = vx.Graph(n, False)
        g2.bulk_load_arrays(u, v)
        r = _run_repairs(vx, g2, int(roots[0]), batches, args, "full")
        rep["always_full"] = {k: r[k] for k in ("repairs_per_s", "mean_batch_ms", "full_recompute_batches", "exact")}
        del g2
        out["repairs"] = rep
    log(f"contrast: {out}")
    del g
    return out


def leg_pagerank(vx, wl, args, no_cpu):
    """Directed R-MAT: the FP32 solve, then the update stream answered two
    ways on two copies of the graph -- the residual-push repair (apply_push:
    localized pushes, warm full sweeps past its push budget) and the warm
    re-solve (full sweeps from the previous ranks) -- each checked per batch
    against the FP64 solve on a third copy, which is itself checked against
    the CPU oracle's FP64 solve on the first and last batch."""
    import torch

    scale = args.scale
    n = 1 << scale
    t0 = time.time()
    s, d = vx.rmat_edges(scale, args.edgefactor, 1)
    g32 = vx.Graph(n, True)
    g32.bulk_load_arrays(s, d)
    gpush = vx.Graph(n, True)
    gpush.bulk_load_arrays(s, d)
    g64 = vx.Graph(n, True)
    g64.bulk_load_arrays(s, d)
    arcs = int(g32.edge_count_directed)
    log(f"PageRank graph: directed R-MAT s{scale}, {arcs} arcs, built in {time.time() - t0:.1f}s")
    pr32 = vx.IncrementalPageRank(g32, 0.85, 32)
    prp = vx.IncrementalPageRank(gpush, 0.85, 32)
    pr64 = vx.IncrementalPageRank(g64, 0.85, 64)
    pr32.solve_l1(200, 1e-7, False)  # warm-up (class lists, allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pr32.solve_l1(200, 1e-7, False)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    iters = int(pr32.last_repair_iterations)
    prp.solve_l1(200, 1e-7, False)
    # the push arm keeps the exact residual of its ranks: one sweep builds it
    # (a self-loop batch changes no arc)
    prp.apply_push_arrays(np.array([0], np.uint32), np.array([0], np.uint32), np.array([1], np.uint8))
    pr64.solve_l1(1000, 1e-11, False)
    errs = {"warm": ([], []), "push": ([], [])}

    def errors():
        b = pr64.values_numpy()
        for key, pr in (("warm", pr32), ("push", prp)):
            a = pr.values_numpy()
            errs[key][0].append(float(np.abs(a - b).sum()))
            errs[key][1].append(float(np.abs(a - b).max()))

    errors()
    bytes_it = 8.0 * arcs + 20.0 * n
    out = {"workload": f"directed R-MAT s{scale} ef{args.edgefactor} seed 1, d=0.85, FP32, L1 residual < 1e-7",
           "arcs": arcs, "solve_ms": 1e3 * t_solve, "iterations": iters, "bytes_per_iteration": bytes_it,
           "ms_per_iteration": 1e3 * t_solve / max(iters, 1),
           "gbps_per_iteration_wall": bytes_it * iters / t_solve / 1e9,
           "roofline_frac_wall": bytes_it * iters / t_solve / 1e9 / HBM_PEAK_GBPS,
           "residual_l1": float(pr32.last_residual_l1)}
    oracle_check = []

    pending = []

    def cpu_check(model):
        from oracle import oracle as orc

        if pending:  # the batches since the last check, in order (the last op on an arc wins)
            model.apply(*(np.concatenate(x) for x in zip(*pending)))
            pending.clear()
        ref, _ = orc.pagerank_full(model, 0.85, 150, 1e-15)
        b = pr64.values_numpy()
        oracle_check.append({"l1": float(np.abs(b - ref).sum()), "max_abs": float(np.abs(b - ref).max())})

    model = None
    if not no_cpu:
        from oracle import oracle as orc

        orc.build()
        model = orc.CsrModel(n, True)
        model.bulk_load_edges(s, d)
    if args.pr_batches > 0:
        eu, ev = wl.canonical_edges(s, d, True)
        stream = wl.BatchStream(eu, ev, n, True, seed=2)
        t_warm, t_push, its, push_info = [], [], [], []
        for i in range(args.pr_batches):
            bs, bd, ba = stream.next(args.batch_size, 0.8)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g32.apply_arrays(bs, bd, ba)
            pr32.solve_l1(200, 1e-7, True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            r = prp.apply_push_arrays(bs, bd, ba, 1e-7)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            t_warm.append(t1 - t0)
            t_push.append(t2 - t1)
            its.append(int(pr32.last_repair_iterations))
            push_info.append(r)
            g64.apply_arrays(bs, bd, ba)
            pr64.solve_l1(1000, 1e-11, True)
            errors()
            if model is not None:
                pending.append((bs, bd, ba))
                if i == 0 or i == args.pr_batches - 1:
                    cpu_check(model)
        out.update({"batches": args.pr_batches, "batch_size": args.batch_size,
                    "warm_resolve": {"mean_batch_ms": 1e3 * float(np.mean(t_warm)),
                                     "batches_per_s": len(t_warm) / sum(t_warm),
                                     "mean_sweeps": float(np.mean(its)),
                                     "max_l1_error": max(errs["warm"][0]), "max_abs_error": max(errs["warm"][1])},
                    "push_repair": {"mean_batch_ms": 1e3 * float(np.mean(t_push)),
                                    "batches_per_s": len(t_push) / sum(t_push),
                                    "fallback_batches": int(sum(x["used_fallback"] for x in push_info)),
                                    "mean_rounds": float(np.mean([x["rounds"] for x in push_info])),
                                    "mean_pushed_arcs": float(np.mean([x["pushed_arcs"] for x in push_info])),
                                    "mean_fallback_sweeps": float(np.mean([x["fallback_sweeps"] for x in push_info])),
                                    "max_l1_error": max(errs["push"][0]), "max_abs_error": max(errs["push"][1])}})
        out["batches_per_s"] = out["push_repair"]["batches_per_s"]
        # smaller batches on the same stream: where the localized pushes pay
        sweep = {}
        for k in (10, 100, 1000):
            tw, tp, fb = [], [], 0
            for _ in range(10):
                bs, bd, ba = stream.next(k, 0.8)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                g32.apply_arrays(bs, bd, ba)
                pr32.solve_l1(200, 1e-7, True)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                r = prp.apply_push_arrays(bs, bd, ba, 1e-7)
                torch.cuda.synchronize()
                t2 = time.perf_counter()
                tw.append(t1 - t0)
                tp.append(t2 - t1)
                fb += int(r["used_fallback"])
                g64.apply_arrays(bs, bd, ba)
                pr64.solve_l1(1000, 1e-11, True)
                errors()
            sweep[str(k)] = {"warm_resolve_ms": 1e3 * float(np.mean(tw)), "push_repair_ms": 1e3 * float(np.mean(tp)),
                             "push_fallback_batches": fb, "batches": 10}
        out["batch_size_sweep"] = sweep
    out.update({"fp64_reference": "GPU FP64 solve to L1 residual 1e-11 on a third copy of the graph",
                "max_l1_error": max(errs["warm"][0] + errs["push"][0]),
                "max_abs_error": max(errs["warm"][1] + errs["push"][1])})
    out["within_1e-6"] = bool(out["max_l1_error"] <= 1e-6 and out["max_abs_error"] <= 1e-6)
    if oracle_check:
        out["fp64_reference_vs_cpu_oracle"] = oracle_check
    if not no_cpu:
        out["cpu_reference"] = cpu_pagerank(s, d, n, 3)
    log(f"pagerank: {out}")
    del pr32, prp, pr64, g32, gpush, g64
    return out


def leg_ppr(vx, args):
    import torch

    from velographx_amd import autograd as ag

    scale = args.ppr_scale
    n = 1 << scale
    s, d = vx.rmat_edges(scale, args.edgefactor, 1)
    g = vx.Graph(n, True)
    g.bulk_load_arrays(s, d)
    arcs = int(g.edge_count_directed)
    K = 32
    rng = np.random.default_rng(1)
    w = torch.tensor((1.0 - rng.random(arcs)).astype(np.float32), device="cuda", requires_grad=True)
    S = torch.zeros((n, K), device="cuda")
    S[torch.tensor(rng.integers(0, n, K), device="cuda"), torch.arange(K, device="cuda")] = 1.0
    S.requires_grad_(True)
    C = torch.tensor(rng.standard_normal((n, K)).astype(np.float32), device="cuda")
    T = args.ppr_iterations

    def step():
        w.grad = None
        S.grad = None
        x = ag.personalized_pagerank(g, w, S, 0.85, T)
        (x * C).sum().backward()
        return x

    step()
    torch.cuda.synchronize()
    t_f = []
    t_all = []
    for _ in range(args.ppr_reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.no_grad():
            ag.personalized_pagerank(g, w.detach(), S.detach(), 0.85, T)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        t_f.append(t1 - t0)
        t_all.append(time.perf_counter() - t1)
    tf = float(np.median(t_f))
    tfb = float(np.median(t_all))
    fwd_bytes = T * (8.0 * arcs + 4.0 * n + 12.0 * K * n)
    bwd_bytes = fwd_bytes + 12.0 * arcs + 8.0 * K * n
    out = {"workload": f"directed R-MAT s{scale} ef{args.edgefactor}, weights U(0,1] row-normalised, "
                       f"K={K} one-hot teleport columns, {T} sweeps, random linear loss",
           "arcs": arcs, "fwd_ms": 1e3 * tf, "fwd_bwd_ms": 1e3 * tfb, "fwd_bwd_per_s": 1.0 / tfb,
           "fwd_gbps": fwd_bytes / tf / 1e9, "fwd_bwd_gbps": (fwd_bytes + bwd_bytes) / tfb / 1e9,
           "roofline_frac_fwd_bwd": (fwd_bytes + bwd_bytes) / tfb / 1e9 / HBM_PEAK_GBPS,
           "grad_w_finite": bool(torch.isfinite(w.grad).all().item()),
           "grad_s_finite": bool(torch.isfinite(S.grad).all().item())}
    log(f"ppr: {out}")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch

        a = torch.ones(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy, "copies": 5}), flush=True)
        return
    if args.pmc_child:
        pmc_child(args)
        return
    if args.legs:
        want = {x.strip() for x in args.legs.split(",") if x.strip()}
        for leg in ("batched", "consolidation", "static", "repairs", "contrast", "batching", "kcore", "pagerank", "ppr"):
            if leg not in want:
                setattr(args, f"no_{leg}", True)
    if args.only_bfs:
        args.no_cpu = args.no_repairs = args.no_batched = args.no_contrast = args.no_consolidation = args.no_static = True
        args.no_kcore = True
        args.no_pagerank = args.no_ppr = args.no_batching = True
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
    lcc, arcs_c = component(vx, wl, g)
    roots = wl.roots_in_component(lcc, args.roots, seed=1)
    m_c = arcs_c / 2.0
    v_c = int(lcc.size)
    bytes_bfs = 4.0 * arcs_c + 12.0 * v_c
    log(f"LCC: {v_c} vertices, {arcs_c} arcs; bytes/BFS {bytes_bfs / 1e6:.1f} MB")

    t_bfs = timed_bfs(vx, g, roots, args.steps, args.warmup, warmup_roots(wl, lcc, roots, args.warmup))
    times, elapsed, kernel_ms_per_bfs = t_bfs["times"], t_bfs["elapsed"], t_bfs["kernel_ms"]
    walker = t_bfs["walker"]
    # every root's distances (outside the timed region): reach, and the
    # direction-optimizing result against a top-down-only traversal
    plain_dist = []
    shapes = []
    for r in roots:
        walker.set_source(int(r))
        plain_dist.append(walker.distances_numpy())
        st = walker.stats()
        shapes.append((st["levels"], st["bottom_up_levels"], st["steps"]))
    reached = {int(np.count_nonzero(x != INF)) for x in plain_dist}
    assert reached == {v_c}, reached
    gteps = m_c * args.steps / sum(times) / 1e9
    ms_per_step = 1e3 * elapsed / args.steps
    achieved = bytes_bfs / (kernel_ms_per_bfs * 1e-3) / 1e9
    log(f"BFS: {ms_per_step:.3f} ms/step, {gteps:.1f} GTEPS, GPU {kernel_ms_per_bfs:.4f} ms/BFS "
        f"over {t_bfs['launches']:.1f} step launches; same roots again {t_bfs['repeat_ms']:.4f} ms; "
        f"levels/bottom-up/steps {shapes}")
    chk = vx.IncrementalBFS(g, int(roots[0]))
    chk.set_tuning(alpha=1e-9, beta=1.0, blocks=0, levels_per_sync=0)
    chk.recompute()
    exact_bfs = bool(np.array_equal(plain_dist[0], chk.distances_numpy()))
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
                        f"from {args.roots} LCC roots one at a time; {args.batched_roots}-root batched pass; "
                        f"{args.batches} x {args.batch_size}-op update batches",
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
            "kernel": "one traversal: k_bfs_init and its k_bfs_step launches (levels, tails, packs)",
            "bytes_per_bfs": bytes_bfs,
            "gpu_ms_per_bfs": kernel_ms_per_bfs,
            "avg_launch_us": 1e3 * t_bfs["launch_ms"],
            "launches_per_bfs": t_bfs["launches"],
            "wall_frac": bytes_bfs / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS,
        },
        "bfs": {"timed": "first traversal from each root (one object re-targeted per step; nothing is kept per "
                         "root, the launch program is steered on the device)",
                "levels": [x[0] for x in shapes], "bottom_up_levels": [x[1] for x in shapes],
                "steps": [x[2] for x in shapes],
                "gpu_ms_per_bfs_same_root_again": t_bfs["repeat_ms"],
                "exact_vs_top_down": exact_bfs},
    }
    if not args.no_batched:
        result["batched"] = leg_batched(vx, wl, g, lcc, m_c, args.batched_roots, plain_dist, gteps)
    del walker, t_bfs
    if not args.no_consolidation:
        result["consolidation"] = leg_consolidation(vx, wl, g, s, d, n, args)
    if not args.no_static:
        result["static"] = leg_static(vx, s, d, roots[0])
    batches = None
    if not args.no_repairs and args.batches > 0:
        result["repairs"], batches = leg_repairs(vx, wl, g, s, d, n, roots[0], args)
    del g
    if not args.no_cpu:
        t0 = time.time()
        result["cpu_baseline"] = cpu_baseline(s, d, n, roots, m_c, args.cpu_budget)
        log(f"cpu baseline in {time.time() - t0:.1f}s: {result['cpu_baseline']}")
        if "consolidation" in result:
            cs = cpu_snapshot(s, d, n)
            if cs is not None:
                result["consolidation"]["cpu_reference"] = cs
                log(f"cpu snapshot: {cs}")
        if "static" in result:
            cst = cpu_static(s, d, roots[0])
            if cst is not None:
                result["static"]["cpu_reference"] = cst
                log(f"cpu static: {cst}")
        if batches is not None:
            cr = cpu_repairs(s, d, n, roots[0], batches, args.cpu_budget)
            if cr is not None:
                result["repairs"]["cpu_reference"] = cr
                log(f"cpu repairs: {cr}")
    if not args.no_contrast:
        result["contrast"] = leg_contrast(vx, wl, args.steps, args)
    if not args.no_batching:
        result["batching"] = leg_batching(vx, wl, args)
    if not args.no_kcore:
        result["kcore"] = leg_kcore(vx, args, not args.no_cpu)
    if not args.no_pagerank:
        result["pagerank"] = leg_pagerank(vx, wl, args, args.no_cpu)
    if not args.no_ppr:
        result["ppr"] = leg_ppr(vx, args)
    if not args.no_traffic:
        t0 = time.time()
        tr = measure_traffic(args, 8)
        log(f"traffic passes in {time.time() - t0:.1f}s: {tr}")
        if tr:
            traffic = tr["FETCH_SIZE"]["bytes_per_bfs"] + tr["WRITE_SIZE"]["bytes_per_bfs"]
            rf = result["roofline"]
            rf["traffic"] = traffic
            rf["traffic_detail"] = tr
            rf["traffic_frac"] = traffic / (kernel_ms_per_bfs * 1e-3) / 1e9 / HBM_PEAK_GBPS
            rf["traffic_over_algorithmic"] = traffic / bytes_bfs
        if "contrast" in result:
            t0 = time.time()
            tr = measure_traffic(args, 8, "er")
            log(f"ER traffic passes in {time.time() - t0:.1f}s: {tr}")
            if tr:
                c = result["contrast"]
                c["traffic"] = tr["FETCH_SIZE"]["bytes_per_bfs"] + tr["WRITE_SIZE"]["bytes_per_bfs"]
                c["traffic_detail"] = tr
                c["traffic_over_algorithmic"] = c["traffic"] / c["bytes_per_bfs"]
        if "pagerank" in result:
            t0 = time.time()
            tr = measure_traffic(args, 16, "pagerank")
            log(f"PageRank traffic passes in {time.time() - t0:.1f}s: {tr}")
            if tr:
                pg = result["pagerank"]
                sweep = tr["FETCH_SIZE"]["bytes_per_sweep"] + tr["WRITE_SIZE"]["bytes_per_sweep"]
                it = tr["FETCH_SIZE"]["bytes_per_k_pr_iter"] + tr["WRITE_SIZE"]["bytes_per_k_pr_iter"]
                pg["traffic"] = {"bytes_per_sweep": sweep, "bytes_per_k_pr_iter": it, "detail": tr,
                                 "algorithmic_bytes_per_sweep": pg["bytes_per_iteration"],
                                 "sweep_over_algorithmic": sweep / pg["bytes_per_iteration"]}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
