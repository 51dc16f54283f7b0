"""Helpers that run scenarios through oracle/_ref/vgx_ref (the compiled
reference engine) and parse its output.  Test infrastructure only."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "vgx_ref")


def available():
    return os.path.exists(REF_BIN) and os.access(REF_BIN, os.X_OK)


class Scenario:
    def __init__(self):
        self.lines = []

    def graph(self, n, directed):
        self.lines.append(f"graph {n} {int(directed)}")

    def bulk(self, edges):
        edges = list(edges)
        self.lines.append(f"bulk {len(edges)}")
        self.lines.extend(f"{u} {v}" for u, v in edges)

    def batch(self, ops):
        ops = list(ops)
        self.lines.append(f"batch {len(ops)}")
        self.lines.extend(f"{u} {v} {int(a)}" for u, v, a in ops)

    def cmd(self, *words):
        self.lines.append(" ".join(str(w) for w in words))

    def run(self):
        out = subprocess.run([REF_BIN], input="\n".join(self.lines) + "\n", capture_output=True,
                             text=True, check=True).stdout
        return parse(out)


def parse(text):
    recs = []
    for line in text.splitlines():
        w = line.split()
        if not w:
            continue
        tag = w[0]
        if tag == "dist":
            recs.append(("dist", np.array([int(x) for x in w[2:]], dtype=np.uint32)))
        elif tag == "bfsdiag":
            recs.append(("bfsdiag", {"reachable_count": int(w[1]), "deletion_candidates": int(w[2]),
                                     "affected_vertices": int(w[3]), "used_full_recompute": w[4] == "1"}))
        elif tag == "values":
            recs.append(("values", np.array([float(x) for x in w[2:]], dtype=np.float64)))
        elif tag == "prdiag":
            recs.append(("prdiag", {"last_repaired_vertices": int(w[1]), "last_repair_iterations": int(w[2]),
                                    "last_residual_l1": float(w[3]), "last_residual_linf": float(w[4]),
                                    "last_full_recompute_converged": w[5] == "1"}))
        elif tag == "prvalid":
            recs.append(("prvalid", {"l1_error": float(w[1]), "linf_error": float(w[2]),
                                     "reference_iterations": int(w[3]), "reference_converged": w[4] == "1",
                                     "within_tolerance": w[5] == "1"}))
        elif tag == "cclabels":
            recs.append(("cclabels", np.array([int(x) for x in w[2:]], dtype=np.uint32)))
        elif tag == "ccdiag":
            recs.append(("ccdiag", int(w[1])))
        elif tag == "kcore":
            recs.append(("kcore", np.array([int(x) for x in w[2:]], dtype=np.uint32)))
        elif tag == "kcorediag":
            recs.append(("kcorediag", int(w[1])))
        elif tag == "sdist":
            recs.append(("sdist", np.array([int(x) for x in w[2:]], dtype=np.uint64)))
        elif tag == "cpolicy":
            recs.append(("cpolicy", (float(w[1]), float(w[2]))))
        elif tag == "csignal":
            recs.append(("csignal", (float(w[1]), float(w[2]), w[3] == "1", w[4] == "1", w[5] == "1")))
        elif tag == "cobserve":
            recs.append(("cobserve", (w[1] == "1", int(w[2]), int(w[3]))))
        elif tag == "neighbors":
            recs.append(("neighbors", [int(x) for x in w[2:]]))
    return recs
