"""The reference's own C++ test programs, compiled against this library.

Test infrastructure only.  Each program of the reference's tests/ listed
below is compiled, read in place from the reference tree, against the C++
facade (include/: the reference's include paths velographx/... are served
by include/velographx/*.hpp) and linked with velographx_amd/libvgx.so, so
that running it on a GPU checks the facade with the reference's own
asserts.  Lines that use parts of the reference outside the traversal path
(triangle counting, weighted graphs, common-neighbour and Jaccard scores)
are dropped on the way into the compiler (fed on stdin); nothing of the
reference is written to disk.  Binaries land in oracle/_ref/facade/.

  python oracle/ref_tests.py          build every program
  python oracle/ref_tests.py --check  also compile python/bindings.cpp
                                      (-fsyntax-only) against the facade
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.environ.get("VGX_REFERENCE", "/root/reference")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "velographx_amd")
OUT = os.path.join(HERE, "_ref", "facade")

_OUTSIDE = re.compile(r"triangle|Triangle|Weighted|weighted|common_neighbor_count|jaccard_similarity")

# program -> True when lines outside the traversal path are dropped
PROGRAMS = {
    "test_storage_independence": True,       # BFS, SSSP, components, k-core, PageRank, alias
    "test_graph_access_adl": False,          # a foreign graph adapted by ADL only
    "test_incremental_bfs_deletion": False,  # BFS repair against full recomputes
    "test_pagerank_localized": False,        # localized PageRank repair and validation
    "test_randomized_pagerank": False,       # randomized batches against full solves
    "test_kcore_localized": False,
    "test_incremental_components_deletion": False,
    "test_consolidation_controller": False,
    "test_static": True,                     # CsrGraph and the one-shot algorithms
}

CXXFLAGS = ["-std=c++20", "-O2", f"-I{INCLUDE}"]


def available() -> bool:
    return os.path.isdir(os.path.join(REF, "tests")) and os.path.isdir(os.path.join(REF, "include", "velographx"))


def source_of(name: str) -> str:
    return os.path.join(REF, "tests", name + ".cpp")


def _filtered(path: str) -> str:
    with open(path, encoding="utf-8") as f:
        return "".join(l for l in f if not _OUTSIDE.search(l))


def _compile(args, stdin_text=None):
    r = subprocess.run(args, input=stdin_text, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(args[:6])} ...\n{r.stdout}{r.stderr}")


def compile_program(name: str, out_dir: str = OUT, syntax_only: bool = False) -> str:
    """Compiles one program; returns the binary's path ('' for syntax_only)."""
    src = source_of(name)
    filtered = PROGRAMS[name]
    args = ["g++", *CXXFLAGS]
    if filtered:
        # stdin has no directory of its own: quoted includes resolve via -I
        args += ["-x", "c++", "-"]
    else:
        args += [src]
    if syntax_only:
        _compile(args + ["-fsyntax-only"], _filtered(src) if filtered else None)
        return ""
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, name)
    rpath = "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIBDIR, out_dir)
    _compile(args + [f"-L{LIBDIR}", "-lvgx", rpath, "-o", exe + ".tmp"], _filtered(src) if filtered else None)
    os.replace(exe + ".tmp", exe)
    return exe


def check_bindings() -> None:
    """The reference's pybind11 module source, minus its triangle and
    weighted-graph lines, compiled -fsyntax-only against the facade."""
    import pybind11

    src = os.path.join(REF, "python", "bindings.cpp")
    args = ["g++", "-std=c++20", "-fsyntax-only", f"-I{INCLUDE}", f"-I{pybind11.get_include()}",
            f"-I{sysconfig.get_paths()['include']}", "-x", "c++", "-"]
    _compile(args, _filtered(src))


def build(jobs: int = 8) -> list[str]:
    if not available():
        print("reference tree not present; skipping oracle/_ref/facade")
        return []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(compile_program, PROGRAMS))


if __name__ == "__main__":
    built = build()
    if "--check" in sys.argv[1:] and available():
        check_bindings()
    print("\n".join(built))
