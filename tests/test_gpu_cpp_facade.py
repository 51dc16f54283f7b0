"""Runs the C++ facade programs on the GPU.

  oracle/_ref/facade/*  the reference's own test programs (tests/*.cpp),
                        compiled against the facade by oracle/ref_tests.py;
                        each exits 0 when its asserts hold
  tests/cpp/bin/*       this repository's programs (tests/cpp/*.cpp):
                        foreign graphs through graph_access, the storage
                        getters, maybe_compact, algorithm objects across a
                        compaction
Programs that were not built (no reference tree where the build ran) are
skipped.
"""
import os
import subprocess
import sys

import pytest

from oracle import ref_tests
from tests.cpp import build as cpp_build

pytestmark = pytest.mark.gpu


# The randomized PageRank program's own size knobs: its host-side reference
# solve reads every row through DynamicGraph::neighbors() (a device round
# trip per row) for up to 2000 iterations per check, so the default 4 seeds
# x 300 batches would take many minutes; 2 seeds x 40 batches keep its
# checks and its fallback rounds.
ENV = {"VELOGRAPHX_PAGERANK_RANDOMIZED_OPS": "40", "VELOGRAPHX_PAGERANK_RANDOMIZED_SEEDS": "2"}


def _run(exe):
    if not os.path.exists(exe):
        pytest.skip(f"{os.path.basename(exe)} not built")
    print(f"running {exe}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100, env={**os.environ, **ENV})
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0, f"{os.path.basename(exe)} exited {r.returncode}: {r.stderr[-2000:]}"


@pytest.mark.parametrize("name", list(ref_tests.PROGRAMS))
def test_reference_program(name):
    _run(os.path.join(ref_tests.OUT, name))


@pytest.mark.parametrize("name", cpp_build.PROGRAMS)
def test_own_program(name):
    _run(os.path.join(cpp_build.BIN, name))
