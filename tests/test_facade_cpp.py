"""The C++ facade against the reference's own C++ sources (CPU: compile only).

The reference's python/bindings.cpp (minus its triangle and weighted-graph
lines) must compile against the facade, and so must the reference test
programs of oracle/ref_tests.py and this repository's own programs under
tests/cpp/; test_gpu_cpp_facade.py runs the built programs on a GPU.
"""
from concurrent.futures import ThreadPoolExecutor

import pytest

from oracle import ref_tests
from tests.cpp import build as cpp_build

needs_reference = pytest.mark.skipif(not ref_tests.available(), reason="reference tree not present")


@needs_reference
def test_reference_bindings_compile_against_facade():
    ref_tests.check_bindings()


@needs_reference
def test_reference_test_programs_compile_against_facade():
    names = list(ref_tests.PROGRAMS)
    with ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(lambda n: ref_tests.compile_program(n, syntax_only=True), names))


def test_own_facade_programs_compile():
    for name in cpp_build.PROGRAMS:
        cpp_build.compile_program(name, syntax_only=True)
