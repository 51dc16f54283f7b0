"""The C ABI library loads on a host without a GPU and exports every entry
point include/vgx.h declares (no compute calls are made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "velographx_amd", "libvgx.so")


def declared():
    text = open(os.path.join(ROOT, "include", "vgx.h")).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(vgx_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared()
    for must in ("vgx_graph_create", "vgx_graph_apply", "vgx_bfs_create", "vgx_bfs_apply", "vgx_bfs_distances",
                 "vgx_pagerank_create", "vgx_pagerank_apply", "vgx_msbfs_run", "vgx_ppr_forward",
                 "vgx_ppr_backward", "vgx_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libvgx.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.vgx_version.restype = ctypes.c_char_p
    assert lib.vgx_version().decode()
    lib.vgx_last_error.restype = ctypes.c_char_p
    assert lib.vgx_last_error() is not None


@pytest.mark.skipif(not os.path.exists(LIB), reason="libvgx.so not built")
def test_null_handles_are_rejected_without_a_device():
    """Argument checks run before any device call: a null handle returns
    VGX_INVALID_ARGUMENT and sets the thread-local message."""
    lib = ctypes.CDLL(LIB)
    lib.vgx_last_error.restype = ctypes.c_char_p
    st = lib.vgx_bfs_recompute(None)
    assert st != 0
    assert lib.vgx_last_error()
