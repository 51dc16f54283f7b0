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
                 "vgx_ppr_backward", "vgx_last_error", "vgx_graph_snapshot", "vgx_graph_probe_neighbors",
                 "vgx_pagerank_static", "vgx_graph_apply_device", "vgx_bfs_apply_device", "vgx_kcore_create",
                 "vgx_kcore_apply", "vgx_components_create", "vgx_components_apply", "vgx_graph_device"):
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
    # the entry points added beside the traversal path check their handles the same way
    for fn in (lib.vgx_kcore_recompute, lib.vgx_components_recompute):
        assert fn(None) != 0 and lib.vgx_last_error()
    out = ctypes.c_void_p()
    assert lib.vgx_graph_snapshot(None, ctypes.byref(out)) != 0
    assert lib.vgx_kcore_create(None, ctypes.byref(out)) != 0
    assert lib.vgx_components_create(None, ctypes.byref(out)) != 0
    ns, dig = ctypes.c_double(), ctypes.c_uint64()
    assert lib.vgx_graph_probe_neighbors(None, ctypes.c_uint64(1), ctypes.c_uint32(1), ctypes.byref(ns),
                                         ctypes.byref(dig)) != 0
