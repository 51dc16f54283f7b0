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
                 "vgx_kcore_apply", "vgx_components_create", "vgx_components_apply", "vgx_graph_device",
                 "vgx_graph_bulk_load_device", "vgx_graph_bulk_load_csr_device", "vgx_msbfs_apply",
                 "vgx_msbfs_apply_device", "vgx_msbfs_set_repair", "vgx_msbfs_repair_stats", "vgx_msbfs_set_params",
                 "vgx_msbfs_get_params"):
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
    # device edge lists and the multi-source repair entry points
    u64 = ctypes.c_uint64
    assert lib.vgx_graph_bulk_load_device(None, None, None, u64(0), u64(1), ctypes.c_uint32(4), None) != 0
    assert lib.vgx_graph_bulk_load_csr_device(None, None, None, u64(0), u64(0), ctypes.c_uint32(4), None) != 0
    assert lib.vgx_msbfs_apply(None, None, None, None, u64(0)) != 0 and lib.vgx_last_error()
    assert lib.vgx_msbfs_apply_device(None, None, None, None, u64(0), None) != 0
    assert lib.vgx_msbfs_set_repair(None, 0, ctypes.c_double(0.1)) != 0
    buf = ctypes.create_string_buffer(256)
    assert lib.vgx_msbfs_repair_stats(None, buf) != 0
    assert lib.vgx_msbfs_set_params(None, buf) != 0 and lib.vgx_msbfs_get_params(None, buf) != 0
    ns, dig = ctypes.c_double(), ctypes.c_uint64()
    assert lib.vgx_graph_probe_neighbors(None, ctypes.c_uint64(1), ctypes.c_uint32(1), ctypes.byref(ns),
                                         ctypes.byref(dig)) != 0
