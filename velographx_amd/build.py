"""Build recipe for the gfx950 engine (in-tree).

  velographx_amd/libvgx.so          HIP kernels + C ABI (hipcc --offload-arch=gfx950)
  velographx_amd/_velographx*.so    pybind11 module over the C++ facade (g++)

Run ``python velographx_amd/build.py`` (or ``__graft_entry__.build()``); it is a
plain script so that it runs before the package can be imported.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libvgx.so")

HIP_SOURCES = ["graph.hip", "bfs.hip", "pagerank.hip", "msbfs.hip", "ppr.hip", "kcore.hip", "components.hip"]
HOST_SOURCES = ["rmat.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are kept out of every code object.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}"]


def ext_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def module_path() -> str:
    return os.path.join(PKG, "_velographx" + ext_suffix())


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{msg}")


def _compile_hip(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def _compile_host(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run(["g++", "-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}", "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_library(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs)) + list(ex.map(_compile_host, HOST_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_bindings() -> str:
    import pybind11

    out = module_path()
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-shared", "-fPIC", "-std=c++20", "-fvisibility=hidden", f"-I{INCLUDE}",
          f"-I{pybind11.get_include()}", f"-I{py_inc}", os.path.join(PKG, "python", "bindings.cpp"),
          f"-L{PKG}", "-lvgx", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build(jobs: int = 8) -> None:
    build_library(jobs)
    build_bindings()


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print(LIB)
    print(module_path())
