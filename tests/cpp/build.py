"""Builds the C++ facade test programs of this directory into tests/cpp/bin/.

Each program includes the facade through the reference's include paths
(include/velographx/...) and links velographx_amd/libvgx.so; it runs on a
GPU (tests/test_gpu_cpp_facade.py) and exits 0 when its checks hold.

  python tests/cpp/build.py
"""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin")
LIBDIR = os.path.join(ROOT, "velographx_amd")
PROGRAMS = ["facade_foreign", "facade_storage"]


def compile_program(name: str, out_dir: str = BIN, syntax_only: bool = False) -> str:
    args = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}",
            os.path.join(HERE, name + ".cpp")]
    if syntax_only:
        args.append("-fsyntax-only")
        exe = ""
    else:
        os.makedirs(out_dir, exist_ok=True)
        exe = os.path.join(out_dir, name)
        args += [f"-L{LIBDIR}", "-lvgx", "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIBDIR, out_dir), "-o", exe + ".tmp"]
    r = subprocess.run(args, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{name}: build failed\n{r.stdout}{r.stderr}")
    if exe:
        os.replace(exe + ".tmp", exe)
    return exe


def build() -> list[str]:
    with ThreadPoolExecutor(max_workers=len(PROGRAMS)) as ex:
        return list(ex.map(compile_program, PROGRAMS))


if __name__ == "__main__":
    print("\n".join(build()))
