"""Build the MI355X shared library (libonionpir.so) in-tree with hipcc.

    python -m onionpirv2_amd.build          # or: python onionpirv2_amd/build.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIB_DIR, "libonionpir.so")

HIP_SOURCES = ["server.hip"]
CXX_SOURCES = ["client.cpp", "facade.cpp", "serve.cpp"]

# packed-fp32 instructions are disabled for every device compilation
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES + [s for s in CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]:
        obj = os.path.join(LIB_DIR, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            cmd = [hipcc()] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-c",
                                          os.path.join(CSRC, src), "-o", obj]
        else:  # host-only C++ (hipcc supplies the HIP headers the shared helpers use)
            cmd = [hipcc(), "-O3", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"),
                   "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_cpp_tests(verbose)
    return LIB


def build_cpp_tests(verbose: bool = False) -> None:
    """C++ programs under tests/cpp that use the drop-in API (libonionpir.so)."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "_build")
    os.makedirs(out_dir, exist_ok=True)
    for src in sorted(os.listdir(src_dir)):
        if not src.endswith(".cpp"):
            continue
        exe = os.path.join(out_dir, src[:-4])
        cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(src_dir, src),
               "-o", exe, "-L", LIB_DIR, "-lonionpir", "-Wl,-rpath,$ORIGIN/../../../onionpirv2_amd/_lib"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
