"""Build libmr.so (the gfx950 HIP library behind include/mr.h) in-tree.

Every ``csrc/kernels/*.hip`` file is compiled with hipcc for gfx950 and every
``csrc/*.cpp`` host file with g++ against the HIP runtime headers; the objects
are linked into ``matrix-reorder_amd/lib/libmr.so``.  The optional CLI driver
``cli/mr_main.cpp`` (the reference's main.cu workflow on the C++ mirror of its
GRAPH interface, include/mr_graph.hpp) is linked against it into lib/mr_main.

Usage: ``python matrix-reorder_amd/build.py [--jobs N] [--no-cli]``
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib"
ROCM = pathlib.Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIPCC_FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + str(ROOT / "include"),
]
CXX_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
    "-I" + str(ROCM / "include"), "-I" + str(ROOT / "include"),
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + res.stdout + res.stderr)


def _compile(src: pathlib.Path) -> pathlib.Path:
    obj = OBJ / (src.stem + (".hip.o" if src.suffix == ".hip" else ".o"))
    if src.suffix == ".hip":
        cmd = [str(ROCM / "bin" / "hipcc"), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(jobs: int | None = None, with_cli: bool = True) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("kernels/*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    out = LIB / "libmr.so"
    tmp = LIB / "libmr.so.tmp"
    _run([str(ROCM / "bin" / "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs], "-lpthread", "-Wl,-soname,libmr.so"])
    os.replace(tmp, out)
    if with_cli:
        # the driver needs only mr.h / mr_graph.hpp and libmr (no HIP headers)
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + str(ROOT / "include"), str(PKG / "cli" / "mr_main.cpp"),
              "-o", str(LIB / "mr_main"), "-L" + str(LIB), "-lmr", "-Wl,-rpath,$ORIGIN"])
    return out


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--no-cli", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, not a.no_cli))
    sys.exit(0)
