"""Builds libaccspmm.so (gfx950 HIP kernels + C ABI + host utilities), the
`mma` driver, and the CPU oracle used by the tests.

Run as ``python -m accspmm_amd.build`` or through ``__graft_entry__.build()``.
Every HIP translation unit is compiled with hipcc for gfx950 only, with packed
FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib" / "libaccspmm.so"
MMA = PKG / "bin" / "mma"
ORACLE = REPO / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result"]

HIP_SOURCES = ["scan.hip", "canon.hip", "builder.hip", "spmm.hip", "sweep.hip", "batch.hip",
               "server.hip", "noalloc.hip", "probe.hip"]
HOST_HIP_SOURCES = ["tune.cpp", "serialize.cpp"]  # host code using the HIP runtime
HOST_SOURCES = ["host.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")


def _compile(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    elif src.name in HOST_HIP_SOURCES:
        cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", "-Wall",
               "-Wno-unused-command-line-argument", "-x", "c++",
               "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
               "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *HOST_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in HIP_SOURCES + HOST_HIP_SOURCES + HOST_SOURCES
            if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> Path:
    MMA.parent.mkdir(exist_ok=True)
    src = CSRC / "mma_main.cpp"
    if not src.exists():
        return MMA
    _run(["g++", *HOST_FLAGS, "-I", str(PKG / "include"), "-I", str(REPO / "include"),
          "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
          str(src), "-o", str(MMA), "-L", str(LIB.parent), "-laccspmm",
          "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,$ORIGIN/../lib",
          "-Wl,-rpath,/opt/rocm/lib"])
    return MMA


def build_test_programs() -> None:
    """C++ programs the GPU tests run (reference-style users of the headers)."""
    src = REPO / "tests" / "cpp" / "format_dropin.cpp"
    if not src.exists():
        return
    _run(["g++", *HOST_FLAGS, "-I", str(PKG / "include"), str(src), "-o",
          str(src.with_suffix("")), "-L", str(LIB.parent), "-laccspmm",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../accspmm_amd/lib",
          "-Wl,-rpath,/opt/rocm/lib"])


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE)])
    return ORACLE / "liboracle.so"


def build_all() -> None:
    build_library()
    build_driver()
    build_test_programs()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}, {MMA}, {ORACLE / 'liboracle.so'}")
