"""Builds libkeepsake_amd.so (gfx950) in-tree and, when asked, the CPU oracle used by the tests.

The library is compiled with explicit hipcc/g++ command lines (no CMake): the HIP translation unit
carries every kernel and the C ABI; the BVH builder and the procedural scene generators are host-only
C++. Output: keepsake_amd/_lib/libkeepsake_amd.so.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libkeepsake_amd.so"

ARCH = os.environ.get("KS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions are kept out of every HIP code object.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-function"]

HIP_SOURCES = ["ks_api.hip", "ks_guide.hip", "ks_bvh_device.hip"]
HOST_SOURCES = ["sbvh_build.cpp", "procedural.cpp", "traversal_count.cpp"]
HEADERS = ["ks_device.h", "ks_traverse.h", "ks_internal.h"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_library(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        _run([CXX, *HOST_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


HOST_TOOL_SOURCES = ["config.cpp", "mesh_asset.cpp", "gltf_asset.cpp", "image_io.cpp", "texture_asset.cpp",
                     "small_pt_task.cpp", "ks_render.cpp"]
TOOL = LIBDIR / "ks_render"


def build_tools() -> Path:
    """The ks_render task runner (C++ host code over the C ABI), linked against the library."""
    cmd = [CXX, *HOST_FLAGS, "-I", str(ROOT / "include")]
    cmd += [str(PKG / "host" / s) for s in HOST_TOOL_SOURCES]
    cmd += ["-L", str(LIBDIR), "-lkeepsake_amd", "-lz", "-Wl,-rpath,$ORIGIN", "-o", str(TOOL)]
    _run(cmd)
    return TOOL


def build_oracle() -> Path:
    _run(["make", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "libks_oracle.so"


def main(argv: list[str]) -> int:
    build_library()
    build_tools()
    if "--no-oracle" not in argv:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
