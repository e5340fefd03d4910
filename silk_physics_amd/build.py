"""Build the native libraries of silk_physics_amd in-tree.

Products (all next to this file, under ``lib/``):
  libsilk_hip.so  HIP device library + C ABI (include/silk_hip.h)
  libsilk.so      C++20 silk::World front end (include/silk/silk.hpp)
  world_api_test  C++ test driver for the World API (tests call it)
  silk_headless   scene-file driver: runs a JSON scene through silk::World and
                  writes the trajectory (csrc/headless)

Device code is compiled for gfx950 only, with packed-FP32 instruction
selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = PKG / "lib" / "obj"
INCLUDE = ROOT / "include"
ARCH = os.environ.get("SILK_OFFLOAD_ARCH", "gfx950")

HIPCC = "/opt/rocm/bin/hipcc"
CXX = "g++"

HIP_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O3", "-std=c++20", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}"]

HIP_SOURCES = sorted((CSRC / "hip").glob("*.hip"))
RUNTIME_SOURCES = sorted((CSRC / "api").glob("*.cpp"))
HOST_SOURCES = sorted((CSRC / "host").glob("*.cpp")) + sorted((CSRC / "cpu").glob("*.cpp"))


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    if proc.stderr.strip():
        # keep warnings visible but do not fail
        lines = [l for l in proc.stderr.splitlines()
                 if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines[:40]) + "\n")


def _compile(src: Path) -> Path:
    obj = OBJ / (src.parent.name + "_" + src.stem + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    elif src.parent.name == "api":
        # runtime sources include hip_runtime.h (host code only)
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [CXX, *HOST_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> dict[str, Path]:
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    sources = HIP_SOURCES + RUNTIME_SOURCES + HOST_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, sources))
    hip_so = LIB / "libsilk_hip.so"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o",
          str(hip_so), *map(str, objs)])
    world_so = LIB / "libsilk.so"
    _run([CXX, *HOST_FLAGS, "-shared", "-o", str(world_so),
          str(CSRC / "frontend" / "world.cpp"), f"-L{LIB}", "-lsilk_hip",
          "-Wl,-rpath,$ORIGIN"])
    headless = LIB / "silk_headless"
    _run([CXX, *HOST_FLAGS, "-o", str(headless),
          str(CSRC / "headless" / "headless.cpp"), str(CSRC / "headless" / "mesh_io.cpp"),
          f"-L{LIB}", "-lsilk", "-lsilk_hip", f"-Wl,-rpath,$ORIGIN"])
    test_bin = LIB / "world_api_test"
    _run([CXX, *HOST_FLAGS, "-o", str(test_bin),
          str(ROOT / "tests" / "cpp" / "world_api_test.cpp"), f"-L{LIB}",
          "-lsilk", "-lsilk_hip", f"-Wl,-rpath,{LIB}"])
    if verbose:
        print(f"built {hip_so}, {world_so}, {test_bin}")
    return {"silk_hip": hip_so, "silk": world_so, "world_api_test": test_bin,
            "silk_headless": headless}


if __name__ == "__main__":
    build(verbose=True)
