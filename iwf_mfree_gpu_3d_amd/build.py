"""Build the gfx950 HIP library (libmfree_hip.so) in-tree with hipcc.

Every source is compiled for gfx950 only, with
  -ffp-contract=off   expressions round as written (the FP64 path follows the reference order)
  -packed-fp32-ops    disabled (packed FP32 instructions are not generated)
and linked into one shared library next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libmfree_hip.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["abi.hip", "k_grid.hip", "k_pair.hip", "k_tool.hip", "k_actions.hip", "scene.cpp", "mesh_io.cpp"]
ARCH = os.environ.get("MFG_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def common_flags() -> list[str]:
    return ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{INCLUDE}", f"-I{CSRC}",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def device_flags() -> list[str]:
    return [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: str) -> str:
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [hipcc(), "-x", "hip", *device_flags(), *common_flags(), "-c", path, "-o", obj]
    else:
        cmd = [hipcc(), *common_flags(), "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link libmfree_hip.so (always rebuilds)."""
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_driver(verbose)
    return LIB


DRIVER_SRC = os.path.join(HERE, "driver", "mfree_gpu_3d.cpp")
DRIVER = os.path.join(HERE, "mfree_gpu_3d")


def build_driver(verbose: bool = False) -> str:
    """The C++ host driver (reference-shaped main over the C++ shim), linked to the library with
    an rpath next to it."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", DRIVER_SRC, "-o", DRIVER + ".tmp",
           f"-L{HERE}", "-lmfree_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
