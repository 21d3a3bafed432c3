"""Builds aieplace_amd/libeplace.so (HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python -m aieplace_amd.build   (or aieplace_amd.build.build())
The library is built in-tree so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libeplace.so")
RUNNER = os.path.join(HERE, "eplace_run")
EXAMPLE = os.path.join(HERE, "placer_example")

SOURCES = ["k_place.hip", "k_field.hip", "k_mlp.hip", "k_batch.hip", "ep_api.cpp", "batch.cpp",
           "generator.cpp", "lefdef.cpp", "legalize.cpp"]
ARCH = os.environ.get("EPLACE_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build aieplace_amd")


def _flags():
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        # no packed-FP32 instructions in device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-command-line-argument", "-Wno-unused-result",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc()] + _flags() + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), 6)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # command-line driver mirroring the reference's AIEplace.exe (main.cpp)
    main_src = os.path.join(CSRC, "main.cpp")
    if os.path.exists(main_src):
        cmd = [hipcc(), "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"), main_src,
               "-o", RUNNER + ".tmp", "-L", HERE, "-leplace", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"eplace_run build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(RUNNER + ".tmp", RUNNER)
    # the C++ Placer example (include/eplace_placer.hpp over the C ABI)
    ex_src = os.path.join(CSRC, "placer_example.cpp")
    if os.path.exists(ex_src):
        cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"), ex_src,
               "-o", EXAMPLE + ".tmp", "-L", HERE, "-leplace", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"placer_example build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(EXAMPLE + ".tmp", EXAMPLE)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
