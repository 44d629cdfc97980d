"""Builds libmpm.so (the HIP engine and its C ABI) for gfx950, in-tree.

Each source compiles to its own object under _build/ (in parallel), then
hipcc links the shared library next to this file.  Run as a script or call
``build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libmpm.so"

ARCH = os.environ.get("MPM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]
# -ffp-contract=off: the scene front-end restates FP32 arithmetic step by step
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}"]

HIP_SOURCES = ["kernels_transfer.hip", "kernels_grid.hip", "kernels_sort.hip", "engine.hip", "corrector.hip",
               "noalloc.hip"]
CXX_SOURCES = ["config.cpp", "scene.cpp", "voxel.cpp"]


def _hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def _compile(cmd: list[str]) -> tuple[list[str], int, str]:
    r = subprocess.run(cmd, capture_output=True, text=True)
    return cmd, r.returncode, r.stdout + r.stderr


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = BUILD / (Path(s).stem + ".o")
        objs.append(o)
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(o)])
    for s in CXX_SOURCES:
        o = BUILD / (Path(s).stem + ".o")
        objs.append(o)
        jobs.append(["g++", *CXX_FLAGS, "-c", str(CSRC / s), "-o", str(o)])
    workers = min(len(jobs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for cmd, rc, out in ex.map(_compile, jobs):
            if verbose and out.strip():
                print(out)
            if rc != 0:
                raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{out}")
    tmp = HERE / "libmpm.so.tmp"
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    cmd, rc, out = _compile(link)
    if rc != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{out}")
    os.replace(tmp, LIB)
    # the headless driver (the reference's program without its window)
    exe = HERE / "mpm_headless"
    cmd, rc, out = _compile(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(CSRC / "headless.cpp"),
                             f"-L{HERE}", "-lmpm", "-Wl,-rpath,$ORIGIN", "-o", str(exe)])
    if rc != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{out}")
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
