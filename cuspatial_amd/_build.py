"""Builds libcuspatial_amd.so for gfx950 in-tree (hipcc, no JIT cache).

    python -m cuspatial_amd._build        # or __graft_entry__.build()

Every .hip file under csrc/ is compiled to an object in csrc/_build/ (in
parallel) and linked into cuspatial_amd/libcuspatial_amd.so.
"""

from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libcuspatial_amd.so")
ARCH = os.environ.get("CUSPATIAL_AMD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",  # FP64 predicate parity: no implicit FMA contraction
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",  # only the QPIP_API entry points are exported
    "-Wall",
    "-Wno-unused-function",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
    os.replace(obj + ".tmp", obj)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_cpp_tests(verbose)
    return LIB


CPP_TESTS = os.path.join(os.path.dirname(HERE), "tests", "cpp")


def build_cpp_tests(verbose: bool = False) -> None:
    """Compiles the C++ interface's test driver(s) under tests/cpp (host code
    linked against the library) into tests/cpp/_build."""
    srcs = sorted(glob.glob(os.path.join(CPP_TESTS, "*.cpp")))
    if not srcs:
        return
    out_dir = os.path.join(CPP_TESTS, "_build")
    os.makedirs(out_dir, exist_ok=True)
    for src in srcs:
        exe = os.path.join(out_dir, os.path.splitext(os.path.basename(src))[0])
        cmd = [HIPCC, "-O2", "-std=c++17", src, "-o", exe + ".tmp", f"-L{HERE}", "-lcuspatial_amd",
               "-Wl,-rpath,$ORIGIN/" + os.path.relpath(HERE, out_dir)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
        os.replace(exe + ".tmp", exe)
        if verbose:
            print("built", exe)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
