"""Build ``tycho2_amd/libtycho2_sweep.so`` for gfx950 (in-tree).

Host set-up code is compiled by g++ with ``-ffp-contract=off`` (bitwise
geometry); the kernels and the C ABI by hipcc for gfx950.  Packed FP32
instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBNAME = "libtycho2_sweep.so"
LIBPATH = os.path.join(HERE, LIBNAME)
DRIVER = os.path.join(HERE, "sweep_hip.x")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
             f"-I{INCLUDE}"]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
              "-Wno-unused-function", f"-I{INCLUDE}"]


def _hipcc() -> str:
    path = os.path.join(ROCM, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n"
                           + res.stdout + res.stderr)
    return res


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for name in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, name)
        obj = os.path.join(BUILD, name + ".o")
        if name.endswith(".cpp"):
            jobs.append(["g++", *HOST_FLAGS, "-c", src, "-o", obj])
        elif name.endswith(".hip"):
            jobs.append([_hipcc(), *HIP_FLAGS, "-c", src, "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        for res in ex.map(_run, jobs):
            if verbose and (res.stdout or res.stderr):
                print(res.stdout + res.stderr)
    objs = [j[-1] for j in jobs]
    tmp = LIBPATH + ".tmp"
    _run([_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp,
          *objs, "-lpthread"])
    shutil.move(tmp, LIBPATH)
    # C++ host driver (the reference's executable with SweepType HIP)
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}",
          os.path.join(HERE, "host", "sweep_hip.cpp"), "-o", DRIVER,
          f"-L{HERE}", "-ltycho2_sweep", "-Wl,-rpath,$ORIGIN"])
    return LIBPATH


if __name__ == "__main__":
    print(build(verbose=True))
