"""Build of the native code (gfx950 HIP library, host helpers, CLI).

Products (all in-tree, git-ignored, shipped with the tree):
  fast-cuda-gpu-dbscan_amd/libdclust.so        HIP kernels + C ABI (include/dclust.h)
  fast-cuda-gpu-dbscan_amd/libdclust_synth.so  synthetic workloads (include/dclust_synth.h)
  fast-cuda-gpu-dbscan_amd/bin/dclust          CLI (main.exe counterpart)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: the FP64 predicate must round exactly like the reference
# oracle (no fused multiply-add); FP32 code uses explicit fmaf where wanted.
# -packed-fp32-ops: never emit packed FP32 instructions.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", INC, "-I", CSRC,
]

HIP_SOURCES = ["api.hip", "index.hip", "cluster.hip", "onestep.hip"]
HOST_SOURCES = ["io.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_dclust(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        objs.append(obj)
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-I", INC, "-c",
                     os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    lib = os.path.join(PKG, "libdclust.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp],
         verbose)
    os.replace(tmp, lib)
    return lib


def build_synth(verbose=False):
    lib = os.path.join(PKG, "libdclust_synth.so")
    tmp = lib + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INC,
          os.path.join(CSRC, "synth.cpp"), "-o", tmp], verbose)
    os.replace(tmp, lib)
    return lib


def build_cli(verbose=False):
    bindir = os.path.join(PKG, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, "dclust")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INC, os.path.join(CSRC, "cli.cpp"),
          "-L", PKG, "-ldclust", "-Wl,-rpath,$ORIGIN/..", "-o", exe], verbose)
    return exe


def build_oracle(verbose=False):
    """The CPU checker (oracle/), and the reference's own oracle compiled in
    place when the reference tree is present in this container."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-s", "liboracle.so"], verbose)
    if os.path.isdir("/root/reference/cpu-dbscan"):
        _run(["make", "-C", odir, "-s", "ref"], verbose)


def build_all(verbose=False):
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        f1 = ex.submit(build_dclust, verbose)
        f2 = ex.submit(build_synth, verbose)
        f3 = ex.submit(build_oracle, verbose)
        f1.result()
        f2.result()
        f3.result()
    build_cli(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
