"""Builds libmfs_core.so (HIP for gfx950 + host plan code) in-tree, and the
torch operator library libmfs_torch.so over it.

    python -m fastmbcui_amd.build

Every HIP translation unit is compiled with packed FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); the host plan code is
built with ``-ffp-contract=off`` so its distance sums round as written.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libmfs_core.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MFS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["mfs_core.hip", "k_plan.hip", "k_factor.hip", "k_solve.hip", "k_points.hip",
               "k_mlp.hip", "mlp_core.hip", "k_caller.hip", "k_fused.hip"]
HOST_SOURCES = ["plan_host.cpp"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


ROOT = os.path.dirname(HERE)
# The reference's op package loads ./build/mfs_torch/csrc/libmfs_torch.so
# relative to the working directory (core/mfs_torch_py/__init__.py:7-8).
SHIM = os.path.join(ROOT, "build", "mfs_torch", "csrc", "libmfs_torch.so")


def build_torch_shim(verbose: bool = False) -> str:
    """libmfs_torch.so: the mfs_torch operator library (host C++, torch headers)
    over libmfs_core.so, at the path the reference's op package loads."""
    import torch
    tdir = os.path.dirname(torch.__file__)
    abi = int(bool(torch._C._GLIBCXX_USE_CXX11_ABI))
    os.makedirs(os.path.dirname(SHIM), exist_ok=True)
    tmp = SHIM + ".tmp"
    r = _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
              f"-D_GLIBCXX_USE_CXX11_ABI={abi}", f"-I{tdir}/include",
              f"-I{tdir}/include/torch/csrc/api/include", "-I/opt/rocm/include",
              os.path.join(CSRC, "torch_shim.cpp"), "-o", tmp, f"-L{tdir}/lib", "-lc10", "-lc10_hip",
              "-ltorch", "-ltorch_cpu", "-ltorch_hip", f"-L{LIB_DIR}", "-lmfs_core",
              "-Wl,-rpath,$ORIGIN/../../../fastmbcui_amd/lib", f"-Wl,-rpath,{tdir}/lib"])
    if verbose and (r.stdout or r.stderr):
        sys.stderr.write(r.stdout + r.stderr)
    os.replace(tmp, SHIM)
    return SHIM


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        out = os.path.join(OBJ, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-o", out, os.path.join(CSRC, src)])
    for src in HOST_SOURCES:
        out = os.path.join(OBJ, src + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", "-o", out, os.path.join(CSRC, src)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    objs = [os.path.join(OBJ, s + ".o") for s in HIP_SOURCES + HOST_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_torch_shim(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
