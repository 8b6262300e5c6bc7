"""Build libxnb_amd.so (HIP kernels + C ABI) in-tree for gfx950.

Every HIP source is compiled with -ffp-contract=off (FP64 arithmetic in source order, the
bitwise contract with the reference formulation) and with packed FP32 instructions
disabled.  Objects go to exanbody_amd/_build/, the library to exanbody_amd/_lib/.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libxnb_amd.so")
ARCH = os.environ.get("XNB_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["xnb_kernels.hip", "xnb_brick.hip", "xnb_api.hip", "xnb_batch.hip"]
CXX_SOURCES = ["xnb_generate.cpp", "xnb_operators.cpp", "xnb_serve.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    for src in CXX_SOURCES:
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++"] + CXX_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
