"""Builds the gfx950 library in-tree: dsmc_erlend_amd/libdsmc_hip.so (kernels + C-ABI) and the
C++ host driver dsmc_erlend_amd/dsmc_gpu. Run as `python -m dsmc_erlend_amd.build`."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdsmc_hip.so")
CLI = os.path.join(HERE, "dsmc_gpu")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]

HIP_SOURCES = ["dsmc_kernels.hip", "dsmc_capi.hip", "surrogate_kernels.hip", "surrogate_capi.hip"]
CXX_SOURCES = ["dsmc_writer.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    cxx = shutil.which("g++") or "g++"
    for src in CXX_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        _run([cxx, *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    host = os.path.join(CSRC, "host", "dsmc_main.cpp")
    if os.path.exists(host):
        _run([cxx, *CXX_FLAGS, "-o", CLI, host, os.path.join(CSRC, "host", "dsmc_host.cpp"),
              "-L" + HERE, "-ldsmc_hip", "-Wl,-rpath,$ORIGIN"], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
