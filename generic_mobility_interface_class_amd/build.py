"""Build the gfx950 library in-tree (hipcc) -- no JIT cache, no pip install.

    python -m generic_mobility_interface_class_amd.build

produces generic_mobility_interface_class_amd/libgmi.so.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libgmi.so")

HIP_SOURCES = [
    "gmi_pair.hip",
    "gmi_capi.hip",
    "gmi_sim.hip",
    "gmi_learn.hip",
    "gmi_io.hip",
    "gmi_seed.hip",
    "gmi_cilia.hip",
    "gmi_dense.hip",
    "gmi_noalloc.hip",
    "gmi_gemm.hip",
]
LINK_FLAGS = ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]

# Packed FP32 ops are disabled for every HIP source: the compiler would
# otherwise form v_pk_* from float code.
HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (not os.path.isabs(cand) or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *srcs, *LINK_FLAGS]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgmi.so")
    os.replace(tmp, LIB)
    return LIB


DRIVER = os.path.join(PKG, "gmi_drive")


def build_driver(verbose: bool = False) -> str:
    """C++ driver of the host class (host/gmi_drive.cpp), linked to libgmi.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(PKG, "host", "gmi_drive.cpp"), "-o", DRIVER + ".tmp",
           f"-L{PKG}", "-lgmi", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("driver build failed")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def build_oracle(verbose: bool = False) -> str:
    od = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-C", od, "-B", "liboracle.so"], capture_output=True, text=True)
    if verbose or r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed")
    return os.path.join(od, "liboracle.so")


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_driver(verbose=True))
