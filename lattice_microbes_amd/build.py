"""Builds the HIP library in-tree for gfx950.

hipcc flags:
  -ffp-contract=off   every float/double operation in the kernels is the
                      single IEEE operation written in the source (the results
                      are compared bit for bit with a serial restatement)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed-FP32 instructions are generated
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liblmx_mpd.so")
# The compiler's per-kernel resource report (VGPRs, scratch, occupancy) of
# the last build, for tests/test_kernel_resources.py.
RESOURCES = os.path.join(OUT_DIR, "kernel_resources.txt")
IO_LIB = os.path.join(OUT_DIR, "liblmx_io.so")

SOURCES = ["mpd_kernels.hip", "mpd_fast.hip", "mpd_capi.cpp"]
HEADERS = ["lmx_device.h", "lmx_kernels.h", "lmx_site.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
# The HDF5 C library (1.10, with the high-level API) of this image.
HDF5_PREFIX = os.environ.get("HDF5_PREFIX", "/opt/conda")
ARCH = "gfx950"

COMMON_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def build_library(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *COMMON_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-shared", "-Wl,-z,defs",
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building liblmx_mpd.so")
    os.replace(tmp, LIB)
    with open(RESOURCES, "w") as fh:
        fh.write("\n".join(l for l in res.stderr.splitlines() if "kernel-resource-usage" in l) + "\n")
    return LIB


def kernel_resources(path: str = RESOURCES) -> dict:
    """{mangled kernel name: {"VGPRs": n, "ScratchSize": bytes per lane,
    "Occupancy": waves per SIMD, ...}} from the last build's report."""
    out, cur = {}, None
    with open(path) as fh:
        for line in fh:
            text = line.split("remark:", 1)[-1].split("[-Rpass", 1)[0].strip()
            if text.startswith("Function Name:"):
                cur = out.setdefault(text.split(":", 1)[1].strip(), {})
            elif cur is not None and ":" in text:
                key, val = text.rsplit(":", 1)
                key = key.split("[", 1)[0].strip()
                try:
                    cur[key] = int(val.strip())
                except ValueError:
                    cur[key] = val.strip()
    return out


def build_io_library(verbose: bool = False) -> str:
    """The SimulationFile (HDF5) library: host C++ only, built with g++."""
    inc, libdir = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5_hl.h")):
        raise RuntimeError(f"HDF5 headers not found under {inc}; set HDF5_PREFIX")
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = IO_LIB + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
           "-I" + inc, os.path.join(CSRC, "lmx_io.cpp"), "-o", tmp, "-L" + libdir, "-lhdf5_hl", "-lhdf5",
           "-Wl,-rpath," + libdir, "-Wl,-z,defs"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building liblmx_io.so")
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_io_library(verbose=True))
