"""Build libusct.so (the solver) and libusct_bw.so (HBM streaming kernels) for
gfx950 with hipcc, and the host-only dataset-file library libusct_h5.so with
g++ (no cmake, no torch extension)."""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "usct.hip")
OUT = os.path.join(HERE, "libusct.so")
H5_SRC = os.path.join(HERE, "csrc", "usct_h5.cpp")
H5_OUT = os.path.join(HERE, "libusct_h5.so")
BW_SRC = os.path.join(HERE, "csrc", "usct_bw.hip")
BW_OUT = os.path.join(HERE, "libusct_bw.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",                     # op order of the arithmetic is part of the contract
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def build(verbose=False):
    """Compile the HIP library in place (always recompiles)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, OUT)
    build_bw(verbose)
    build_h5(verbose)
    return OUT


def build_bw(verbose=False):
    """Compile the HBM streaming kernels (libusct_bw.so) that measure the device's
    bandwidth ceiling and calibrate the traffic counters."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = BW_OUT + ".tmp"
    r = subprocess.run([hipcc, *HIPCC_FLAGS, "-o", tmp, BW_SRC], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, BW_OUT)
    return BW_OUT


def build_h5(verbose=False):
    """Compile the dataset-file reader/writer (host code: OpenMP + zlib)."""
    cxx = os.environ.get("CXX", "g++")
    tmp = H5_OUT + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", "-o", tmp, H5_SRC, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, H5_OUT)
    return H5_OUT


if __name__ == "__main__":
    print(build(verbose=True))
