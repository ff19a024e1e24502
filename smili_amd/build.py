"""Builds the package's native libraries in-tree (``python -m smili_amd.build``):
libsmili_hip.so and libsmili_unet.so (HIP, gfx950) and libsmili_lbfgsb.so
(the host optimizer, C++ / OpenMP)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "smili_abi.hip")
OUT = os.path.join(HERE, "libsmili_hip.so")
UNET_SRC = os.path.join(HERE, "csrc", "smili_unet.hip")
UNET_OUT = os.path.join(HERE, "libsmili_unet.so")
LBFGSB_SRC = os.path.join(HERE, "csrc", "smili_lbfgsb.cpp")
LBFGSB_OUT = os.path.join(HERE, "libsmili_lbfgsb.so")
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-mavx2", "-mfma"]

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc(src, out, verbose):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = out + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    if verbose:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        if msg.strip():
            print(msg)
    os.replace(tmp, out)
    return out


def build_lbfgsb(verbose=False):
    """The host L-BFGS-B (g++, no GPU code)."""
    cxx = os.environ.get("CXX", "g++")
    tmp = LBFGSB_OUT + ".tmp"
    cmd = [cxx, *CXX_FLAGS, "-o", tmp, LBFGSB_SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed: " + " ".join(cmd))
    if verbose and (r.stdout + r.stderr).strip():
        print(r.stdout + r.stderr)
    os.replace(tmp, LBFGSB_OUT)
    return LBFGSB_OUT


def build_hip(verbose=False):
    """Every native library: the imaging cost (libsmili_hip.so), the
    warm-start network (libsmili_unet.so) and the host optimizer
    (libsmili_lbfgsb.so)."""
    build_lbfgsb(verbose)
    _hipcc(UNET_SRC, UNET_OUT, verbose)
    return _hipcc(SRC, OUT, verbose)


if __name__ == "__main__":
    print(build_hip(verbose=True))
