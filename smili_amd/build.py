"""Builds libsmili_hip.so in-tree for gfx950 (``python -m smili_amd.build``)."""
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


def build_hip(verbose=False):
    """Both libraries: the imaging cost (libsmili_hip.so) and the
    warm-start network (libsmili_unet.so)."""
    _hipcc(UNET_SRC, UNET_OUT, verbose)
    return _hipcc(SRC, OUT, verbose)


if __name__ == "__main__":
    print(build_hip(verbose=True))
