"""Builds the oracle's compiled parts (test infrastructure only): the C +
OpenMP restatement of the cost and, where the reference tree is present,
the reference's own regularizer module (`make -C oracle ref`)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = "/root/reference/src/fortlib"


def build_all():
    mk = os.path.join(HERE, "Makefile")
    if os.path.exists(mk):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        if os.path.exists(os.path.join(REF_SRC, "image.f90")):
            subprocess.run(["make", "-s", "-C", HERE, "ref", "REF=" + REF_SRC], check=True)
