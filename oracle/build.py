"""Builds the oracle's compiled parts (test infrastructure only)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build_all():
    mk = os.path.join(HERE, "Makefile")
    if os.path.exists(mk):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
