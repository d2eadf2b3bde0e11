"""Builds the native libraries in-tree (``make -B`` in ``csrc/``)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs: int = 8, force: bool = True) -> None:
    cmd = ["make", "-C", CSRC, "-j%d" % jobs] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    build()
