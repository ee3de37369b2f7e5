"""Build libfresco_hip.so (gfx950) in-tree with hipcc via fresco_amd/csrc/Makefile."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(force: bool = True, jobs: int = 8) -> str:
    args = ["make", "-C", CSRC, "-j%d" % jobs]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return os.path.join(HERE, "libfresco_hip.so")


if __name__ == "__main__":
    print(build())
