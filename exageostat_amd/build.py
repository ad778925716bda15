"""Build libexageostat_gpu.so in-tree (hipcc, gfx950)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, clean: bool = False, force: bool = False) -> str:
    """Runs the Makefile; force recompiles every source even where the
    objects are newer than their sources."""
    if clean:
        subprocess.run(["make", "-C", HERE, "clean"], check=True)
    cmd = ["make", "-C", HERE, f"-j{jobs}"] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)
    return os.path.join(HERE, "libexageostat_gpu.so")


if __name__ == "__main__":
    print(build())
