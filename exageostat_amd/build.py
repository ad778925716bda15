"""Build libexageostat_gpu.so in-tree (hipcc, gfx950)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, clean: bool = False) -> str:
    if clean:
        subprocess.run(["make", "-C", HERE, "clean"], check=True)
    subprocess.run(["make", "-C", HERE, f"-j{jobs}"], check=True)
    return os.path.join(HERE, "libexageostat_gpu.so")


if __name__ == "__main__":
    print(build())
