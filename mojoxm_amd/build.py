"""Compile libmxm_hip.so for gfx950 in-tree (mojoxm_amd/libmxm_hip.so)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int | None = None, force: bool = True) -> str:
    """Run the package Makefile; ``force`` rebuilds every object."""
    jobs = jobs or min(16, os.cpu_count() or 4)
    cmd = ["make", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(HERE, "libmxm_hip.so")


if __name__ == "__main__":
    print(build())
