"""Build step for libvbh.so (hipcc, gfx950) -- run by __graft_entry__.build()."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, force: bool = True) -> str:
    csrc = os.path.join(HERE, "csrc")
    if force:
        subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", csrc], check=True)
    return os.path.join(HERE, "libvbh.so")
