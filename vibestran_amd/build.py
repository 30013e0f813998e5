"""Build step for libvbh.so (hipcc, gfx950) -- run by __graft_entry__.build()."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# The reference's include tree: the SolverBackend subclass of
# include/hip_pcg_solver_backend.hpp is compiled against its headers where they
# lie (nothing is copied); absent on machines without the reference.
REFERENCE_INCLUDE = "/root/reference/include"
BOUNDARY_SRC = os.path.join(ROOT, "tests", "cpp", "reference_boundary_test.cpp")
BOUNDARY_BIN = os.path.join(ROOT, "tests", "cpp", "reference_boundary_test.bin")


def build_reference_boundary(out=BOUNDARY_BIN) -> str | None:
    """Compiles tests/cpp/reference_boundary_test.cpp against the reference's
    headers and libvbh.so; returns the binary path, or None without the
    reference."""
    if not os.path.isdir(REFERENCE_INCLUDE):
        return None
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I" + REFERENCE_INCLUDE, "-I" + os.path.join(ROOT, "include"), BOUNDARY_SRC,
           "-o", out + ".tmp", "-L" + HERE, "-lvbh",
           "-Wl,-rpath,$ORIGIN/../../vibestran_amd:" + HERE, "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


def build(jobs: int = 8, force: bool = True) -> str:
    csrc = os.path.join(HERE, "csrc")
    if force:
        subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", csrc], check=True)
    build_reference_boundary()
    return os.path.join(HERE, "libvbh.so")
