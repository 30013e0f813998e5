"""The C++ host mirror (include/vbh_backend.hpp) on the GPU: solves, block
solve, SolverError paths and mixed precision, run as its own program."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
CPP_TEST = os.path.join(os.path.dirname(__file__), "cpp", "backend_test.bin")


def test_cpp_backend_program():
    r = subprocess.run([CPP_TEST], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout + r.stderr
