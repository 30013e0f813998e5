import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvbh on the GPU)")


@pytest.fixture(scope="session")
def gpu_backend_factory():
    """Creates HIP backends; fails loudly when the HIP extension is missing."""
    from vibestran_amd import HipMixedPcgSolverBackend, HipPcgSolverBackend, _lib

    _lib.load()  # raises if libvbh.so is absent

    def make(mixed=False, **kw):
        cls = HipMixedPcgSolverBackend if mixed else HipPcgSolverBackend
        be = cls.try_create(**kw)
        if be is None:
            pytest.fail("no HIP device visible for a gpu-marked test")
        return be

    return make
