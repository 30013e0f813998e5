"""GPU load vectors through the C ABI: FORCE/MOMENT + PLOAD4 against the host
mirror and against the oracle's element face loads, GRAV (consistent mass)
against the oracle, on solids and shells."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from vibestran_amd import _lib as L, decks
from vibestran_amd import model as M
from vibestran_amd.assembly import assemble_model, assemble_rhs_device, dof_map_device
from tests.test_gpu_assembly import mixed_mesh

pytestmark = pytest.mark.gpu


def test_force_and_pload4_match_host(gpu_backend_factory):
    be = gpu_backend_factory()
    m = decks.config2_hex_jitter(nx=6, ny=4, nz=3, seed=3)
    m.loads[0].nodal[::7, 2] = np.linspace(-5, 5, len(m.loads[0].nodal[::7]))
    eq, n = dof_map_device(be, m)
    assemble_model(be, m, eq, n)
    F = assemble_rhs_device(be, m, eq, n, m.loads[0])
    Fh = M.assemble_rhs(m, eq, n, m.loads[0])
    assert np.abs(F - Fh).max() <= 1e-13 * np.abs(Fh).max()


def test_force_and_pload4_match_oracle(gpu_backend_factory):
    """FORCE + PLOAD4 on a jittered hex block, built independently of the host
    mirror: the oracle's CHEXA8 face load (linear_static_pressure.cpp:154-196,
    430-517) per loaded element, scattered through the equation map."""
    be = gpu_backend_factory()
    m = decks.config2_hex_jitter(nx=8, ny=5, nz=4, seed=11)
    pl = m.loads[0].pload4
    pl.p = np.random.default_rng(4).uniform(0.5, 2.0, (len(pl.elem), 4))
    m.loads[0].nodal[::5, 0] = np.linspace(-3, 3, len(m.loads[0].nodal[::5]))
    eq, n = dof_map_device(be, m)
    assemble_model(be, m, eq, n)
    F = assemble_rhs_device(be, m, eq, n, m.loads[0])
    nodal = np.array(m.loads[0].nodal, dtype=np.float64, copy=True)
    conn = m.conn2d()
    for i, e in enumerate(pl.elem):
        f8 = O.pload4_hexa8(m.xyz[conn[e]], int(pl.g1[i]), int(pl.g34[i]), pl.p[i])
        np.add.at(nodal[:, 0:3], conn[e], f8)
    Fo = np.zeros(n)
    mask = eq >= 0
    np.add.at(Fo, eq[mask], nodal[mask])
    assert np.abs(F - Fo).max() <= 1e-13 * np.abs(Fo).max()


@pytest.mark.parametrize("mk", [mixed_mesh, lambda: decks.cquad4_tube(n_circ=16, n_axial=8, n_cases=1)])
def test_grav_matches_oracle(gpu_backend_factory, mk):
    be = gpu_backend_factory()
    m = mk()
    m.rho = np.random.default_rng(2).uniform(1e-9, 8e-9, m.ne)
    eq, n = dof_map_device(be, m)
    assemble_model(be, m, eq, n)
    case = M.LoadCase(nodal=np.zeros((m.nn, 6)), grav=np.array([0.0, -3.0, -9810.0]))
    F = assemble_rhs_device(be, m, eq, n, case)
    Fo = O.grav_rhs(n, m, eq, case.grav)
    assert np.abs(F - Fo).max() <= 1e-12 * np.abs(Fo).max()


def test_rhs_into_device_column(gpu_backend_factory):
    be = gpu_backend_factory()
    m = mixed_mesh()
    m.rho = np.full(m.ne, 7.85e-9)
    eq, n = dof_map_device(be, m)
    assemble_model(be, m, eq, n)
    cases = [M.LoadCase(nodal=np.zeros((m.nn, 6)), grav=np.array([0.0, 0.0, -9810.0 * (j + 1)]))
             for j in range(3)]
    lib = L.load()
    buf = C.c_void_p()
    assert lib.vbh_device_alloc(be._h, 3 * n * 8, C.byref(buf)) == 0
    try:
        for j, cs in enumerate(cases):
            assemble_rhs_device(be, m, eq, n, cs, F_device=buf.value + j * n * 8)
        out = np.empty((n, 3), order="F")
        assert lib.vbh_memcpy_d2h(be._h, out.ctypes.data_as(C.c_void_p), buf, out.nbytes) == 0
    finally:
        lib.vbh_device_free(be._h, buf)
    Fo = O.grav_rhs(n, m, eq, cases[0].grav)
    for j in range(3):
        assert np.abs(out[:, j] - (j + 1) * Fo).max() <= 1e-12 * (j + 1) * np.abs(Fo).max()
