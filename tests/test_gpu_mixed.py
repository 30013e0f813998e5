"""Mixed hex/penta/tet mesh end to end on the GPU: device numbering and
assembly, GRAV + FORCE load vectors, mixed-precision block PCG over the load
cases, and the compliance adjoint, each against the CPU oracle."""
import numpy as np
import pytest

import oracle as O
from vibestran_amd import decks
from vibestran_amd import model as M
from vibestran_amd.assembly import (assemble_model, assemble_rhs_device, dof_map_device,
                                    element_sensitivity)

pytestmark = pytest.mark.gpu


def test_mixed_block_end_to_end(gpu_backend_factory):
    m = decks.mixed_block(nx=10, ny=6, nz=5, n_cases=3)
    be = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=50000)
    eq, n = dof_map_device(be, m)
    eq_o, n_o = M.analysis_dof_map(m)
    assert n == n_o and np.array_equal(eq, eq_o)
    K = assemble_model(be, m, eq, n).fetch()
    oc = O.assemble_model(n, m, eq)
    rp, ci, v = oc.arrays()
    assert np.array_equal(K.row_ptr, rp) and np.array_equal(K.col_ind, ci)
    assert np.array_equal(K.values.view(np.int64), v.view(np.int64))
    F = np.stack([assemble_rhs_device(be, m, eq, n, lc) for lc in m.loads], axis=1)
    for j, lc in enumerate(m.loads):
        Fo = O.grav_rhs(n, m, eq, lc.grav) + M.assemble_rhs(m, eq, n, M.LoadCase(nodal=lc.nodal))
        assert np.abs(F[:, j] - Fo).max() <= 1e-12 * np.abs(Fo).max()
    U = be.solve_block(K, F)
    for j in range(F.shape[1]):
        assert O.true_relative_residual(oc, U[:, j], F[:, j]) < 1e-8
    g = element_sensitivity(be, m, 0, U[:, 0], U[:, 0])
    go = O.element_sensitivity(n, m, eq, 0, U[:, 0], U[:, 0])
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()
