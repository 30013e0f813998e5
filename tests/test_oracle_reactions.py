"""Element internal forces f = sum_e K_e u_e and reactions R = f - F (no
reference counterpart; the oracle restatement is pinned here):
  * rigid translations (and, for solids, rotations) are in every K_e's null
    space, so sum_i f_i = 0 and sum_i x_i x f_i = 0 for ANY u;
  * at free DOFs f equals the assembled CSR times u (K u);
  * the reference's chexa8_axial_force deck: the four base nodes carry
    R_z = -250 each (sigma_z A / 4), x/y reactions vanish; ctetra4_axial_force:
    base reactions total -1000 in z; at free DOFs R is the (tiny) residual."""
import os

import numpy as np
import pytest

import oracle as O
from e2e_csv import GOLD
from vibestran_amd import bdf, decks
from vibestran_amd import model as M
from tests.test_gpu_assembly import mixed_mesh
from tests.test_gpu_shell import warped_mesh


def all_free(m):
    """Equation map with every active DOF free (no SPC)."""
    mm = M.Model(**{**m.__dict__, "spc_nodes": np.zeros(0, np.int32),
                    "spc_masks": np.zeros(0, np.int32)})
    return mm, *M.analysis_dof_map(mm)


@pytest.mark.parametrize("mk,solid", [(lambda: decks.hex_block(3, 2, 2, jitter=0.15), True),
                                      (mixed_mesh, True),
                                      (lambda: warped_mesh(form=0), False),
                                      (lambda: decks.cquad4_tube(n_circ=10, n_axial=4, n_cases=1), False)])
def test_internal_forces_self_equilibrated(mk, solid):
    m, eq, n = all_free(mk())
    u = np.random.default_rng(3).standard_normal(n)
    f = O.element_forces(m, eq, u)
    scale = np.abs(f).max()
    assert np.abs(f[:, :3].sum(axis=0)).max() < 1e-9 * scale * m.nn
    if solid:
        mom = np.cross(m.xyz, f[:, :3]).sum(axis=0) + f[:, 3:].sum(axis=0)
        assert np.abs(mom).max() < 1e-9 * scale * m.nn * np.abs(m.xyz).max()


@pytest.mark.parametrize("mk", [lambda: decks.hex_block(3, 2, 2, jitter=0.15), mixed_mesh,
                                lambda: warped_mesh(form=1)])
def test_free_rows_are_K_times_u(mk):
    m = mk()
    eq, n = M.analysis_dof_map(m)
    u = np.random.default_rng(5).standard_normal(n)
    f = O.element_forces(m, eq, u)
    Ku = O.assemble_model(n, m, eq).multiply(u)
    free = eq >= 0
    assert np.abs(f[free] - Ku[eq[free]]).max() < 1e-11 * np.abs(Ku).max()


def solve_deck(name):
    m = bdf.read_bdf(os.path.join(GOLD, name + ".bdf"))
    eq, n = M.analysis_dof_map(m)
    u, _, _ = O.pcg_jacobi(O.assemble_model(n, m, eq), M.assemble_rhs(m, eq, n, m.loads[0]),
                           tol=1e-13, max_iters=100000)
    full = np.arange(6 * m.nn, dtype=np.int32).reshape(m.nn, 6)
    F_full = M.assemble_rhs(m, full, 6 * m.nn, m.loads[0]).reshape(m.nn, 6)
    R = O.element_forces(m, eq, u) - F_full
    return m, eq, R, F_full


def test_axial_cube_reactions_closed_form():
    m, eq, R, F = solve_deck("chexa8_axial_force")
    idx = {int(g): i for i, g in enumerate(m.node_ids)}
    base = [idx[g] for g in (1, 2, 3, 4)]
    assert np.allclose(R[base, 2], -250.0, rtol=1e-9)
    assert np.abs(R[:, :2]).max() < 1e-8 * 250.0
    free = eq >= 0
    assert np.abs(R[free]).max() < 1e-8 * 250.0
    assert np.allclose(R[:, :3].sum(axis=0), -F[:, :3].sum(axis=0), atol=1e-8 * 1000)


def test_tet_cube_reactions_balance_the_load():
    m, eq, R, F = solve_deck("ctetra4_axial_force")
    free = eq >= 0
    assert np.abs(R[free]).max() < 1e-8 * np.abs(F).max()
    assert np.allclose(R[:, :3].sum(axis=0), [0.0, 0.0, -1000.0], atol=1e-7)
