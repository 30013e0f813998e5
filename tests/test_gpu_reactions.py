"""GPU element internal forces (vbh_element_forces) and reactions
(vbh_reactions) against the oracle restatement, and the equilibrium of a
solved model: at free DOFs R is the (small) solve residual, and the reactions
balance every applied load (FORCE, PLOAD4, GRAV)."""
import os

import numpy as np
import pytest

import oracle as O
from e2e_csv import GOLD
from vibestran_amd import bdf, decks
from vibestran_amd import model as M
from vibestran_amd import results as R
from vibestran_amd.assembly import assemble_model, assemble_rhs_device, dof_map_device
from tests.test_gpu_assembly import mixed_mesh
from tests.test_gpu_shell import warped_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mk", [
    lambda: decks.hex_block(5, 3, 3, jitter=0.15, jitter_seed=6),
    lambda: M.Model(**{**decks.hex_block(4, 3, 2, jitter=0.1).__dict__, "solid_form": 1}),
    mixed_mesh,
    lambda: warped_mesh(form=0), lambda: warped_mesh(form=1),
    lambda: decks.cquad4_tube(n_circ=12, n_axial=6, n_cases=1),
])
def test_element_forces_match_oracle(gpu_backend_factory, mk):
    m = mk()
    be = gpu_backend_factory()
    eq, n = dof_map_device(be, m)
    assemble_model(be, m, eq, n)
    u = np.random.default_rng(8).standard_normal(n)
    f = R.element_forces(be, m, u)
    fo = O.element_forces(m, eq, u)
    assert np.abs(f - fo).max() <= 1e-11 * np.abs(fo).max()


def test_axial_cube_reactions(gpu_backend_factory):
    m = bdf.read_bdf(os.path.join(GOLD, "chexa8_axial_force.bdf"))
    be = gpu_backend_factory(tolerance=1e-12, max_iters=10000)
    eq, n = dof_map_device(be, m)
    K = assemble_model(be, m, eq, n)
    u = be.solve(K, assemble_rhs_device(be, m, eq, n, m.loads[0]))
    Rn = R.reactions(be, m, u, m.loads[0])
    idx = {int(g): i for i, g in enumerate(m.node_ids)}
    assert np.allclose(Rn[[idx[g] for g in (1, 2, 3, 4)], 2], -250.0, rtol=1e-8)
    assert np.abs(Rn[eq >= 0]).max() < 1e-7 * 250.0


def test_reactions_balance_force_pressure_gravity(gpu_backend_factory):
    m = decks.hex_block(8, 3, 3, jitter=0.1, jitter_seed=3, E_sigma=0.2)
    ijk = m.meta["ijk"]
    nodal = np.zeros((m.nn, 6))
    nodal[ijk[:, 0] == 8, 2] = -10.0
    top = np.nonzero(np.arange(m.ne) % 3 == 2)[0]
    pl = M.Pload4(elem=top, g1=np.full(len(top), 4), g34=np.full(len(top), 6),
                  p=np.full((len(top), 4), 0.5))
    case = M.LoadCase(nodal=nodal, pload4=pl, grav=np.array([0.0, -9.81e3, 0.0]))
    be = gpu_backend_factory(tolerance=1e-12, max_iters=20000)
    eq, n = dof_map_device(be, m)
    K = assemble_model(be, m, eq, n)
    F = assemble_rhs_device(be, m, eq, n, case)
    u = be.solve(K, F)
    Rn = R.reactions(be, m, u, case)
    full = np.arange(6 * m.nn, dtype=np.int32).reshape(m.nn, 6)
    F_full = M.assemble_rhs(m, full, 6 * m.nn, M.LoadCase(nodal=nodal, pload4=pl)).reshape(m.nn, 6)
    F_full[:, :3] += O.grav_rhs(6 * m.nn, m, full, case.grav).reshape(m.nn, 6)[:, :3]
    scale = np.abs(F_full).max()
    assert np.abs(Rn[eq >= 0]).max() < 1e-7 * scale              # residual at free DOFs
    clamp = ijk[:, 0] == 0
    assert np.abs(Rn[~clamp]).max() < 1e-7 * scale
    tot = Rn[:, :3].sum(axis=0) + F_full[:, :3].sum(axis=0)      # force balance
    assert np.abs(tot).max() < 1e-8 * np.abs(F_full[:, :3].sum(axis=0)).max()
    Ro = O.element_forces(m, eq, u) - F_full
    assert np.abs(Rn - Ro).max() < 1e-9 * scale
