"""GPU result recovery through the C ABI (vbh_recover_stress,
vbh_recover_displacements) against the oracle's recover_results restatement,
and the reference's e2e decks run end to end on the GPU (device assembly,
PCG, recovery) through the CSV files and the e2e runner's CSV checks.
At the full 605k-DOF hex size the check is the constant-strain patch: every
centroid and corner row of the jittered mesh equals the closed-form stress."""
import os

import numpy as np
import pytest

import oracle as O
from e2e_csv import GOLD, apply_checks, load_csv_results
from vibestran_amd import deck, decks
from vibestran_amd import model as M
from vibestran_amd import results as R
from vibestran_amd.assembly import (analysis_numbering, assemble_model, assemble_rhs_device,
                                    dof_map_device)
from tests.test_gpu_assembly import mixed_mesh
from tests.test_gpu_shell import warped_mesh

pytestmark = pytest.mark.gpu


def _parity(be, m, seed=3):
    eq, n = dof_map_device(be, m)
    assemble_model(be, m, eq, n)
    u = np.random.default_rng(seed).standard_normal(n) * 1e-3
    cen, cor = R.recover_stress(be, m, u)
    ocen, ocor = O.recover_stress(m, eq, u)
    for g, o in ((cen, ocen), (cor, ocor)):
        scale = np.abs(o).max(axis=tuple(range(o.ndim - 1)))  # per column
        assert np.all(np.abs(g - o) <= 1e-11 * scale + 1e-300), np.abs(g - o).max()
    disp = R.recover_displacements(be, m, u)
    assert np.array_equal(disp, np.where(eq >= 0, u[np.maximum(eq, 0)], 0.0))
    # centroid-only call gives the same centroid rows
    cen2, _ = R.recover_stress(be, m, u, corners=False)
    assert np.array_equal(cen2, cen)


@pytest.mark.parametrize("mk", [
    lambda: decks.hex_block(6, 4, 3, jitter=0.15, jitter_seed=4),
    lambda: M.Model(**{**decks.hex_block(5, 3, 3, jitter=0.1).__dict__, "solid_form": 1}),
    mixed_mesh,
    lambda: warped_mesh(form=0), lambda: warped_mesh(form=1),
    lambda: decks.cquad4_tube(n_circ=12, n_axial=6, n_cases=1),
    lambda: decks.mixed_block(nx=6, ny=4, nz=4, n_cases=1),
])
def test_recover_matches_oracle(gpu_backend_factory, mk):
    _parity(gpu_backend_factory(), mk())


def test_recover_needs_mesh(gpu_backend_factory):
    from vibestran_amd import _lib as L
    be = gpu_backend_factory()
    x = np.zeros(16)
    assert L.load().vbh_recover_stress(be._h, L.dp(x), L.dp(x), None) == L.VBH_ERR_STATE


DECKS = sorted(f[:-4] for f in os.listdir(GOLD) if f.endswith(".bdf"))


@pytest.mark.parametrize("name", DECKS)
def test_e2e_deck_on_gpu_through_csv(gpu_backend_factory, name, tmp_path):
    # the native reader and the analysis numbering with MPC elimination and
    # CD frames (the cquad4_ring_* decks have CORD2C nodes and oblique MPCs)
    m = deck.read_deck(os.path.join(GOLD, name + ".bdf"))
    be = gpu_backend_factory(tolerance=1e-12, max_iters=200000)
    eq, n, n_red = analysis_numbering(be, m)
    assert (n_red < n) == (m.mpc is not None)
    K = assemble_model(be, m, eq, n)
    assert K.n == n_red
    F = assemble_rhs_device(be, m, eq, n_red, m.loads[0])
    u = be.solve(K, F)
    disp = M.to_output_frames(m, R.recover_displacements(be, m, u))
    cen, _ = R.recover_stress(be, m, u, corners=False)
    nc, ec = R.write_csv(str(tmp_path / name), m, [(m.subcase_id(0), disp, cen)])
    node_data, elem_data = load_csv_results(nc, ec)
    assert apply_checks(name, node_data, elem_data) > 0


def test_full_size_hex_constant_strain(gpu_backend_factory):
    m = decks.config2_hex_jitter()
    be = gpu_backend_factory()
    eq, n = dof_map_device(be, m)
    assemble_model(be, m, eq, n)
    rng = np.random.default_rng(12)
    A = rng.normal(size=(3, 3)) * 1e-3
    field = m.xyz @ A.T
    u = np.zeros(n)
    for d in range(3):
        sel = eq[:, d] >= 0
        u[eq[sel, d]] = field[sel, d]
    # constrained nodes sit at x = 0 where an affine field would not vanish:
    # only elements away from the clamp see the pure field
    cen, cor = R.recover_stress(be, m, u)
    c2 = m.conn2d()
    free = np.all(m.xyz[c2, 0] > 0.0, axis=1)
    lam = m.E * m.nu / ((1 + m.nu) * (1 - 2 * m.nu))
    mu = m.E / (2 * (1 + m.nu))
    eps = 0.5 * (A + A.T)
    tr = np.trace(eps)
    want = np.stack([lam * tr + 2 * mu * eps[0, 0], lam * tr + 2 * mu * eps[1, 1],
                     2 * mu * eps[0, 1], lam * tr + 2 * mu * eps[2, 2], 2 * mu * eps[1, 2],
                     2 * mu * eps[0, 2]], axis=1)
    scale = np.abs(want).max()
    assert free.sum() > 0.95 * m.ne
    assert np.abs(cen[free, :6] - want[free]).max() < 1e-8 * scale
    assert np.abs(cor[free, :, :6] - want[free, None, :]).max() < 1e-8 * scale
