"""GPU CQUAD4 (MITC4+ and Mindlin) assembly through the C ABI against the CPU
oracle, and the reference's shell e2e decks solved end to end on the GPU.

The device element matrices follow the oracle's operation order entry by
entry without floating-point contraction, and every CSR slot is summed in
element order on both sides, so the assembled lower CSR -- pattern after the
reference's exact-zero drop (sparse_matrix.cpp:22-33) and values -- must be
bit-identical to the oracle's."""
import json
import os

import numpy as np
import pytest

import oracle as O
from e2e_csv import uses_frames_or_mpc
from vibestran_amd import bdf, decks
from vibestran_amd import model as M
from vibestran_amd.assembly import assemble_model

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "e2e")
SHELL_DECKS = sorted(f[:-4] for f in os.listdir(GOLD)
                     if f.startswith("cquad4") and f.endswith(".bdf") and not uses_frames_or_mpc(f[:-4]))
DOFS = {"T1": 0, "T2": 1, "T3": 2, "R1": 3, "R2": 4, "R3": 5}


def csr_dict(rp, ci, v):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return dict(zip(zip(rows.tolist(), ci.tolist()), v.tolist()))


def compare(K, oc):
    """Bit-exact pattern and values; returns the number of differing slots (0)."""
    rp, ci, v = oc.arrays()
    a, b = csr_dict(K.row_ptr, K.col_ind, K.values), csr_dict(rp, ci, v)
    ndiff = len(set(a) ^ set(b))
    assert ndiff == 0, f"{ndiff} slots differ from the oracle's pattern"
    assert np.array_equal(K.row_ptr, rp) and np.array_equal(K.col_ind, ci)
    assert np.array_equal(K.values.view(np.int64), v.view(np.int64)), \
        f"values differ by up to {np.abs(K.values - v).max() / np.abs(v).max():.3e} of max|K|"
    return ndiff


def check_shell(be, m):
    eq, n = M.analysis_dof_map(m)
    K = assemble_model(be, m, eq, n).fetch()
    oc = O.assemble_model(n, m, eq)
    assert K.n == n
    compare(K, oc)
    return eq, n, K, oc


def warped_mesh(nx=6, ny=4, seed=8, form=0):
    """Jittered grid on the saddle z = 0.15 x y, clamped at x = 0."""
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    x = i.reshape(-1).astype(float)
    y = j.reshape(-1).astype(float)
    inner = (i.reshape(-1) > 0) & (i.reshape(-1) < nx) & (j.reshape(-1) > 0) & (j.reshape(-1) < ny)
    x = x + inner * rng.uniform(-0.15, 0.15, x.size)
    y = y + inner * rng.uniform(-0.15, 0.15, y.size)
    xyz = np.stack([x, y, 0.15 * x * y], axis=1)
    ie, je = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ie, je = ie.reshape(-1), je.reshape(-1)
    node = lambda a, b: a * (ny + 1) + b
    conn = np.stack([node(ie, je), node(ie + 1, je), node(ie + 1, je + 1), node(ie, je + 1)], axis=1)
    ne = len(conn)
    spc = np.nonzero(i.reshape(-1) == 0)[0].astype(np.int32)
    return M.Model(node_ids=np.arange(1, xyz.shape[0] + 1, dtype=np.int32) * 2, xyz=xyz,
                   elem_type=np.full(ne, M.CQUAD4, np.int32),
                   elem_ptr=(np.arange(ne + 1) * 4).astype(np.int32),
                   conn=conn.reshape(-1).astype(np.int32), E=rng.uniform(5e4, 9e4, ne),
                   nu=np.full(ne, 0.33), rho=np.zeros(ne), thickness=rng.uniform(0.05, 0.3, ne),
                   spc_nodes=spc, spc_masks=np.full(len(spc), 0b111111, np.int32), shell_form=form)


@pytest.mark.parametrize("name", SHELL_DECKS)
def test_shell_deck_assembly_and_solve(gpu_backend_factory, name):
    be = gpu_backend_factory(tolerance=1e-12, max_iters=200000)
    m = bdf.read_bdf(os.path.join(GOLD, name + ".bdf"))
    eq, n, K, oc = check_shell(be, m)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u = be.solve(K, F)
    uo, _, _ = O.pcg_jacobi(oc, F, tol=1e-12, max_iters=200000)
    assert np.abs(u - uo).max() <= 1e-8 * np.abs(uo).max()
    idx = {int(g): i for i, g in enumerate(m.node_ids)}
    exp = json.load(open(os.path.join(GOLD, name + ".expected.json")))
    for c in exp["checks"]:
        if c["type"] != "node_displacement":
            continue
        e = eq[idx[c["node_id"]], DOFS[c["dof"]]]
        val = 0.0 if e < 0 else u[e]
        tol = c["rel_tol"] * abs(c["expected"]) if "rel_tol" in c else c["abs_tol"]
        assert abs(val - c["expected"]) <= tol, (c, val)


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("size", [(24, 10), (200, 100)])
def test_tube_assembly_parity(gpu_backend_factory, form, size):
    be = gpu_backend_factory()
    m = decks.cquad4_tube(n_circ=size[0], n_axial=size[1], n_cases=2)
    m.shell_form = form
    check_shell(be, m)


@pytest.mark.parametrize("form", [0, 1])
def test_warped_shell_assembly_and_solve(gpu_backend_factory, form):
    be = gpu_backend_factory(tolerance=1e-11, max_iters=200000)
    m = warped_mesh(form=form)
    eq, n, K, oc = check_shell(be, m)
    F = np.random.default_rng(2).standard_normal(n)
    u = be.solve(K, F)
    uo, _, _ = O.pcg_jacobi(oc, F, tol=1e-11, max_iters=200000)
    assert np.abs(u - uo).max() <= 1e-7 * np.abs(uo).max()


def test_shell_error_paths(gpu_backend_factory):
    from vibestran_amd.backend import SolverError
    be = gpu_backend_factory()
    m = warped_mesh(nx=2, ny=2)
    eq, n = M.analysis_dof_map(m)
    bad = M.Model(**{**m.__dict__, "thickness": np.where(np.arange(m.ne) == 1, 0.0, m.thickness)})
    with pytest.raises(ValueError, match="non-positive thickness"):
        assemble_model(be, bad, eq, n)
    twisted = M.Model(**{**m.__dict__, "conn": m.conn.reshape(-1, 4)[:, [0, 2, 1, 3]].reshape(-1).copy()})
    with pytest.raises(SolverError, match="CQUAD4|Jacobian"):
        assemble_model(be, twisted, eq, n)
