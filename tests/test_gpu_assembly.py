"""GPU DOF numbering and assembly through the C ABI against the CPU oracle:
DOF map, CSR pattern and values bit-identical (element matrices in the
oracle's operation order, slots summed in element order)."""
import json
import os

import numpy as np
import pytest

import oracle as O
from vibestran_amd import bdf, decks
from vibestran_amd import model as M
from vibestran_amd.assembly import (assemble_model, assemble_rhs_device, assembly_stats,
                                    dof_map_device)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "e2e")


def oracle_K(m, eq, n):
    form = "SRI" if m.solid_form == 1 else "EAS"
    return O.assemble_solids(n, m.elem_type, m.elem_ptr, m.conn, m.xyz, eq, m.E, m.nu, form)


def tet_mesh(nx=3, ny=2, nz=2, seed=4):
    """5-tet split of a jittered hex block (alternating split for conformity)."""
    hm = decks.hex_block(nx, ny, nz, jitter=0.1, jitter_seed=seed)
    c = hm.conn2d()
    A = [[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]]
    B = [[0, 1, 2, 5], [0, 2, 3, 7], [0, 5, 7, 4], [2, 5, 6, 7], [0, 2, 5, 7]]
    tets = []
    for e in range(len(c)):
        i, j, k = np.unravel_index(e, (nx, ny, nz))
        for t in (A if (i + j + k) % 2 == 0 else B):
            q = c[e, t]
            X = hm.xyz[q]
            if np.linalg.det(np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) < 0:
                q = q[[0, 2, 1, 3]]
            tets.append(q)
    tets = np.array(tets, np.int32)
    m = M.Model(node_ids=hm.node_ids, xyz=hm.xyz, elem_type=np.full(len(tets), M.CTETRA4, np.int32),
                elem_ptr=(np.arange(len(tets) + 1) * 4).astype(np.int32), conn=tets.reshape(-1),
                E=np.full(len(tets), 7e4), nu=np.full(len(tets), 0.33), rho=np.zeros(len(tets)),
                spc_nodes=hm.spc_nodes, spc_masks=hm.spc_masks)
    return m


def wedge_mesh(nx=3, ny=2, nz=2):
    hm = decks.hex_block(nx, ny, nz, jitter=0.08, jitter_seed=9)
    c = hm.conn2d()
    w = np.concatenate([c[:, [0, 1, 2, 4, 5, 6]], c[:, [0, 2, 3, 4, 6, 7]]])
    ne = len(w)
    return M.Model(node_ids=hm.node_ids, xyz=hm.xyz, elem_type=np.full(ne, M.CPENTA6, np.int32),
                   elem_ptr=(np.arange(ne + 1) * 6).astype(np.int32), conn=w.reshape(-1).astype(np.int32),
                   E=np.full(ne, 2e5), nu=np.full(ne, 0.3), rho=np.zeros(ne),
                   spc_nodes=hm.spc_nodes, spc_masks=hm.spc_masks)


def mixed_mesh():
    """Hexa core, wedge and tet regions sharing nodes (one model)."""
    hm = decks.hex_block(4, 2, 2, jitter=0.1, jitter_seed=12, E_sigma=0.2)
    c = hm.conn2d()
    ijk = np.array(np.unravel_index(np.arange(len(c)), (4, 2, 2))).T
    et, conn, E, nu = [], [], [], []
    for e in range(len(c)):
        if ijk[e, 0] < 2:
            et.append(M.CHEXA8); conn.extend(c[e]); E.append(hm.E[e]); nu.append(0.3)
        elif ijk[e, 0] == 2:
            for w in ([0, 1, 2, 4, 5, 6], [0, 2, 3, 4, 6, 7]):
                et.append(M.CPENTA6); conn.extend(c[e, w]); E.append(7e4); nu.append(0.33)
        else:
            for t in ([0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]):
                et.append(M.CTETRA4); conn.extend(c[e, t]); E.append(2e5); nu.append(0.3)
    ep = np.cumsum([0] + [M.NODES_PER[t] for t in et]).astype(np.int32)
    return M.Model(node_ids=hm.node_ids, xyz=hm.xyz, elem_type=np.array(et, np.int32), elem_ptr=ep,
                   conn=np.array(conn, np.int32), E=np.array(E), nu=np.array(nu),
                   rho=np.zeros(len(et)), spc_nodes=hm.spc_nodes, spc_masks=hm.spc_masks)


def check_assembly(be, m, form_override=None):
    eq, n = M.analysis_dof_map(m)
    K = assemble_model(be, m, eq, n).fetch()
    oc = oracle_K(m, eq, n)
    rp, ci, v = oc.arrays()
    assert K.n == n and K.nnz == len(ci)
    assert np.array_equal(K.row_ptr, rp) and np.array_equal(K.col_ind, ci)
    # element matrices in the oracle's operation order, slots summed in element order
    assert np.array_equal(K.values.view(np.int64), v.view(np.int64)), \
        np.abs(K.values - v).max() / np.abs(v).max()
    return eq, n, K, oc


def test_dof_map_device_bit_exact(gpu_backend_factory):
    be = gpu_backend_factory()
    m = decks.hex_block(6, 3, 4)
    rng = np.random.default_rng(3)
    m.node_ids = rng.permutation(np.arange(1, m.nn + 1) * 3).astype(np.int32)
    extra = rng.choice(m.nn, size=20, replace=False).astype(np.int32)
    m.spc_nodes = np.concatenate([m.spc_nodes, extra])
    m.spc_masks = np.concatenate([m.spc_masks, rng.integers(1, 64, size=20).astype(np.int32)])
    eq_d, n_d = dof_map_device(be, m)
    eq_o, n_o = O.analysis_dof_map(m.node_ids, m.elem_type, m.elem_ptr, m.node_ids[m.conn],
                                   m.node_ids[m.spc_nodes], m.spc_masks)
    assert n_d == n_o and np.array_equal(eq_d, eq_o)


@pytest.mark.parametrize("form", [0, 1])
def test_config1_assembly_parity(gpu_backend_factory, form):
    be = gpu_backend_factory()
    m = decks.config1_hex_cantilever()
    m.solid_form = form
    eq, n, K, oc = check_assembly(be, m)
    st = assembly_stats(be)
    assert st.nnz == oc.nnz and st.nnz_topological >= st.nnz


def test_jittered_lognormal_assembly_parity(gpu_backend_factory):
    be = gpu_backend_factory()
    m = decks.config2_hex_jitter(nx=8, ny=5, nz=4, seed=5)
    check_assembly(be, m)


@pytest.mark.parametrize("mk", [tet_mesh, wedge_mesh, mixed_mesh])
def test_other_solids_assembly_parity(gpu_backend_factory, mk):
    be = gpu_backend_factory()
    check_assembly(be, mk())


def test_partial_spc_deck_assembly_and_solve(gpu_backend_factory):
    be = gpu_backend_factory(tolerance=1e-12)
    m = bdf.read_bdf(os.path.join(GOLD, "chexa8_axial_force.bdf"))
    eq, n, K, oc = check_assembly(be, m)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u = be.solve(K, F)
    idx = {int(g): i for i, g in enumerate(m.node_ids)}
    for g in (5, 6, 7, 8):
        assert abs(u[eq[idx[g], 2]] - 1e-3) <= 1e-6 * 1e-3
    assert abs(u[eq[idx[6], 0]] + 3e-4) <= 1e-6 * 3e-4


def test_assembled_solve_matches_oracle(gpu_backend_factory):
    be = gpu_backend_factory(tolerance=1e-10, max_iters=20000)
    m = decks.config1_hex_cantilever()
    eq, n = M.analysis_dof_map(m)
    K = assemble_model(be, m, eq, n)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u = be.solve(K, F)
    ud = np.load(os.path.join(os.path.dirname(GOLD), "c1_direct.npz"))["u"]
    assert np.abs(u - ud).max() / np.abs(ud).max() < 1e-8


@pytest.mark.parametrize("mesh", ["hex-jitter", "tube", "mixed"])
def test_assembly_is_deterministic(gpu_backend_factory, mesh):
    """Every CSR slot is summed in element order (no atomics), so two
    assemblies of the same mesh -- in fresh contexts -- are bit-identical, and
    so are the mixed-precision solves on them (iterations and u)."""
    m = {"hex-jitter": lambda: decks.config2_hex_jitter(nx=24, ny=10, nz=10, seed=8),
         "tube": lambda: decks.cquad4_tube(n_circ=24, n_axial=30, n_cases=1),
         "mixed": lambda: decks.mixed_block(nx=6, ny=4, nz=4, n_cases=1)}[mesh]()
    runs = []
    F = None
    for _ in range(2):
        be = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=100000)
        eq, n = dof_map_device(be, m)
        K = assemble_model(be, m, eq, n)
        Kh = K.fetch()
        if F is None:  # one load vector for both solves (GRAV is assembled on the device)
            F = assemble_rhs_device(be, m, eq, n, m.loads[0])
        u = be.solve(K, F)
        runs.append((Kh.row_ptr.copy(), Kh.col_ind.copy(), Kh.values.copy(), u,
                     be.last_iteration_count()))
        be.close()
    (rp0, ci0, v0, u0, it0), (rp1, ci1, v1, u1, it1) = runs
    assert np.array_equal(rp0, rp1) and np.array_equal(ci0, ci1)
    assert np.array_equal(v0.view(np.int64), v1.view(np.int64))
    assert it0 == it1 and np.array_equal(u0, u1)


def _partial_spc(m, seed, masks):
    rng = np.random.default_rng(seed)
    extra = rng.choice(m.nn, size=max(1, m.nn // 4), replace=False).astype(np.int32)
    m.spc_nodes = np.concatenate([m.spc_nodes, extra]).astype(np.int32)
    m.spc_masks = np.concatenate([m.spc_masks, rng.choice(masks, size=len(extra))]).astype(np.int32)
    return m


@pytest.mark.parametrize("mesh", ["hex-shuffled-spc", "tube-spc", "mixed", "tets"])
def test_node_and_pair_patterns_identical(gpu_backend_factory, monkeypatch, mesh):
    """The node-pair pattern pass (meshes whose nodes all carry 3 or 6 DOFs)
    and the element-pair pass (VBH_PATTERN=dof) give the same lower CSR, bit
    for bit, and the same topological count, on meshes with shuffled node IDs,
    partial SPCs (non-contiguous active DOF sets), shells and several element
    kinds."""
    def hex_shuffled():
        m = decks.config2_hex_jitter(nx=9, ny=4, nz=5, seed=4)
        m.node_ids = np.random.default_rng(2).permutation(np.arange(1, m.nn + 1) * 7).astype(np.int32)
        return _partial_spc(m, 6, [1, 2, 4, 3, 5, 6])
    m = {"hex-shuffled-spc": hex_shuffled,
         "tube-spc": lambda: _partial_spc(decks.cquad4_tube(n_circ=12, n_axial=9, n_cases=1), 8,
                                          [1, 8, 16, 32, 9, 34, 56]),
         "mixed": mixed_mesh, "tets": tet_mesh}[mesh]()
    eq, n = M.analysis_dof_map(m)
    out = []
    for mode in ("node", "dof"):
        if mode == "dof":
            monkeypatch.setenv("VBH_PATTERN", "dof")
        else:
            monkeypatch.delenv("VBH_PATTERN", raising=False)
        be = gpu_backend_factory()
        K = assemble_model(be, m, eq, n).fetch()
        out.append((K.row_ptr.copy(), K.col_ind.copy(), K.values.copy(),
                    assembly_stats(be).nnz_topological))
        be.close()
    (rp0, ci0, v0, t0), (rp1, ci1, v1, t1) = out
    assert np.array_equal(rp0, rp1) and np.array_equal(ci0, ci1) and t0 == t1
    assert np.array_equal(v0.view(np.int64), v1.view(np.int64))
    if mesh != "tube-spc":
        rp, ci, v = oracle_K(m, eq, n).arrays()
        assert np.array_equal(rp0, rp) and np.array_equal(ci0, ci)


def _random_solid_mesh(seed):
    """Jittered hex block whose cells become a random mix of CHEXA8, two
    CPENTA6 and five CTETRA4 (conforming alternate splits), with some cells
    left out (their nodes may then belong to no element), shuffled node IDs
    and random partial SPCs."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in rng.integers(2, 5, size=3))
    hm = decks.hex_block(nx, ny, nz, jitter=0.1, jitter_seed=seed)
    c = hm.conn2d()
    ijk = np.array(np.unravel_index(np.arange(len(c)), (nx, ny, nz))).T
    kinds = rng.choice(4, size=len(c), p=[0.4, 0.2, 0.3, 0.1])  # hex, wedges, tets, none
    et, conn = [], []
    for e in range(len(c)):
        if kinds[e] == 0:
            et.append(M.CHEXA8); conn.extend(c[e])
        elif kinds[e] == 1:
            for w in ([0, 1, 2, 4, 5, 6], [0, 2, 3, 4, 6, 7]):
                et.append(M.CPENTA6); conn.extend(c[e, w])
        elif kinds[e] == 2:
            split = ([[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]]
                     if ijk[e].sum() % 2 == 0 else
                     [[0, 1, 2, 5], [0, 2, 3, 7], [0, 5, 7, 4], [2, 5, 6, 7], [0, 2, 5, 7]])
            for t in split:
                q = c[e, t]
                X = hm.xyz[q]
                if np.linalg.det(np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) < 0:
                    q = q[[0, 2, 1, 3]]
                et.append(M.CTETRA4); conn.extend(q)
    if not et:
        et.append(M.CHEXA8); conn.extend(c[0])
    ep = np.cumsum([0] + [M.NODES_PER[t] for t in et]).astype(np.int32)
    ne = len(et)
    m = M.Model(node_ids=rng.permutation(np.arange(1, hm.nn + 1) * 3).astype(np.int32), xyz=hm.xyz,
                elem_type=np.array(et, np.int32), elem_ptr=ep, conn=np.array(conn, np.int32),
                E=rng.uniform(1e5, 3e5, ne), nu=np.full(ne, 0.3), rho=np.zeros(ne),
                spc_nodes=hm.spc_nodes, spc_masks=hm.spc_masks)
    return _partial_spc(m, seed + 100, [1, 2, 4, 3, 5, 6, 7])


@pytest.mark.parametrize("seed", range(6))
def test_node_pair_pattern_random_meshes(gpu_backend_factory, monkeypatch, seed):
    """Randomised meshes for the node-pair pattern pass: solids of random
    kinds with holes, shuffled IDs and partial SPCs (even seeds), shell tubes
    of random size with partial SPCs (odd seeds); the pass and the
    element-pair pass agree bit for bit, and solids match the oracle's CSR."""
    if seed % 2 == 0:
        m = _random_solid_mesh(seed)
    else:
        rng = np.random.default_rng(seed)
        m = _partial_spc(decks.cquad4_tube(n_circ=int(rng.integers(6, 14)),
                                           n_axial=int(rng.integers(3, 9)), n_cases=1),
                         seed + 200, [1, 2, 8, 16, 32, 9, 34, 56, 63])
    eq, n = M.analysis_dof_map(m)
    if n == 0:
        pytest.skip("every DOF constrained")
    out = []
    for mode in ("node", "dof"):
        if mode == "dof":
            monkeypatch.setenv("VBH_PATTERN", "dof")
        else:
            monkeypatch.delenv("VBH_PATTERN", raising=False)
        be = gpu_backend_factory()
        K = assemble_model(be, m, eq, n).fetch()
        out.append((K.row_ptr.copy(), K.col_ind.copy(), K.values.copy(),
                    assembly_stats(be).nnz_topological))
        be.close()
    (rp0, ci0, v0, t0), (rp1, ci1, v1, t1) = out
    assert np.array_equal(rp0, rp1) and np.array_equal(ci0, ci1) and t0 == t1
    assert np.array_equal(v0.view(np.int64), v1.view(np.int64))
    if seed % 2 == 0:
        rp, ci, v = oracle_K(m, eq, n).arrays()
        assert np.array_equal(rp0, rp) and np.array_equal(ci0, ci)
        assert np.array_equal(v0.view(np.int64), v.view(np.int64))
