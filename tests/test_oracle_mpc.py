"""The oracle's MPC elimination (oracle/mpc.py, restating the reference's
MpcHandler, src/core/mpc_handler.cpp) pinned without restating it:

- the eliminated solution equals the Lagrange-multiplier solution of the
  constrained problem (K u + C^T lambda = F, C u = 0, a dense solve), which
  fixes T's signs, indices, the reduced numbering, T^T K T, T^T F and T u;
- the element-level T^T Ke T (apply_to_stiffness) sums to the global
  T^T K_full T;
- hand-derived cases for the rules the reference adds: a dependent DOF that
  is SPC-constrained skips its equation, a dependency cycle is an error, a
  term on another equation's dependent DOF drops out (mpc_handler.cpp:146-156);
- the reference's cquad4_ring_* e2e fixtures (oblique MPCs, CORD2C nodes)
  and, for CD frames, the slice deck rewritten with its oblique constraint as
  an SPC in a cylindrical CD frame, which must give the same solution."""
import os

import numpy as np
import pytest

import oracle as O
from e2e_csv import GOLD, deck_checks
from oracle.mpc import MpcOracle, assemble_model_mpc, nodal_rhs_mpc
from vibestran_amd import deck, decks
from vibestran_amd import model as M


def kkt_solution(model, eq, n, equations, nodal):
    K = O.assemble_model(n, model, eq).to_scipy().toarray()
    F = np.zeros(n)
    mask = eq >= 0
    np.add.at(F, eq[mask], nodal[mask])
    rows = []
    for terms in equations:
        r = np.zeros(n)
        for g, d, c in terms:
            if eq[g, d - 1] >= 0:
                r[eq[g, d - 1]] += c
        rows.append(r)
    C = np.array(rows)
    A = np.block([[K, C.T], [C, np.zeros((len(rows), len(rows)))]])
    x = np.linalg.solve(A, np.concatenate([F, np.zeros(len(rows))]))
    return x[:n], K, F


def eliminated_solution(model, eq, n, equations, nodal):
    mpc = MpcOracle(eq, n, equations)
    Kr = assemble_model_mpc(model, eq, mpc).to_scipy().toarray()
    Fr = nodal_rhs_mpc(model, eq, nodal, mpc)
    return mpc.recover(np.linalg.solve(Kr, Fr)), mpc, Kr, Fr


def random_mpcs(model, eq, rng, count):
    """Independent equations: dependents on one node set, other terms on
    another, the dependent term strictly largest."""
    free_nodes = [i for i in range(model.nn) if eq[i, 0] >= 0]
    rng.shuffle(free_nodes)
    deps, inds = free_nodes[:count], free_nodes[count:]
    eqs = []
    for g in deps:
        d = int(rng.integers(1, 4))
        terms = [(g, d, float(rng.choice([-1, 1]) * rng.uniform(2.0, 3.0)))]
        for _ in range(int(rng.integers(1, 4))):
            terms.append((int(rng.choice(inds)), int(rng.integers(1, 4)), float(rng.uniform(-1.5, 1.5))))
        eqs.append(terms)
    return eqs


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_elimination_matches_lagrange_multipliers(seed):
    rng = np.random.default_rng(seed)
    m = decks.hex_block(3, 2, 2, jitter=0.15, jitter_seed=seed)
    eq, n = M.analysis_dof_map(m)
    eqs = random_mpcs(m, eq, rng, 5)
    nodal = np.zeros((m.nn, 6))
    nodal[rng.integers(0, m.nn, 4), rng.integers(0, 3, 4)] = rng.uniform(-100, 100, 4)
    u_kkt, K, F = kkt_solution(m, eq, n, eqs, nodal)
    u, mpc, Kr, Fr = eliminated_solution(m, eq, n, eqs, nodal)
    assert mpc.n_red == n - len(eqs) and mpc.skipped == 0
    assert np.max(np.abs(u - u_kkt)) <= 1e-10 * np.max(np.abs(u_kkt))
    # the constraints hold exactly in the recovered vector's arithmetic
    for terms in eqs:
        assert abs(sum(c * u[eq[g, d - 1]] for g, d, c in terms)) <= 1e-12 * np.max(np.abs(u))
    T = mpc.transform()
    assert np.allclose(Kr, T.T @ K @ T, rtol=0, atol=1e-10 * np.abs(K).max())
    assert np.allclose(Fr, T.T @ F, rtol=0, atol=1e-12 * np.abs(F).max())


def tiny():
    m = decks.hex_block(1, 1, 1)
    eq, n = M.analysis_dof_map(m)
    return m, eq, n


def test_spc_constrained_dependent_is_skipped():
    m, eq, n = tiny()
    clamped = int(m.spc_nodes[0])
    free = [i for i in range(m.nn) if eq[i, 0] >= 0]
    mpc = MpcOracle(eq, n, [[(clamped, 1, 5.0), (free[0], 1, 1.0)]])
    assert mpc.skipped == 1 and mpc.n_red == n and not mpc.eliminations


def test_dependency_cycle_is_an_error():
    m, eq, n = tiny()
    a, b = [i for i in range(m.nn) if eq[i, 0] >= 0][:2]
    with pytest.raises(O.OracleError, match="MPC dependency graph contains a cycle"):
        MpcOracle(eq, n, [[(a, 1, 2.0), (b, 1, -1.0)], [(b, 1, 2.0), (a, 1, -1.0)]])


def test_chained_term_drops_out():
    # 2 u_a - u_b = 0 (a dependent), 3 u_b - u_c = 0 (b dependent): a's term
    # on b is removed after renumbering, so u_a is recovered as 0 and u_b as u_c / 3.
    m, eq, n = tiny()
    a, b, c = [i for i in range(m.nn) if eq[i, 0] >= 0][:3]
    mpc = MpcOracle(eq, n, [[(a, 1, 2.0), (b, 1, -1.0)], [(b, 1, 3.0), (c, 1, -1.0)]])
    assert mpc.n_red == n - 2
    assert mpc.t_column(int(eq[a, 0])) == []
    rc = mpc.index_map[int(eq[c, 0])]
    assert mpc.t_column(int(eq[b, 0])) == [(rc, 1.0 / 3.0)]
    u_red = np.arange(1.0, mpc.n_red + 1.0)
    u = mpc.recover(u_red)
    assert u[eq[a, 0]] == 0.0 and u[eq[b, 0]] == (1.0 / 3.0) * u_red[rc]


def node_dof_value(m, eq, u, node_id, dof):
    i = int(np.nonzero(m.node_ids == node_id)[0][0])
    e = eq[i, dof]
    return 0.0 if e < 0 else u[e]


def ring_solution(m):
    eq, n = M.analysis_dof_map(m)
    eq, n, mp = M.analysis_mpc_system(m, eq, n)
    mpc = MpcOracle(eq, n, mp.equations() if mp is not None else [])
    K = assemble_model_mpc(m, eq, mpc)
    F = nodal_rhs_mpc(m, eq, m.loads[0].nodal, mpc)
    u_red, _, _ = O.pcg_jacobi(K, F, tol=1e-14, max_iters=100000)
    return eq, mpc.recover(u_red), mpc


@pytest.mark.parametrize("name", ["cquad4_ring_60deg_slice", "cquad4_ring_6fold_full"])
def test_ring_fixtures(name):
    m = deck.read_deck(os.path.join(GOLD, name + ".bdf"))
    eq, u, mpc = ring_solution(m)
    assert mpc.n_red < mpc.n_full
    dofs = {"T1": 0, "T2": 1, "T3": 2, "R1": 3, "R2": 4, "R3": 5}
    cen, _ = O.recover_stress(m, eq, u, corners=False)
    ran = 0
    for c in deck_checks(name):
        if c["type"] == "node_displacement":
            v = node_dof_value(m, eq, u, c["node_id"], dofs[c["dof"]])
        elif c["type"] == "element_stress":
            v = cen[int(np.nonzero(m.elem_ids == c["elem_id"])[0][0]), 9]
        else:
            continue
        tol = c["rel_tol"] * abs(c["expected"]) if "rel_tol" in c else c["abs_tol"]
        assert abs(v - c["expected"]) <= tol, (c, v)
        ran += 1
    assert ran >= 6


def slice_with_cd_frame():
    """The 60-degree slice with its oblique MPCs replaced by CD = 10
    (cylindrical) on nodes 2 and 8 and an SPC of their tangential T2."""
    with open(os.path.join(GOLD, "cquad4_ring_60deg_slice.bdf")) as fh:
        lines = fh.read().splitlines()
    out = []
    for ln in lines:
        if ln.startswith("MPC,"):
            continue
        if ln.startswith("GRID,2,") or ln.startswith("GRID,8,"):
            ln = ln + ",10"
        out.append(ln)
        if ln.startswith("SPC1,1,2,1,7"):
            out.append("SPC1,1,2,2,8")
    return "\n".join(out) + "\n"


def test_cd_frame_spc_equals_oblique_mpc():
    ref = deck.read_deck(os.path.join(GOLD, "cquad4_ring_60deg_slice.bdf"))
    cdm = deck.read_deck(slice_with_cd_frame(), is_text=True)
    assert cdm.mpc is None and list(cdm.cd[[1, 3]]) == [10, 10]
    eq_r, u_r, _ = ring_solution(ref)
    eq_c, u_c, mpc_c = ring_solution(cdm)
    # the tangential SPC in the CD frame becomes the deck's oblique equation
    assert mpc_c.n_red == mpc_c.n_full - 2
    for nid in (1, 2, 7, 8):
        for d in range(6):
            a = node_dof_value(ref, eq_r, u_r, nid, d)
            b = node_dof_value(cdm, eq_c, u_c, nid, d)
            assert abs(a - b) <= 1e-12 * np.max(np.abs(u_r)), (nid, d, a, b)
    # output in the CD frame: radial displacement, zero tangential
    disp = np.zeros((cdm.nn, 6))
    mask = eq_c >= 0
    disp[mask] = u_c[eq_c[mask]]
    out = M.to_output_frames(cdm, disp)
    i8 = int(np.nonzero(cdm.node_ids == 8)[0][0])
    assert abs(out[i8, 1]) <= 1e-15 * np.abs(disp).max()
    assert abs(out[i8, 0] - np.hypot(disp[i8, 0], disp[i8, 1])) <= 1e-12 * np.abs(disp).max()


def plate_deck(rigid_cards, nx=4, ny=3):
    """CQUAD4 plate in the XY plane, x = 0 edge clamped, plus the given rigid
    element cards and their extra grids; loads on grid 9999. Lever arms stay
    below 1 so that every rigid equation keeps its unit-coefficient grid as
    the dependent term (the largest |coefficient| is chosen)."""
    lines = ["SOL 101", "CEND", "SUBCASE 1", "  LOAD = 1", "  SPC = 1", "BEGIN BULK",
             "MAT1,1,70000.,,.3", "PSHELL,1,1,0.5"]
    nid = lambda i, j: 1 + i + (nx + 1) * j
    for j in range(ny + 1):
        for i in range(nx + 1):
            lines.append(f"GRID,{nid(i, j)},,{1.0 * i},{0.25 * j + 0.02 * i * j},0.")
    e = 1
    for j in range(ny):
        for i in range(nx):
            lines.append(f"CQUAD4,{e},1,{nid(i, j)},{nid(i + 1, j)},{nid(i + 1, j + 1)},{nid(i, j + 1)}")
            e += 1
    lines.append("SPC1,1,123456," + ",".join(str(nid(0, j)) for j in range(ny + 1)))
    lines += rigid_cards(nid, nx, ny)
    lines += ["FORCE,1,9999,0,100.,0.3,-0.2,1.", "MOMENT,1,9999,0,50.,1.,0.5,-0.4", "ENDDATA"]
    return "\n".join(lines) + "\n"


def solve_model(m):
    eq, n = M.analysis_dof_map(m)
    eq, n, mp = M.analysis_mpc_system(m, eq, n)
    mpc = MpcOracle(eq, n, mp.equations())
    K = assemble_model_mpc(m, eq, mpc).to_scipy().toarray()
    u = mpc.recover(np.linalg.solve(K, nodal_rhs_mpc(m, eq, m.loads[0].nodal, mpc)))
    disp = np.zeros((m.nn, 6))
    disp[eq >= 0] = u[eq[eq >= 0]]
    return disp, {int(g): i for i, g in enumerate(m.node_ids)}


def test_rbe2_moves_its_grids_rigidly():
    cards = lambda nid, nx, ny: [
        "GRID,9999,,4.3,0.5,0.4",
        "RBE2,101,9999,123456," + ",".join(str(nid(nx, j)) for j in range(ny + 1)) + ",1.-6"]
    m = deck.read_deck(plate_deck(cards), is_text=True)
    assert m.rigid_mpc is not None and m.rigid_mpc.n == 6 * 4
    assert m.rigid_active[-1] == 63
    disp, idx = solve_model(m)
    h = idx[9999]
    scale = np.abs(disp).max()
    for j in range(4):
        g = idx[1 + 4 + 5 * j]
        r = m.xyz[g] - m.xyz[h]
        rigid = disp[h, :3] + np.cross(disp[h, 3:], r)
        assert np.abs(disp[g, :3] - rigid).max() <= 1e-10 * scale
        assert np.abs(disp[g, 3:] - disp[h, 3:]).max() <= 1e-10 * scale
    assert np.abs(disp[h]).max() > 0.0


def test_rbe3_reference_grid_follows_the_weighted_average():
    # group 1: weight 2, translations of two edge grids; group 2: weight 0.5,
    # all six components of two other edge grids (lever terms omega_j x r_j)
    cards = lambda nid, nx, ny: [
        "GRID,9999,,4.2,0.45,0.3",
        f"RBE3,201,,9999,123456,2.0,123,{nid(nx, 0)},{nid(nx, 1)},0.5,123456,{nid(nx, 2)},{nid(nx, 3)}"]
    m = deck.read_deck(plate_deck(cards), is_text=True)
    assert m.rigid_mpc.n == 6
    disp, idx = solve_model(m)
    h = idx[9999]
    g1 = [idx[5], idx[10]]
    g2 = [idx[15], idx[20]]
    W = 2.0 * 2 + 0.5 * 2
    avg = sum(2.0 * disp[g, :3] for g in g1) + sum(
        0.5 * (disp[g, :3] + np.cross(disp[g, 3:], m.xyz[g] - m.xyz[h])) for g in g2)
    scale = np.abs(disp).max()
    assert np.abs(disp[h, :3] - avg / W).max() <= 1e-10 * scale
    # rotations: only the group with rotational components, over W_rot
    assert np.abs(disp[h, 3:] - sum(0.5 * disp[g, 3:] for g in g2) / (0.5 * 2)).max() <= 1e-10 * scale
