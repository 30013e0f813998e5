"""MPC elimination on the GPU (vbh_set_mpc, the device K_red = T^T K T,
F_red = T^T F, u_full = T u_red) against the oracle's element-level
restatement of the reference's MpcHandler (oracle/mpc.py):

- T is bit-identical (the same -a_j / a_dep arithmetic);
- K_red and F_red agree to 1e-12 of their largest entries (summation order
  differs: the reference transforms element by element, the device the
  assembled operator);
- the solved and expanded u agrees with the oracle's eliminated solution;
- models: random equations with a chained pair, tie + slide equations on a
  hex block, the 6-fold ring deck (oblique MPCs), a CQUAD4 plate with RBE2
  and RBE3 elements;
- error paths: a dependency cycle, a bad term, a numbering mismatch;
- at the 605k-DOF hex size (2.6k tie + 4.8k slide equations) the properties
  the elimination fixes: tied DOFs equal their master exactly, slide
  equations hold to rounding, and the reduced residual T^T (K u - F) meets
  the solver tolerance."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from oracle.mpc import MpcOracle, assemble_model_mpc, nodal_rhs_mpc
from vibestran_amd import _lib as L
from vibestran_amd import decks
from vibestran_amd import model as M
from vibestran_amd import results as R
from vibestran_amd.assembly import (analysis_numbering, assemble_model, assemble_rhs_device,
                                    assembly_stats, mpc_expand, mpc_stats, mpc_transform, set_mpc)
from vibestran_amd.backend import SolverError

pytestmark = pytest.mark.gpu


def random_mpc_block(seed=4):
    rng = np.random.default_rng(seed)
    m = decks.hex_block(4, 3, 3, jitter=0.15, jitter_seed=seed)
    eq, _ = M.analysis_dof_map(m)
    free = [i for i in range(m.nn) if eq[i, 0] >= 0]
    rng.shuffle(free)
    deps, inds = free[:12], free[12:]
    eqs = []
    for g in deps:
        terms = [(g, int(rng.integers(1, 4)), float(rng.choice([-1, 1]) * rng.uniform(2.0, 3.0)))]
        for _ in range(int(rng.integers(1, 4))):
            terms.append((int(rng.choice(inds)), int(rng.integers(1, 4)), float(rng.uniform(-1.5, 1.5))))
        eqs.append(terms)
    # one chained pair: its term on the other dependent drops out
    a, b = inds[0], inds[1]
    eqs.append([(a, 1, 2.0), (b, 1, -1.0)])
    eqs.append([(b, 1, 3.0), (inds[2], 1, -1.0)])
    m.mpc = M.Mpc.from_equations(eqs)
    nodal = np.zeros((m.nn, 6))
    nodal[free[-1], 2] = -100.0
    nodal[free[-2], 0] = 40.0
    m.loads.append(M.LoadCase(nodal=nodal))
    return m


def ring_full():
    from vibestran_amd import deck
    import os
    from e2e_csv import GOLD
    return deck.read_deck(os.path.join(GOLD, "cquad4_ring_6fold_full.bdf"))


def rbe_plate():
    from vibestran_amd import deck
    from tests.test_oracle_mpc import plate_deck
    cards = lambda nid, nx, ny: [
        "GRID,9999,,4.3,0.5,0.4", "GRID,9998,,4.2,0.45,0.3",
        "RBE2,101,9999,123456," + ",".join(str(nid(nx, j)) for j in (0, 1)),
        f"RBE3,201,,9998,123456,2.0,123,{nid(nx, 2)},{nid(nx - 1, 2)},0.5,123456,{nid(nx, 3)},"
        f"{nid(nx - 1, 3)}"]
    return deck.read_deck(plate_deck(cards).replace("ENDDATA", "FORCE,1,9998,0,20.,0.,1.,1.\nENDDATA"),
                          is_text=True)


MODELS = {
    "random_block": random_mpc_block,
    "tie_slide_block": lambda: decks.hex_block_mpc(5, 3, 4),
    "ring_full": ring_full,
    "rbe_plate": rbe_plate,
    # the tied end face makes the centre grid's block row ~240 blocks long:
    # past the sliced-ELL cap, so the overflow pass carries the rest
    "long_row_block": lambda: decks.hex_block_mpc(8, 10, 10),
}


def dense_lower_sym(rp, ci, v, n):
    L_ = sp.csr_matrix((v, ci, rp), shape=(n, n)).toarray()
    return L_ + np.tril(L_, -1).T


@pytest.mark.parametrize("name", sorted(MODELS))
def test_mpc_matches_oracle(gpu_backend_factory, name):
    m = MODELS[name]()
    be = gpu_backend_factory(tolerance=1e-13, max_iters=100000)
    eq, n, n_red = analysis_numbering(be, m)
    eq0, n0 = M.analysis_dof_map(m)
    _, _, mp = M.analysis_mpc_system(m, eq0, n0)
    omp = MpcOracle(eq, n, mp.equations())
    assert n_red == omp.n_red < n
    st = mpc_stats(be)
    assert (st.n_full, st.n_reduced, st.n_eliminations) == (n, n_red, len(omp.eliminations))
    # T, bit for bit
    tp, tc, tv = mpc_transform(be)
    for i in range(n):
        got = list(zip(tc[tp[i]:tp[i + 1]].tolist(), tv[tp[i]:tp[i + 1]].tolist()))
        assert got == [(int(r), float(c)) for r, c in omp.t_column(i)], i
    # K_red
    K = assemble_model(be, m, eq, n)
    assert K.n == n_red
    K.fetch()
    Kg = dense_lower_sym(K.row_ptr, K.col_ind, K.values, n_red)
    ok = assemble_model_mpc(m, eq, omp)
    Ko = ok.to_scipy().toarray()
    assert np.abs(Kg - Ko).max() <= 1e-12 * np.abs(Ko).max()
    assert assembly_stats(be).mpc_ms > 0.0
    # F_red
    F = assemble_rhs_device(be, m, eq, n_red, m.loads[0])
    Fo = nodal_rhs_mpc(m, eq, m.loads[0].nodal, omp)
    assert np.abs(F - Fo).max() <= 1e-12 * np.abs(Fo).max()
    # the operator in the reduced numbering: SpMV and block solve
    x = np.random.default_rng(0).standard_normal(n_red)
    assert np.abs(be.spmv(x) - Ko @ x).max() <= 1e-12 * np.abs(Ko @ x).max()
    U = be.solve_block(K, np.stack([F, -2.0 * F], axis=1))
    # solve, expand, recover
    u = be.solve(K, F)
    assert np.abs(U[:, 0] - u).max() <= 1e-9 * np.abs(u).max()
    assert np.abs(U[:, 1] + 2.0 * u).max() <= 2e-9 * np.abs(u).max()
    uo = omp.recover(np.linalg.solve(Ko, Fo))
    uf = mpc_expand(be, u, n)
    assert np.abs(uf - uo).max() <= 1e-9 * np.abs(uo).max()
    assert np.array_equal(uf, omp.recover(u))          # same term order, same arithmetic
    disp = R.recover_displacements(be, m, u)
    assert np.array_equal(disp, np.where(eq >= 0, uf[np.maximum(eq, 0)], 0.0))
    cen, _ = R.recover_stress(be, m, u, corners=False)
    ocen, _ = O.recover_stress(m, eq, uf, corners=False)
    assert np.abs(cen - ocen).max() <= 1e-11 * np.abs(ocen).max()


def test_clearing_restores_the_plain_operator(gpu_backend_factory):
    m = decks.hex_block_mpc(3, 2, 2)
    be = gpu_backend_factory()
    eq, n, n_red = analysis_numbering(be, m)
    assert assemble_model(be, m, eq, n).n == n_red < n
    assert set_mpc(be, m, eq, n, None) == n
    assert assemble_model(be, m, eq, n).n == n


def test_set_mpc_errors(gpu_backend_factory):
    m = decks.hex_block(2, 1, 1)
    be = gpu_backend_factory()
    eq, n = M.analysis_dof_map(m)
    free = [i for i in range(m.nn) if eq[i, 0] >= 0]
    a, b = free[:2]
    cyc = M.Mpc.from_equations([[(a, 1, 2.0), (b, 1, -1.0)], [(b, 1, 2.0), (a, 1, -1.0)]])
    with pytest.raises(SolverError, match="MPC dependency graph contains a cycle"):
        set_mpc(be, m, eq, n, cyc)
    bad = M.Mpc.from_equations([[(a, 7, 1.0), (b, 1, -1.0)]])
    with pytest.raises(ValueError, match="component must be 1..6"):
        set_mpc(be, m, eq, n, bad)
    ok = M.Mpc.from_equations([[(a, 1, 2.0), (b, 1, -1.0)]])
    assert set_mpc(be, m, eq, n, ok) == n - 1
    lib = L.load()
    a32 = [np.ascontiguousarray(x, dtype=np.int32) for x in (eq, m.elem_type, m.elem_ptr, m.conn)]
    f64 = [np.ascontiguousarray(x, dtype=np.float64) for x in (m.xyz, m.E, m.nu)]
    rc = lib.vbh_assemble(be._h, n + 1, m.nn, L.dp(f64[0]), L.ip(a32[0]), m.ne, L.ip(a32[1]),
                          L.ip(a32[2]), L.ip(a32[3]), L.dp(f64[1]), L.dp(f64[2]), None, 0, 0)
    assert rc == L.VBH_ERR_ARG and "n_full" in L.last_error()


def test_full_size_tie_and_slide(gpu_backend_factory):
    m = decks.hex_block_mpc(120, 40, 40)
    be = gpu_backend_factory()
    eq, n, n_red = analysis_numbering(be, m)
    print(f"hex 120x40x40 with MPCs: n_full {n}, n_reduced {n_red}, {m.mpc.n} equations", flush=True)
    K = assemble_model(be, m, eq, n)
    print(f"  K_red = T^T K T: {assembly_stats(be).mpc_ms:.1f} ms", flush=True)
    F = assemble_rhs_device(be, m, eq, n_red, m.loads[0])
    u = be.solve(K, F)
    assert be.stats().format == 3     # node blocks kept: dependents are identity rows
    uf = mpc_expand(be, u, n)
    umax = np.abs(uf).max()
    c = m.meta["centre"]
    for terms in m.mpc.equations():
        if len(terms) == 2 and terms[1][0] == c:       # tie: exactly the master's value
            (g, d, _), _ = terms
            assert uf[eq[g, d - 1]] == uf[eq[c, d - 1]]
    slides = [t for t in m.mpc.equations() if t[1][0] != c]
    res = [abs(sum(k * uf[eq[g, d - 1]] for g, d, k in t)) for t in slides]
    assert max(res) <= 1e-14 * umax * 4
    # reduced residual T^T (K u - F) from the element forces on the device
    f = R.element_forces(be, m, u)
    Ffull = np.zeros(n)
    mask = eq >= 0
    np.add.at(Ffull, eq[mask], m.loads[0].nodal[mask])
    Ku = np.zeros(n)
    np.add.at(Ku, eq[mask], f[mask])
    tp, tc, tv = mpc_transform(be)
    T = sp.csr_matrix((tv, tc, tp), shape=(n, n_red))
    r = T.T @ (Ku - Ffull)
    assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(F), np.linalg.norm(r) / np.linalg.norm(F)


def test_long_rows_mixed_precision(gpu_backend_factory):
    """A block row longer than the sliced-ELL cap through the mixed solver
    (FP32 SpMV overflow pass and FP64 refresh overflow pass): the true FP64
    residual of K_red meets the tolerance and the solution matches FP64's."""
    m = decks.hex_block_mpc(8, 10, 10)
    sols = {}
    for mixed in (False, True):
        be = gpu_backend_factory(mixed=mixed, tolerance=1e-10, max_iters=100000)
        eq, n, n_red = analysis_numbering(be, m)
        K = assemble_model(be, m, eq, n)
        F = assemble_rhs_device(be, m, eq, n_red, m.loads[0])
        u = be.solve(K, F)
        K.fetch()
        Kd = dense_lower_sym(K.row_ptr, K.col_ind, K.values, n_red)
        assert np.linalg.norm(F - Kd @ u) <= 1e-9 * np.linalg.norm(F)
        assert be.stats().format == 3
        sols[mixed] = u
    assert np.abs(sols[True] - sols[False]).max() <= 1e-7 * np.abs(sols[False]).max()
