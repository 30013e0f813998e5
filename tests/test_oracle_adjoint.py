"""The FP64 CPU adjoint (oracle) pinned by central finite differences of the
full assemble -> direct-solve chain, by the analytic dK/dt against finite
differences of the element matrix, and by the symmetry of the contraction."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from vibestran_amd import decks
from vibestran_amd import model as M
from tests.test_gpu_shell import warped_mesh
from tests.test_oracle_shell import PLATE, rot


def solve(m, eq, n, F):
    K = O.assemble_model(n, m, eq)
    return spla.spsolve(K.to_scipy().tocsc(), F)


@pytest.mark.parametrize("form", ["MITC4", "MINDLIN"])
def test_dke_dt_matches_fd(form):
    X = PLATE @ (rot("x", 0.4) @ rot("z", 0.2)).T
    t = 0.4
    h = 1e-6 * t
    d = O.quad4_dke_dt(X, 7e4, 0.33, t, form)
    fd = (O.quad4_ke(X, 7e4, 0.33, t + h, form) - O.quad4_ke(X, 7e4, 0.33, t - h, form)) / (2 * h)
    assert np.abs(d - fd).max() <= 1e-7 * np.abs(d).max()


def test_simp_compliance_gradient_fd():
    m = decks.config4_simp(nx=6, ny=3, nz=2, seed=5)
    eq, n = M.analysis_dof_map(m)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u = solve(m, eq, n, F)
    g = O.element_sensitivity(n, m, eq, 0, u, u)  # dC/dE_e
    rng = np.random.default_rng(0)
    for e in rng.choice(m.ne, 6, replace=False):
        h = 1e-4 * m.E[e]  # truncation O(h^2) ~ 1e-8, solve roundoff / h ~ 1e-7
        Ep, Em = m.E.copy(), m.E.copy()
        Ep[e] += h
        Em[e] -= h
        mp = M.Model(**{**m.__dict__, "E": Ep})
        mm = M.Model(**{**m.__dict__, "E": Em})
        fd = (F @ solve(mp, eq, n, F) - F @ solve(mm, eq, n, F)) / (2 * h)
        assert abs(g[e] - fd) <= 2e-6 * abs(fd) + 1e-12 * abs(g).max(), (e, g[e], fd)


@pytest.mark.parametrize("form", [0, 1])
def test_shell_tip_thickness_gradient_fd(form):
    m = warped_mesh(nx=4, ny=3, form=form)
    eq, n = M.analysis_dof_map(m)
    F = np.random.default_rng(4).standard_normal(n)
    tip = n - 3  # one translational DOF on the far edge
    L = np.zeros(n)
    L[tip] = 1.0
    u = solve(m, eq, n, F)
    lam = solve(m, eq, n, L)
    g = O.element_sensitivity(n, m, eq, 1, u, lam)  # d(u_tip)/dt_e
    for e in range(m.ne):
        h = 1e-3 * m.thickness[e]  # thin shell: |u_tip| ~ 1e2, solve noise / h dominates below
        tp, tm = m.thickness.copy(), m.thickness.copy()
        tp[e] += h
        tm[e] -= h
        fp = solve(M.Model(**{**m.__dict__, "thickness": tp}), eq, n, F)[tip]
        fm = solve(M.Model(**{**m.__dict__, "thickness": tm}), eq, n, F)[tip]
        fd = (fp - fm) / (2 * h)
        # absolute part: FD noise of the direct solve (|u_tip| eps / h) and O(h^2)
        assert abs(g[e] - fd) <= 1e-5 * abs(fd) + 2e-7 * abs(g).max(), (e, g[e], fd)


def test_contraction_is_symmetric():
    m = warped_mesh(nx=3, ny=2)
    eq, n = M.analysis_dof_map(m)
    rng = np.random.default_rng(1)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    for p in (0, 1):
        a = O.element_sensitivity(n, m, eq, p, u, v)
        b = O.element_sensitivity(n, m, eq, p, v, u)
        assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max()
