"""Oracle Jacobi-PCG pinned by the reference's PCG test closed forms
(tests/unit/test_pcg_solver.cpp:80-222) and scipy direct solves (the role of
EigenSolverBackend there), plus the error paths of cuda_pcg_backend_impl.hpp."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from vibestran_amd import decks
from vibestran_amd import model as M

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def tridiag(n, d, off, lower_only=False):
    b = O.SparseMatrixBuilder(n)
    for i in range(n):
        b.add(i, i, d)
        if i > 0:
            b.add(i, i - 1, off)
            if not lower_only:
                b.add(i - 1, i, off)
    return b.build_csr()


def diag(n, v):
    b = O.SparseMatrixBuilder(n)
    for i in range(n):
        b.add(i, i, v)
    return b.build_csr()


def direct(K, F):
    return spla.spsolve(K.to_scipy().tocsc(), F)


def test_diagonal_exact():  # test_pcg_solver.cpp:80-95
    b = O.SparseMatrixBuilder(2)
    b.add(0, 0, 4.0)
    b.add(1, 1, 9.0)
    u, it, rel = O.pcg_jacobi(b.build_csr(), [8.0, 27.0])
    assert np.allclose(u, [2.0, 3.0], atol=1e-9) and it <= 1


@pytest.mark.parametrize("lower", [False, True])
def test_tridiagonal_vs_direct(lower):  # 100-118
    K = tridiag(5, 3.0, -1.0, lower)
    F = np.ones(5)
    u, it, rel = O.pcg_jacobi(K, F)
    assert np.allclose(u, direct(K, F), atol=1e-7)


def test_banded_500_vs_direct():  # 123-149
    n = 500
    b = O.SparseMatrixBuilder(n)
    for i in range(n):
        b.add(i, i, float(n + 1))
        if i > 0:
            b.add(i, i - 1, -1.0)
            b.add(i - 1, i, -1.0)
        if i > 1:
            b.add(i, i - 2, -0.5)
            b.add(i - 2, i, -0.5)
    K = b.build_csr()
    F = np.array([(i % 7) + 1 for i in range(n)], float)
    u, it, rel = O.pcg_jacobi(K, F)
    assert np.allclose(u, direct(K, F), atol=1e-6) and rel < 1e-7


def test_stiff_diagonal():  # 151-165
    u, it, rel = O.pcg_jacobi(diag(100, 1.3e9), np.full(100, 333.0))
    assert np.allclose(u, 333.0 / 1.3e9, rtol=1e-6)


def test_zero_rhs():  # 169-178 (CUDA variant 361-371: iters == 0)
    u, it, rel = O.pcg_jacobi(tridiag(10, 3.0, -1.0), np.zeros(10))
    assert np.all(u == 0.0) and it == 0


def test_large_tridiagonal_converges():  # 199-222
    n = 1000
    K = tridiag(n, 2.0, -1.0)
    F = np.ones(n)
    u, it, rel = O.pcg_jacobi(K, F)
    ud = direct(K, F)
    assert np.allclose(u, ud, atol=1e-6, rtol=1e-10) and it <= n


def test_errors():
    with pytest.raises(O.OracleError, match="zero diagonal"):
        b = O.SparseMatrixBuilder(2)
        b.add(0, 0, 1.0)
        b.add(1, 1, 0.0)
        O.pcg_jacobi(b.build_csr(), [1.0, 1.0])
    with pytest.raises(O.OracleError, match="missing diagonal"):
        b = O.SparseMatrixBuilder(2)
        b.add(0, 0, 1.0)
        b.add(1, 0, 1.0)
        O.pcg_jacobi(b.build_csr(), [1.0, 1.0])
    with pytest.raises(O.OracleError, match="did not converge"):
        O.pcg_jacobi(tridiag(50, 2.0, -1.0), np.ones(50), max_iters=3)
    with pytest.raises(O.OracleError):  # indefinite: p.Ap <= 0 breakdown
        b = O.SparseMatrixBuilder(2)
        b.add(0, 0, 1.0)
        b.add(1, 1, -1.0)
        O.pcg_jacobi(b.build_csr(), [0.0, 1.0])


def test_config1_cantilever_vs_direct_fixture():
    m = decks.config1_hex_cantilever()
    eq, n = M.analysis_dof_map(m)
    assert n == 7260
    K = O.assemble_solids(n, m.elem_type, m.elem_ptr, m.conn, m.xyz, eq, m.E, m.nu)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    u, it, rel = O.pcg_jacobi(K, F, tol=1e-10)
    g = np.load(os.path.join(GOLD, "c1_direct.npz"))
    assert g["n"] == n
    ud = g["u"]
    assert np.abs(u - ud).max() / np.abs(ud).max() < 1e-8
    assert O.true_relative_residual(K, u, F) < 1e-9
