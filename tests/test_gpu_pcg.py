"""GPU Jacobi-PCG through the C ABI, mirroring the reference's CUDA PCG tests
(tests/unit/test_pcg_solver.cpp:237-427) with scipy direct solves in the role
of EigenSolverBackend, and parity against the CPU oracle."""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from vibestran_amd import CsrData, SolverError, decks
from vibestran_amd import model as M

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def csr_of(oc):
    rp, ci, v = oc.arrays()
    return CsrData.from_arrays(rp, ci, v, lower_only=oc.stores_lower_triangle_only())


def tridiag(n, d, off, lower_only=False):
    b = O.SparseMatrixBuilder(n)
    for i in range(n):
        b.add(i, i, d)
        if i > 0:
            b.add(i, i - 1, off)
            if not lower_only:
                b.add(i - 1, i, off)
    return b.build_csr()


def direct(oc, F):
    return spla.spsolve(oc.to_scipy().tocsc(), F)


@pytest.fixture(scope="module")
def be(gpu_backend_factory):
    return gpu_backend_factory()


def test_diagonal_exact(be):  # test_pcg_solver.cpp:268-282
    b = O.SparseMatrixBuilder(2)
    b.add(0, 0, 4.0)
    b.add(1, 1, 9.0)
    u = be.solve(csr_of(b.build_csr()), np.array([8.0, 27.0]))
    assert np.allclose(u, [2.0, 3.0], atol=1e-7) and be.last_iteration_count() <= 1


@pytest.mark.parametrize("lower", [False, True])
def test_tridiagonal_agrees_with_direct(be, lower):  # 286-318
    oc = tridiag(8, 3.0, -1.0, lower)
    F = np.ones(8)
    u = be.solve(csr_of(oc), F)
    assert np.allclose(u, direct(oc, F), atol=1e-6)


def test_larger_tridiagonal(be):  # 322-338
    oc = tridiag(200, 2.0, -1.0)
    F = np.ones(200)
    u = be.solve(csr_of(oc), F)
    assert np.allclose(u, direct(oc, F), atol=1e-5) and be.last_relative_residual() < 1e-7


def test_stiff_diagonal(be):  # 342-357
    b = O.SparseMatrixBuilder(100)
    for i in range(100):
        b.add(i, i, 1.3e9)
    u = be.solve(csr_of(b.build_csr()), np.full(100, 333.0))
    assert np.allclose(u, 333.0 / 1.3e9, rtol=1e-5)


def test_zero_rhs(be):  # 361-371
    u = be.solve(csr_of(tridiag(10, 3.0, -1.0)), np.zeros(10))
    assert np.all(u == 0.0) and be.last_iteration_count() == 0


def test_banded_500(be):  # 375-398
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
    oc = b.build_csr()
    F = np.array([(i % 7) + 1 for i in range(n)], float)
    assert np.allclose(be.solve(csr_of(oc), F), direct(oc, F), atol=1e-5)


def test_names(be):  # 402-408
    assert "PCG" in be.name() and be.device_name()


def test_error_paths(gpu_backend_factory):
    be = gpu_backend_factory(max_iters=3)
    with pytest.raises(SolverError, match="zero diagonal"):
        be.solve(CsrData.from_arrays([0, 1, 2], [0, 1], [1.0, 0.0]), np.ones(2))
    with pytest.raises(SolverError, match="missing diagonal"):
        be.solve(CsrData.from_arrays([0, 1, 2], [0, 0], [1.0, 1.0]), np.ones(2))
    with pytest.raises(SolverError, match="empty"):
        be.solve(CsrData.from_arrays([0], [], []), np.ones(0))
    with pytest.raises(SolverError, match="size"):
        be.solve(csr_of(tridiag(4, 2.0, -1.0)), np.ones(3))
    with pytest.raises(SolverError, match="did not converge after 3"):
        be.solve(csr_of(tridiag(50, 2.0, -1.0)), np.ones(50))
    be2 = gpu_backend_factory()
    with pytest.raises(SolverError, match="did not converge"):  # p.Ap <= 0 breakdown
        # indefinite [[1,2],[2,1]]: r.z = 2 > 0 but p.Ap = -2
        be2.solve(CsrData.from_arrays([0, 1, 3], [0, 0, 1], [1.0, 2.0, 1.0], lower_only=True),
                  np.array([1.0, -1.0]))
    with pytest.raises(SolverError, match="non-positive initial r"):  # impl:1859-1862
        be2.solve(CsrData.from_arrays([0, 1, 2], [0, 1], [1.0, -1.0], lower_only=True),
                  np.array([0.0, 1.0]))


def test_mixed_error_paths(gpu_backend_factory):
    """The reference's outcomes in mixed precision, where pauses for an FP64
    refresh run inside the iteration kernels: the iteration limit, a zero
    load, p.Ap <= 0 and a non-positive initial r.z."""
    be = gpu_backend_factory(mixed=True, max_iters=3)
    with pytest.raises(SolverError, match="did not converge after 3"):
        be.solve(csr_of(tridiag(50, 2.0, -1.0)), np.ones(50))
    be2 = gpu_backend_factory(mixed=True)
    u = be2.solve(csr_of(tridiag(20, 2.0, -1.0)), np.zeros(20))
    assert not u.any() and be2.last_iteration_count() == 0
    with pytest.raises(SolverError, match="did not converge"):  # p.Ap <= 0 breakdown
        be2.solve(CsrData.from_arrays([0, 1, 3], [0, 0, 1], [1.0, 2.0, 1.0], lower_only=True),
                  np.array([1.0, -1.0]))
    with pytest.raises(SolverError, match="non-positive initial r"):
        be2.solve(CsrData.from_arrays([0, 1, 2], [0, 1], [1.0, -1.0], lower_only=True),
                  np.array([0.0, 1.0]))
    # the context stays usable after each failure
    oc = tridiag(200, 2.0, -1.0)
    u = be2.solve(csr_of(oc), np.ones(200))
    assert np.allclose(u, direct(oc, np.ones(200)), atol=5e-3)


@pytest.fixture(scope="module")
def c1():
    m = decks.config1_hex_cantilever()
    eq, n = M.analysis_dof_map(m)
    oc = O.assemble_solids(n, m.elem_type, m.elem_ptr, m.conn, m.xyz, eq, m.E, m.nu)
    F = M.assemble_rhs(m, eq, n, m.loads[0])
    return m, eq, n, oc, F


@pytest.mark.parametrize("fmt", [-1, 0, 3])
def test_config1_fp64_parity_with_oracle(gpu_backend_factory, c1, fmt):
    m, eq, n, oc, F = c1
    be = gpu_backend_factory(tolerance=1e-10, max_iters=20000)
    be.set_format(fmt)
    u = be.solve(csr_of(oc), F)
    uo, it_o, rel_o = O.pcg_jacobi(oc, F, tol=1e-10, max_iters=20000)
    assert abs(be.last_iteration_count() - it_o) <= 2
    assert np.abs(u - uo).max() / np.abs(uo).max() < 1e-9
    ud = np.load(os.path.join(GOLD, "c1_direct.npz"))["u"]
    assert np.abs(u - ud).max() / np.abs(ud).max() < 1e-8
    assert O.true_relative_residual(oc, u, F) < 1e-9
    assert be.stats().format == (3 if fmt == -1 else fmt)


def test_config1_spmv_parity(gpu_backend_factory, c1):
    m, eq, n, oc, F = c1
    for fmt in (0, 3):
        be = gpu_backend_factory()
        be.set_format(fmt)
        be.set_matrix(csr_of(oc))
        x = np.random.default_rng(1).standard_normal(n)
        y = be.spmv(x)
        yo = oc.multiply(x)
        assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()


def test_config1_mixed_reaches_1e8(gpu_backend_factory, c1):
    m, eq, n, oc, F = c1
    be = gpu_backend_factory(mixed=True, tolerance=1e-8, max_iters=20000)
    u = be.solve(csr_of(oc), F)
    rel = O.true_relative_residual(oc, u, F)
    assert rel < 1e-8 and be.last_relative_residual() < 1e-8
    ud = np.load(os.path.join(GOLD, "c1_direct.npz"))["u"]
    assert np.abs(u - ud).max() / np.abs(ud).max() < 1e-6
    assert be.stats().refinements >= 1


def test_mixed_tridiagonal(gpu_backend_factory):  # 412-425 (reference bound 5e-3)
    be = gpu_backend_factory(mixed=True)
    oc = tridiag(200, 2.0, -1.0)
    F = np.ones(200)
    u = be.solve(csr_of(oc), F)
    assert np.allclose(u, direct(oc, F), atol=5e-3) and be.last_estimated_error() < 1e-4


def _sell_cases():
    """C1 (BSR3), a shell tube (BSR6), and a jittered hex of 40,625 nodes
    whose node IDs are shuffled, so that block columns lie more than 32,767
    nodes apart and the sliced ELL keeps 32-bit block columns."""
    from vibestran_amd.assembly import assemble_model, dof_map_device
    hx = decks.config2_hex_jitter(nx=64, ny=24, nz=24, seed=3)
    hx.node_ids = np.random.default_rng(9).permutation(hx.node_ids).astype(np.int32)
    out = []
    for m in (decks.config1_hex_cantilever(), decks.cquad4_tube(n_circ=16, n_axial=24, n_cases=1), hx):
        from vibestran_amd import HipPcgSolverBackend
        be = HipPcgSolverBackend.try_create()
        eq, n = dof_map_device(be, m)
        K = assemble_model(be, m, eq, n).fetch()
        csr = CsrData.from_arrays(K.row_ptr.copy(), K.col_ind.copy(), K.values.copy(), lower_only=True)
        out.append((csr, M.assemble_rhs(m, eq, n, m.loads[0])))
        be.close()
    return out


@pytest.mark.parametrize("mixed", [False, True])
def test_sell_layouts_give_identical_solves(gpu_backend_factory, monkeypatch, mixed):
    """The three sliced-ELL layouts -- relative (one block-column offset per
    slot and slice; VBH_SELL_REL), absolute with 16-bit node0-relative block
    columns, absolute with 32-bit block columns (VBH_SELL_IDX16=0) -- sum every
    row's products in the same column order (the relative layout's extra zero
    blocks add exact zeros), so solves and iteration counts are bit-identical
    on the same operator. The shuffled-numbering hex keeps the absolute 32-bit
    layout (its slices see too many offsets for the relative one)."""
    for csr, F in _sell_cases():
        res = []
        for rel, idx16 in (("1", "1"), ("0", "1"), ("0", "0")):
            monkeypatch.setenv("VBH_SELL_REL", rel)
            monkeypatch.setenv("VBH_SELL_IDX16", idx16)
            be = gpu_backend_factory(mixed=mixed, tolerance=1e-8, max_iters=50000)
            u = be.solve(csr, F)
            res.append((u, be.last_iteration_count()))
        (u0, it0) = res[0]
        for u, it in res[1:]:
            assert it == it0 and np.array_equal(u, u0)
