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


def _reversed_rows(rp, ci, v):
    """The same lower CSR with each row's entries in reverse order: rows that
    are not increasing go through vbh_set_csr's host conversion, sorted rows
    through its device conversion."""
    ci2, v2 = ci.copy(), v.copy()
    for r in range(len(rp) - 1):
        a, b = rp[r], rp[r + 1]
        ci2[a:b] = ci[a:b][::-1]
        v2[a:b] = v[a:b][::-1]
    return ci2, v2


def _small_tube():
    m = decks.cquad4_tube(n_circ=24, n_axial=30, n_cases=1)
    eq, n = M.analysis_dof_map(m)
    oc = O.assemble_model(n, m, eq)
    return oc, M.assemble_rhs(m, eq, n, m.loads[0])


@pytest.mark.parametrize("case,fmt", [("c1", -1), ("c1", 3), ("c1", 6), ("tube", -1),
                                      ("tube", 3)])
def test_set_csr_device_conversion_matches_host(gpu_backend_factory, c1, case, fmt):
    """vbh_set_csr converts sorted lower storage (the reference's CsrData) on
    the device; its operator must equal the host conversion's bit for bit:
    same block format, same SpMV, same PCG iterates."""
    if case == "c1":
        oc, F = c1[3], c1[4]
    else:
        oc, F = _small_tube()
    rp, ci, v = oc.arrays()
    assert oc.stores_lower_triangle_only()
    ci_r, v_r = _reversed_rows(rp, ci, v)
    x = np.random.default_rng(3).standard_normal(len(rp) - 1)
    res = []
    for cols, vals in ((ci, v), (ci_r, v_r)):
        for mixed in (False, True):
            be = gpu_backend_factory(mixed=mixed, tolerance=1e-9, max_iters=20000)
            be.set_format(fmt)
            K = CsrData.from_arrays(rp, cols, vals, lower_only=True)
            be.set_matrix(K)
            y = be.spmv(x)
            u = be.solve(K, F)
            res.append((be.stats().format, y, u, be.last_iteration_count()))
    dev, host = res[:2], res[2:]
    for a, b in zip(dev, host):
        assert a[0] == b[0] and a[0] in (3, 6)
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[2], b[2]) and a[3] == b[3]
    yo = oc.multiply(x)
    assert np.abs(dev[0][1] - yo).max() <= 1e-13 * np.abs(yo).max()


def test_set_csr_device_conversion_keeps_lower_csr(gpu_backend_factory, c1):
    """After a device conversion the context hands back the caller's CSR."""
    from vibestran_amd.assembly import DeviceCsr
    oc = c1[3]
    rp, ci, v = oc.arrays()
    be = gpu_backend_factory()
    be.set_matrix(CsrData.from_arrays(rp, ci, v, lower_only=True))
    K = DeviceCsr(be, len(rp) - 1, len(ci)).fetch()
    assert np.array_equal(K.row_ptr, rp) and np.array_equal(K.col_ind, ci)
    assert np.array_equal(K.values, v)


def test_set_csr_error_order(gpu_backend_factory):
    """The first failing row decides the error, as in the reference's row loop
    (cuda_pcg_backend_impl.hpp:2076-2101); the context keeps its operator."""
    oc = tridiag(40, 4.0, -1.0, lower_only=True)
    rp, ci, v = (a.copy() for a in oc.arrays())
    be = gpu_backend_factory()
    good = CsrData.from_arrays(rp, ci, v, lower_only=True)
    u0 = be.solve(good, np.ones(40))

    def bad(edit):
        c2, v2 = ci.copy(), v.copy()
        edit(c2, v2)
        return CsrData.from_arrays(rp, c2, v2, lower_only=True)

    def diag_of(r):
        return rp[r + 1] - 1  # the diagonal ends each lower row

    def zero_then_range(c2, v2):  # zero diagonal at row 7 before a bad column at row 9
        v2[diag_of(7)] = 0.0
        c2[rp[9]] = 99

    def range_then_zero(c2, v2):  # bad column at row 5 before a zero diagonal at row 8
        c2[rp[5]] = -1
        v2[diag_of(8)] = 0.0

    def missing_then_upper(c2, v2):  # row 6 loses its diagonal, row 12 gets an upper entry
        c2[diag_of(6)] = 4
        c2[rp[12]] = 13

    with pytest.raises(SolverError, match="zero diagonal at row 7 "):
        be.solve(bad(zero_then_range), np.ones(40))
    with pytest.raises(ValueError, match="column index out of range"):
        be.solve(bad(range_then_zero), np.ones(40))
    with pytest.raises(SolverError, match="missing diagonal at row 6 "):
        be.solve(bad(missing_then_upper), np.ones(40))
    with pytest.raises(ValueError, match="row_ptr decreases"):
        rp2 = rp.copy()
        rp2[3], rp2[4] = rp2[4], rp2[3]
        be.solve(CsrData.from_arrays(rp2, ci, v, lower_only=True), np.ones(40))
    u1 = be.solve(good, np.ones(40))  # the operator installed before the failures
    assert np.array_equal(u0, u1)


def _expected_stream_bytes(rp, ci, n, B):
    """Operator stream bytes of the sliced-ELL layouts, computed from the lower
    CSR pattern alone: the full symmetric block pattern, each 64-row slice's
    widest block row (absolute layout) and its distinct block-column offsets
    (relative layout, chosen when within 10 % of the absolute slot count).
    Returns (fp64, mixed) for the relative layout, or None if absolute."""
    rows = np.repeat(np.arange(n), np.diff(rp))
    nbr = -(-n // B)
    bi = np.concatenate([rows // B, ci // B, np.arange(nbr)])
    bj = np.concatenate([ci // B, rows // B, np.arange(nbr)])
    pair = np.unique(bi.astype(np.int64) * nbr + bj)
    I, J = pair // nbr, pair % nbr
    rowlen = np.bincount(I, minlength=nbr)
    ns = -(-nbr * B // 64)
    if rowlen.max() > 128:
        return None
    s0, s1 = (B * I) // 64, (B * I + B - 1) // 64
    sl = np.concatenate([s0, s1[s1 != s0]])
    off = np.concatenate([J - I, (J - I)[s1 != s0]])
    key = np.unique(sl * (2 * nbr) + (off + nbr))
    ksl, koff = key // (2 * nbr), key % (2 * nbr) - nbr
    rel64 = len(key)
    rel32 = rel64 - int(np.count_nonzero(koff == 0))
    rs0, rs1 = (B * np.arange(nbr)) // 64, (B * np.arange(nbr) + B - 1) // 64
    kmax = np.zeros(ns, np.int64)
    np.maximum.at(kmax, rs0, rowlen)
    np.maximum.at(kmax, rs1, rowlen)
    if np.bincount(ksl, minlength=ns).max() > 128 or rel64 > 1.1 * kmax.sum():
        return None
    fp64 = rel64 * 64 * B * 8 + rel64 * 4 + 4 * (ns + 1)
    mixed = rel32 * 64 * B * 4 + rel32 * 4 + 4 * (ns + 1) + nbr * B * B * 8
    return fp64, mixed


def test_relative_sell_slots_match_pattern(gpu_backend_factory):
    """The relative sliced ELL's slot counts, built from the device BSR, equal
    those computed from the CSR pattern on the host (C1 BSR3, a shell tube
    BSR6, a small tube through forced 3x3 blocks)."""
    import ctypes as C
    from vibestran_amd import _lib as L
    m1 = decks.config1_hex_cantilever()
    eq, n = M.analysis_dof_map(m1)
    c1 = O.assemble_solids(n, m1.elem_type, m1.elem_ptr, m1.conn, m1.xyz, eq, m1.E, m1.nu)
    tube, _ = _small_tube()
    seen_rel = 0
    for oc, fmt, B in ((c1, -1, 3), (tube, -1, 6), (tube, 3, 3)):
        rp, ci, v = oc.arrays()
        n = len(rp) - 1
        be = gpu_backend_factory(max_iters=100000)
        be.set_format(fmt)
        be.solve(CsrData.from_arrays(rp, ci, v, lower_only=True), np.ones(n))
        assert be.stats().format == B
        exp = _expected_stream_bytes(rp, ci, n, B)
        got = []
        for prec in (L.VBH_PREC_FP64, L.VBH_PREC_MIXED):
            b = C.c_longlong(0)
            assert L.load().vbh_operator_stream_bytes(be._h, prec, C.byref(b)) == L.VBH_OK
            got.append(b.value)
        if exp is not None:
            seen_rel += 1
            assert tuple(got) == exp, (B, got, exp)
    assert seen_rel >= 2


def test_set_csr_rejects_operators_beyond_32_bit_indices(gpu_backend_factory):
    """A lower CSR whose full symmetric operator would hold 2^31 or more
    entries is refused before any entry is read (only row_ptr's two ends are
    looked at here), and the context stays usable."""
    import ctypes as C
    from vibestran_amd import _lib as L
    be = gpu_backend_factory()
    oc = tridiag(20, 4.0, -1.0, lower_only=True)
    u0 = be.solve(csr_of(oc), np.ones(20))
    rp = np.array([0, 2 ** 30], dtype=np.int32)
    ci = np.zeros(1, dtype=np.int32)
    v = np.ones(1)
    rc = L.load().vbh_set_csr(be._h, 1, 2 ** 30, L.ip(rp), L.ip(ci), L.dp(v), 1)
    assert rc == L.VBH_ERR_ARG and "2^31" in L.last_error()
    assert np.array_equal(be.solve(csr_of(oc), np.ones(20)), u0)


def test_set_csr_full_and_duplicate_storage(gpu_backend_factory, c1):
    """The same operator handed over as the oracle's lower CSR (device
    conversion), as its expanded full symmetric CSR, and as a lower CSR with
    every off-diagonal entry split into two halves (both host conversions,
    duplicates summed, exactly): same block format and bit-identical SpMV and
    PCG iterates."""
    oc, F = c1[3], c1[4]
    rp, ci, v = oc.arrays()
    full = oc.expanded_symmetric()
    frp, fci, fv = full.arrays()
    assert not full.stores_lower_triangle_only()
    # duplicates: each off-diagonal lower entry v -> v/2 twice (exact halves)
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = ci != rows
    reps = np.where(off, 2, 1)
    drp = np.concatenate([[0], np.cumsum(np.add.reduceat(reps, rp[:-1]))]).astype(np.int32)
    dci = np.repeat(ci, reps).astype(np.int32)
    dv = np.repeat(np.where(off, v / 2.0, v), reps)
    x = np.random.default_rng(5).standard_normal(n)
    yo = oc.multiply(x)
    res = []
    for K in (CsrData.from_arrays(rp, ci, v, lower_only=True),
              CsrData.from_arrays(frp, fci, fv, lower_only=False),
              CsrData.from_arrays(drp, dci, dv, lower_only=True)):
        be = gpu_backend_factory(tolerance=1e-10, max_iters=20000)
        be.set_matrix(K)
        y = be.spmv(x)
        assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
        u = be.solve(K, F)
        res.append((be.stats().format, be.last_iteration_count(), u, y))
    assert np.array_equal(res[1][3], res[0][3]) and np.array_equal(res[2][3], res[0][3])
    assert res[0][0] == res[1][0] == res[2][0] == 3
    assert res[0][1] == res[1][1] == res[2][1]
    for r in res[1:]:
        assert np.array_equal(r[2], res[0][2])
